"""Ensemble / serving layer on the GPU: footprint accounting, a member's
result independent of what runs beside it, HBM-budget admission and
arrival-time ordering."""
import numpy as np
import pytest

import atmospheric_flows_amd as af

pytestmark = pytest.mark.gpu

FR = 1.0 / np.sqrt(9.81)


@pytest.mark.parametrize("k,nx,nz,prec", [(2, 16, 8, 0), (3, 32, 16, 1)])
def test_footprint_matches_allocations(k, nx, nz, prec):
    h = af.witch_of_agnesi(nx)
    ctx = af.DeviceContext(k, nx, nz, h=h, Ma=1.0, Fr=FR, dt=4.0, precision=prec)
    assert ctx.device_bytes() == af.device_bytes(k, nx, nz, h=h, precision=prec)


def run_alone(m):
    ctx = af.DeviceContext(m.degree, m.nx, m.nz, h=af.witch_of_agnesi(m.nx, h0=m.h0), Ma=1.0, Fr=FR, **m.params)
    rho, u, p = ctx.init_hill(N_bv=m.N_bv, U=m.U, seed=m.seed)
    ctx.upload_state(rho, u, p)
    ctx.imex_step(m.nsteps)
    return ctx.download_state()


def test_member_result_does_not_depend_on_the_mix():
    x = af.Member(2, 32, 16, nsteps=3, h0=250.0, U=8.0, N_bv=0.011, seed=5, params=dict(dt=8.0, precision=1))
    others = [af.Member(3 if i % 2 else 2, 16 * (1 + i % 3), 8 * (1 + i % 3), nsteps=2 + i % 3, h0=100.0 + 50 * i,
                        seed=100 + i, params=dict(dt=8.0, precision=i % 2)) for i in range(6)]
    ens = af.Ensemble(slots=4, keep_states=True)
    ids = [ens.submit(m) for m in others[:3]] + [ens.submit(x)] + [ens.submit(m) for m in others[3:]]
    ens.run()
    for i in ids:
        r = ens.result(i)
        assert r["status"] == 0, r["error"]
        assert r["steps_done"] == ens.members[i].nsteps
    mixed = ens.state(ids[3])
    alone = run_alone(x)
    for a, b in zip(mixed, alone):
        assert np.array_equal(a, b)
    solo = af.Ensemble(slots=1, keep_states=True)
    solo.submit(x)
    solo.run()
    for a, b in zip(solo.state(0), alone):
        assert np.array_equal(a, b)


def test_admission_respects_the_hbm_budget_and_arrivals():
    big = af.Member(3, 64, 32, nsteps=2, seed=1, params=dict(dt=8.0))
    small = af.Member(2, 16, 8, nsteps=2, seed=2, params=dict(dt=8.0))
    nb = af.device_bytes(3, 64, 32, h=af.witch_of_agnesi(64))
    ns = af.device_bytes(2, 16, 8, h=af.witch_of_agnesi(16))
    ens = af.Ensemble(slots=4, hbm_budget_bytes=nb + ns + 1)
    bigs = [ens.submit(af.Member(**{**big.__dict__, "seed": 10 + i})) for i in range(3)]
    late = ens.submit(af.Member(**{**small.__dict__, "arrival_s": 0.4}))
    too_big = ens.submit(af.Member(3, 128, 64, nsteps=1, params=dict(dt=4.0)))
    ens.run()
    rb = [ens.result(i) for i in bigs]
    for r in rb:
        assert r["status"] == 0, r["error"]
    # at most one big member holds HBM at any time
    iv = sorted((r["admit_s"], r["finish_s"]) for r in rb)
    for (a0, a1), (b0, b1) in zip(iv, iv[1:]):
        assert b0 >= a1 - 1e-6
    rl = ens.result(late)
    assert rl["status"] == 0 and rl["admit_s"] >= 0.4 - 1e-6
    assert ens.result(too_big)["status"] == af._native.ATMO_OOM


def test_ensemble_mesh_set_alone_and_in_the_mix():
    """The ensemble workload's mesh set (64x32, 128x64, 256x128 at Q2 and Q3,
    twice each with other drawn parameters: 12 members, mixed precision): each
    member's state after its steps is bit for bit the same run in a mix of all
    twelve on 8 slots as run alone in its own context."""
    rng = np.random.default_rng(77)
    members = []
    for rep in range(2):
        for nx, nz in ((64, 32), (128, 64), (256, 128)):
            for k in (2, 3):
                U = float(rng.uniform(5, 15))
                dt = min(512.0 / nx, 0.35 * (40000.0 / nx) / (3.0 * U * k * np.sqrt(2.0)))
                members.append(af.Member(k, nx, nz, nsteps=2 + rep, h0=float(rng.uniform(100, 600)), U=U,
                                         N_bv=float(rng.uniform(0.008, 0.012)), seed=700 + len(members),
                                         params=dict(dt=dt, eps=1e-10, precision=1)))
    ens = af.Ensemble(slots=8, keep_states=True)
    ids = [ens.submit(m) for m in members]
    ens.run()
    for i, m in zip(ids, members):
        r = ens.result(i)
        assert r["status"] == 0, r["error"]
        alone = run_alone(m)
        for a, b in zip(ens.state(i), alone):
            assert np.array_equal(a, b), (m.nx, m.degree)


def test_autotune_picks_a_measured_slot_count():
    """The batch-shape autotuner runs the first members on each candidate
    slot count, keeps the fastest, and the ensemble then runs with it."""
    ens = af.Ensemble(slots=1, keep_states=True)
    ms = [af.Member(2 + i % 2, 16 * (1 + i % 2), 8 * (1 + i % 2), nsteps=3, seed=40 + i, params=dict(dt=8.0))
          for i in range(6)]
    for m in ms:
        ens.submit(m)
    best, rates = ens.autotune(candidates=(1, 2, 4), probe_steps=2, sample=6)
    assert best in (1, 2, 4) and all(r > 0 for r in rates.values())
    assert rates[best] == max(rates.values())
    ens.run()
    for i, m in enumerate(ms):
        r = ens.result(i)
        assert r["status"] == 0 and r["steps_done"] == 3
    # the probes leave the members' own results untouched: member 0 equals its run alone
    for a, b in zip(ens.state(0), run_alone(ms[0])):
        assert np.array_equal(a, b)


def test_autotune_rate_tolerance_picks_fewest_slots_near_the_best():
    """With a rate tolerance the autotuner keeps the fewest slots whose
    measured rate is within that fraction of the best rate; out-of-range
    tolerances are refused."""
    ens = af.Ensemble(slots=3)
    for i in range(6):
        ens.submit(af.Member(2, 16, 8, nsteps=2, seed=60 + i, params=dict(dt=8.0)))
    best, rates = ens.autotune(candidates=(4, 1, 2), probe_steps=2, sample=6, rate_tolerance=0.5)
    top = max(rates.values())
    assert top > 0
    assert best == min(c for c, r in rates.items() if r >= 0.5 * top)
    with pytest.raises(af.AtmoError):
        ens.autotune(candidates=(1, 2), probe_steps=1, sample=2, rate_tolerance=1.0)
