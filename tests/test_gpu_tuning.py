"""Kernel-shape autotuner (atmo_tune_kernels, atmo_ensemble_tune_kernels).

The tunable shapes (the Arnoldi kernel's waves per SIMD; the FP32 multigrid
sweep on the matrix cores with 16 cells per wave or on the vector ALUs) change
no arithmetic, so a tuned context or ensemble must give results bitwise equal
to the untuned one; the tuner must report a timing for every candidate and
choose the faster one.
"""
import numpy as np
import pytest

import atmospheric_flows_amd as af

pytestmark = pytest.mark.gpu

G = 9.81
FR = 1.0 / np.sqrt(G)


def run_steps(ctx, n):
    rho, u, p = ctx.init_hill(seed=7)
    ctx.upload_state(rho, u, p)
    ctx.imex_step(n)
    return ctx.download_state()


@pytest.mark.parametrize("k,nx,nz", [(2, 64, 32), (3, 32, 16)])
def test_context_tuning_choices_and_results(k, nx, nz):
    h = af.witch_of_agnesi(nx)
    common = dict(h=h, Ma=1.0, Fr=FR, dt=4.0, eps=1e-10, precision=1)
    ref = run_steps(af.DeviceContext(k, nx, nz, **common), 2)
    ctx = af.DeviceContext(k, nx, nz, **common)
    t = ctx.tune_kernels(nrep=10)
    print(t)
    assert all(x > 0 for x in t["arnoldi_us"])
    assert t["arnoldi_waves_per_simd"] == (3 if t["arnoldi_us"][0] < t["arnoldi_us"][1] else 4)
    assert t["sweep_us_vector_alus"] > 0 and t["sweep_us_matrix_cores"] > 0
    assert t["mg_matrix_cores"] == (1 if t["sweep_us_matrix_cores"] <= t["sweep_us_vector_alus"] else 0)
    got = run_steps(ctx, 2)
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    # every shape by hand gives the same results too
    for waves in (3, 4):
        for mx in (0, 1):
            c = af.DeviceContext(k, nx, nz, arnoldi_waves_per_simd=waves, mg_matrix_cores=mx, **common)
            for a, b in zip(run_steps(c, 1), run_steps(af.DeviceContext(k, nx, nz, **common), 1)):
                assert np.array_equal(a, b), (waves, mx)


def test_fp64_form_is_not_tuned():
    nx, nz = 32, 16
    ctx = af.DeviceContext(2, nx, nz, h=af.witch_of_agnesi(nx), Ma=1.0, Fr=FR, dt=4.0, precision=0)
    t = ctx.tune_kernels(nrep=5)
    assert t["arnoldi_waves_per_simd"] == 0 and t["mg_matrix_cores"] == -1


def test_ensemble_tuning_per_class():
    meshes = [(64, 32, 2), (32, 16, 3), (64, 32, 2), (32, 16, 3), (64, 32, 2)]

    def build(tune):
        ens = af.Ensemble(slots=3, keep_states=True)
        for i, (nx, nz, k) in enumerate(meshes):
            ens.submit(af.Member(degree=k, nx=nx, nz=nz, h0=300.0 + 50 * i, U=10.0, N_bv=0.01, seed=11 + i, nsteps=2,
                                 params=dict(dt=4.0, eps=1e-10, precision=1)))
        classes = ens.tune_kernels(nrep=5) if tune else None
        ens.run()
        return ens, classes

    tuned, classes = build(True)
    print(classes)
    assert len(classes) == 2
    for c in classes:
        assert c["arnoldi_waves_per_simd"] in (3, 4) and c["mg_matrix_cores"] in (0, 1)
    plain, _ = build(False)
    for i in range(len(meshes)):
        a, b = tuned.result(i), plain.result(i)
        assert a["status"] == 0 and b["status"] == 0
        for x, y in zip(tuned.state(i), plain.state(i)):
            assert np.array_equal(x, y), i
