"""atmo_dot_noalloc, the scalar product that allocates nothing (its block
partials live in the library's static device array), against atmo_vec_dot on
the same data: bitwise equal, on two calls in a row, the second with a smaller
input (the shared array keeps the first call's partials)."""
import ctypes

import numpy as np
import pytest

import atmospheric_flows_amd as af
from atmospheric_flows_amd import _native as N

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            return ctypes.CDLL(name)
        except OSError:
            continue
    pytest.fail("libamdhip64 not found")


def _dev(v):
    p = ctypes.c_void_p()
    N.check(N.load().atmo_vec_device_ptr(v.handle, ctypes.byref(p)))
    return p


def test_noalloc_dot_matches_context_dot():
    ctx = af.DeviceContext(3, 16, 8, h=af.witch_of_agnesi(16), Ma=1.0, Fr=1.0 / np.sqrt(9.81), dt=4.0)
    hip = _hip()
    L = N.load()
    rng = np.random.default_rng(11)
    res = ctx.vector(af.BLOCK_PRESSURE)  # its first entry receives the result
    for block in (af.BLOCK_VELOCITY, af.BLOCK_PRESSURE):  # 2 N_s, then N_s doubles
        n = ctx.vector(block).size
        x, y = rng.standard_normal(n), rng.standard_normal(n)
        dx, dy = ctx.vector(block, x), ctx.vector(block, y)
        want = dx * dy  # atmo_vec_dot: the context's reduction buffer
        print(f"[case] atmo_dot_noalloc n={n}", flush=True)
        st = L.atmo_dot_noalloc(_dev(dx), _dev(dy), ctypes.c_size_t(n), None, _dev(res))
        assert st == N.ATMO_OK, L.atmo_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        got = res.download()[0]
        print(f"n={n}: noalloc {got!r}, context {want!r}", flush=True)
        assert got == want
        assert got == pytest.approx(float(x @ y), rel=1e-13, abs=1e-13 * np.sqrt(n))


def test_noalloc_dot_rejects_bad_arguments():
    L = N.load()
    assert L.atmo_dot_noalloc(None, None, ctypes.c_size_t(4), None, None) == N.ATMO_INVALID_ARG
    big = ctypes.c_size_t((1 << 36) + 1)
    one = ctypes.c_void_p(1)
    assert L.atmo_dot_noalloc(one, one, big, None, one) == N.ATMO_INVALID_ARG
    assert b"ATMO_DOT_NOALLOC_MAX_N" in L.atmo_last_error(None)
