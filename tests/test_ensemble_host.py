"""Host-side checks of the ensemble layer (no GPU): the device footprint
used for admission, and argument errors."""
import numpy as np
import pytest

import atmospheric_flows_amd as af
from atmospheric_flows_amd import _native as N


def test_footprint_scales_with_the_mesh_and_precision():
    b1 = af.device_bytes(3, 64, 32)
    b2 = af.device_bytes(3, 128, 64)
    ns1, ns2 = 64 * 32 * 16, 128 * 64 * 16
    # ~ 100-130 FP64 fields of N_s on the fine level plus the multigrid levels
    assert 90 * 8 * ns1 < b1 < 140 * 8 * ns1
    assert abs(b2 / b1 - 4.0) < 0.1  # fixed-size reduction and record buffers aside
    # the mixed form adds exactly the FP64 inverses of A's cell blocks, symmetric and packed:
    # N (N + 1) / 2 doubles per cell for N = (k+1)^2
    bm = af.device_bytes(3, 128, 64, precision=1)
    assert bm - b2 == 128 * 64 * 136 * 8
    # a longer GMRES basis adds exactly its vectors (leading dimension N_s + 32)
    assert af.device_bytes(3, 128, 64, gmres_restart=30) - b2 == 2 * (ns2 + 32) * 8


def test_footprint_of_the_headline_mesh_fits_hbm_many_times():
    b = af.device_bytes(3, 512, 256, precision=1)
    assert 1.0e9 < b < 3.0e9  # ~ 2 GB: about 100 such members fit in 256 GB of HBM


def test_footprint_rejects_bad_meshes():
    with pytest.raises(af.AtmoError):
        af.device_bytes(4, 64, 32)
    with pytest.raises(af.AtmoError):
        af.device_bytes(2, 0, 32)


def test_ensemble_arguments_are_checked():
    import ctypes
    h = ctypes.c_void_p()
    assert N.load().atmo_ensemble_create(0, 1e9, 0, ctypes.byref(h)) == N.ATMO_INVALID_ARG
    assert N.load().atmo_ensemble_create(4, -1.0, 0, ctypes.byref(h)) == N.ATMO_INVALID_ARG
