"""edus_wannier_hk_noalloc: the Wannier contraction with every device buffer in
the library's fixed scratch array (no device allocation in the call), against
edus_wannier_hk through the C-ABI: two calls in a row, the second smaller, and
the rejection of a model too large for the scratch array."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import N, ROOT

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
N.hip.edus_wannier_hk_noalloc.argtypes = [_ip, C.c_int32, C.c_int32, _ip, _dp, _dp, _dp, _dp]


def _scratch_bytes():
    """EDUS_NOALLOC_SCRATCH_BYTES of include/edus.h."""
    text = open(os.path.join(ROOT, "include", "edus.h")).read()
    m = re.search(r"#define EDUS_NOALLOC_SCRATCH_BYTES\s+\(\(int64_t\)(\d+) << (\d+)\)", text)
    return int(m.group(1)) << int(m.group(2))


def _model(nw, rmax, seed):
    rng = np.random.default_rng(seed)
    R = np.array([[0, a, b] for a in range(-rmax, rmax + 1) for b in range(-rmax, rmax + 1)], dtype=np.int32)
    nR = len(R)
    HR = rng.normal(size=(nR, nw, nw)) + 1j * rng.normal(size=(nR, nw, nw))
    rR = rng.normal(size=(nR, nw, nw, 3)) + 1j * rng.normal(size=(nR, nw, nw, 3))
    return R, HR, rR


def _call(fn, nk3, R, HR, rR, nw, with_d=True):
    nk3 = np.ascontiguousarray(nk3, dtype=np.int32)
    nk = int(np.prod(nk3))
    H = np.full((nk, nw, nw), np.nan, dtype=np.complex128)
    D = np.full((nk, nw, nw, 3), np.nan, dtype=np.complex128)
    HR = np.ascontiguousarray(HR)
    rR = np.ascontiguousarray(rR)
    rc = fn(nk3.ctypes.data_as(_ip), len(R), nw, R.ctypes.data_as(_ip), HR.view(np.float64).ctypes.data_as(_dp),
            rR.view(np.float64).ctypes.data_as(_dp), H.view(np.float64).ctypes.data_as(_dp),
            D.view(np.float64).ctypes.data_as(_dp) if with_d else None)
    return rc, H, D


@pytest.mark.parametrize("with_d", [True, False])
def test_noalloc_matches_allocating_entry_point_twice(with_d):
    cases = [((1, 24, 20), 12, 3, 1), ((1, 7, 9), 5, 1, 2)]  # the second call is smaller
    for nk3, nw, rmax, seed in cases:
        R, HR, rR = _model(nw, rmax, seed)
        rc0, H0, D0 = _call(N.hip.edus_wannier_hk, nk3, R, HR, rR, nw, with_d)
        rc1, H1, D1 = _call(N.hip.edus_wannier_hk_noalloc, nk3, R, HR, rR, nw, with_d)
        print(f"nk3={nk3} nw={nw} nR={len(R)} with_d={with_d}: rc {rc0} {rc1}", flush=True)
        assert rc0 == 0 and rc1 == 0, N.hip.edus_last_error()
        np.testing.assert_array_equal(H1, H0)
        if with_d:
            np.testing.assert_array_equal(D1, D0)
        else:
            assert np.all(np.isnan(D1))


def test_noalloc_chunks_a_grid_larger_than_one_chunk():
    """A grid whose outputs exceed the scratch left after the model: the call
    walks it in chunks and still matches."""
    nw, rmax = 16, 2
    R, HR, rR = _model(nw, rmax, 3)
    per_k = 128 * nw * nw  # contraction output, H and D of one k
    nk_side = int(np.ceil(np.sqrt(_scratch_bytes() / per_k))) + 8
    nk3 = (1, nk_side, nk_side)
    rc0, H0, D0 = _call(N.hip.edus_wannier_hk, nk3, R, HR, rR, nw)
    rc1, H1, D1 = _call(N.hip.edus_wannier_hk_noalloc, nk3, R, HR, rR, nw)
    assert rc0 == 0 and rc1 == 0, N.hip.edus_last_error()
    np.testing.assert_array_equal(H1, H0)
    np.testing.assert_array_equal(D1, D0)


def test_noalloc_rejects_a_model_above_the_scratch_size():
    nw = 32
    cap = _scratch_bytes()
    nR = cap // (64 * nw * nw) + 1  # H(R) and r(R) alone exceed the array
    R = np.zeros((nR, 3), dtype=np.int32)
    HR = np.zeros((nR, nw, nw), dtype=np.complex128)
    rR = np.zeros((nR, nw, nw, 3), dtype=np.complex128)
    nk3 = np.array([1, 4, 4], dtype=np.int32)
    H = np.zeros((16, nw, nw), dtype=np.complex128)
    rc = N.hip.edus_wannier_hk_noalloc(nk3.ctypes.data_as(_ip), nR, nw, R.ctypes.data_as(_ip),
                                       HR.view(np.float64).ctypes.data_as(_dp), rR.view(np.float64).ctypes.data_as(_dp),
                                       H.view(np.float64).ctypes.data_as(_dp), None)
    assert rc == -4
    assert b"scratch" in N.hip.edus_last_error()
