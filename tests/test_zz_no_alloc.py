"""thistle_unique_total_noalloc (device pointers, static scratch, no
allocation) against thistle_unique_total on two calls in a row, the second
smaller, so that the second runs on the scratch the first left behind."""
import ctypes as C
import os

import numpy as np
import pytest

from thistle_amd import lib as L
from thistle_amd import scattering as sc

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library itself loaded (one runtime per process)."""
    L.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert paths, "libthistle_hip.so did not load libamdhip64"
    h = C.CDLL(sorted(paths)[0])
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


H2D, D2H = 1, 2


class Dev:
    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(int(nbytes), 8)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, dtype, n):
        out = np.zeros(n, dtype=dtype)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes,
                                      D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _rows(seed, nrow, ncol):
    rng = np.random.default_rng(seed)
    base = rng.integers(1, 40, size=(max(nrow // 3, 1), ncol), dtype=np.int32)
    mat = base[rng.integers(0, base.shape[0], size=nrow)]
    return np.asfortranarray(mat), rng.standard_normal(nrow)


def _noalloc(hip, lib, mat, tot):
    nrow, ncol = mat.shape
    dm, dt = Dev(hip, 4 * mat.size), Dev(hip, 8 * nrow)
    drm, drt = Dev(hip, 4 * mat.size), Dev(hip, 8 * nrow)
    diu, dir_, dc = Dev(hip, 4 * nrow), Dev(hip, 4 * nrow), Dev(hip, 8)
    dm.put(mat.ravel(order="F"))
    dt.put(tot)
    o = L.make_opts()
    rc = lib.thistle_unique_total_noalloc(dm.p, nrow, ncol, dt.p, drm.p, drt.p, diu.p, dir_.p,
                                          dc.p, C.byref(o))
    assert rc == 0
    assert hip.hipDeviceSynchronize() == 0
    u = int(dc.get(np.int64, 1)[0])
    rm = drm.get(np.int32, nrow * ncol).reshape(ncol, nrow).T[:u]
    out = (u, rm, drt.get(np.float64, nrow)[:u], diu.get(np.int32, nrow)[:u],
           dir_.get(np.int32, nrow))
    for d in (dm, dt, drm, drt, diu, dir_, dc):
        d.free()
    return out


@pytest.mark.parametrize("ncol", [4])
def test_noalloc_matches_unique_total_twice(ncol):
    hip = _hip()
    lib = L.load()
    for call, nrow in enumerate((20000, 3000)):
        mat, tot = _rows(100 + call, nrow, ncol)
        print(f"noalloc call {call}: nrow={nrow} ncol={ncol}", flush=True)
        u, rm, rt, iu, ir = _noalloc(hip, lib, mat, tot)
        ref = sc.unique_total(mat, tot)
        rmat, rtot, iuni, irec = ref
        assert u == rmat.shape[0]
        assert np.array_equal(rm, rmat)
        assert np.array_equal(rt, rtot)
        assert np.array_equal(iu, iuni) and np.array_equal(ir, irec)


def test_noalloc_rejects_oversize():
    lib = L.load()
    o = L.make_opts()
    c = C.c_int64()
    rc = lib.thistle_unique_total_noalloc(None, (1 << 18) + 1, 4, None, None, None, None, None,
                                          C.byref(c), C.byref(o))
    assert rc == 1
