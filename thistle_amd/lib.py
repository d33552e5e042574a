"""ctypes binding of libthistle_hip.so (the C ABI in include/thistle_scatter.h).

The library is built in-tree (``thistle_amd/libthistle_hip.so``).  Loading
fails loudly when it is missing: there is no CPU fallback for the device path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libthistle_hip.so")

THISTLE_OK = 0
STATUS = {0: "ok", 1: "invalid argument", 2: "a GTO has l, m or n > 4",
          3: "quartet LLmax + 1 > 20", 4: "a group is not one contiguous single-shell run",
          5: "HIP / rocBLAS failure", 6: "device allocation failed",
          7: "calculation type must be 1, 2, 5, 6, 7 or 8 (total / elastic j0, total / "
             "elastic electron, total / elastic j2)",
          8: "cutoffmd above THISTLE_CUTOFFMD_MAX (1e-18): the factorised contraction keeps "
             "every MD product"}
CUTOFFMD_MAX = 1e-18
# 1E-30 as a default real: the 2-RDM selection threshold of the reference's
# builders (src/twordms.f90:514, src/one_rdm.f90:1312)
RDM_CUTOFF = 1.0000000031710769e-30
FP64, MIXED = 0, 1

EXPORTS = [
    "thistle_total_scattering_calculation", "thistle_total_from_twordm",
    "thistle_total_scattering_adjoint",
    "thistle_elastic_from_onerdm", "thistle_twordm_dense", "thistle_twordm_select",
    "thistle_createtwordm", "thistle_nuclei_electron",
    "thistle_mcci_to_bit", "thistle_createtwordm_bit", "thistle_one_rdm_bit",
    "thistle_createtwordm_bit_fci",
    "thistle_total_scattering_calculation_2",
    "thistle_onerdm", "thistle_createonerdm", "thistle_quartet_list",
    "thistle_quartet_list_screened",
    "thistle_unique_total", "thistle_unique_total_noalloc", "thistle_pair_tables", "thistle_plan_create",
    "thistle_plan_execute", "thistle_plan_result", "thistle_plan_timings",
    "thistle_plan_destroy", "thistle_plan_set_quartets", "thistle_plan_kernel_timings", "thistle_plan_mixed_fallback", "thistle_batch_create", "thistle_batch_execute",
    "thistle_batch_info", "thistle_batch_tuning", "thistle_batch_destroy", "thistle_device_count",
    "thistle_device_synchronize",
    "thistle_version",
]


class ThistleError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = int(code)
        super().__init__(f"{where}: status {code} ({STATUS.get(int(code), 'unknown')})")


class Basis(C.Structure):
    _fields_ = [("ngto", C.c_int32), ("ng", C.c_int32), ("maxl", C.c_int32),
                ("ga", C.c_void_p), ("l", C.c_void_p), ("m", C.c_void_p), ("n", C.c_void_p),
                ("xx", C.c_void_p), ("yy", C.c_void_p), ("zz", C.c_void_p),
                ("group", C.c_void_p), ("nmo", C.c_int32), ("mmod", C.c_void_p),
                ("nq", C.c_int32), ("q", C.c_void_p), ("cutoffz", C.c_double),
                ("cutoffmd", C.c_double), ("cutoffcentre", C.c_double)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("rdm_cutoff", C.c_double),
                ("stream", C.c_void_p), ("device", C.c_int32), ("waves", C.c_int32)]


_LIB = None


def load():
    """Load the in-tree library; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` (the scattering path has no CPU fallback)")
        lib = C.CDLL(LIBPATH)
        _declare(lib)
        _LIB = lib
    return _LIB


def _declare(L):
    p, i32, i64, d = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.thistle_total_scattering_calculation.argtypes = [i32, i32, p, p, i32, i32, i32, p, i32,
                                                       i32, i32, p, p, p, p, p]
    L.thistle_total_scattering_calculation.restype = i32
    L.thistle_total_from_twordm.argtypes = [p, i64, p, p, p, p]
    L.thistle_total_from_twordm.restype = i32
    L.thistle_total_scattering_adjoint.argtypes = [p, i64, p, p, i32, p, p, p, p, p, p]
    L.thistle_total_scattering_adjoint.restype = i32
    L.thistle_elastic_from_onerdm.argtypes = [p, i32, p, p, p]
    L.thistle_elastic_from_onerdm.restype = i32
    L.thistle_twordm_dense.argtypes = [p, i32, i32, p, i32, i32, i32, p]
    L.thistle_twordm_dense.restype = i32
    L.thistle_twordm_select.argtypes = [p, i32, d, p, p, i64]
    L.thistle_twordm_select.restype = i64
    L.thistle_createtwordm.argtypes = [p, i32, i32, p, i32, i32, i32, d, p, p, p, i64]
    L.thistle_createtwordm.restype = i64
    L.thistle_mcci_to_bit.argtypes = [p, p, p, p, i64, i32, p, p, p]
    L.thistle_mcci_to_bit.restype = i32
    L.thistle_createtwordm_bit.argtypes = [p, p, p, p, i64, d, p, p, p, i64]
    L.thistle_createtwordm_bit.restype = i64
    L.thistle_createtwordm_bit_fci.argtypes = [p, p, p, i64, p, i64, p, p, p, d, p, p, p, i64]
    L.thistle_createtwordm_bit_fci.restype = i64
    L.thistle_one_rdm_bit.argtypes = [p, p, p, i64, p, i32, p, p]
    L.thistle_one_rdm_bit.restype = i32
    L.thistle_total_scattering_calculation_2.argtypes = [i32, p, i64, p, p, p, p, i32, p, p, p]
    L.thistle_total_scattering_calculation_2.restype = i32
    L.thistle_nuclei_electron.argtypes = [p, i32, p, i32, p, p, p, p]
    L.thistle_nuclei_electron.restype = i32
    L.thistle_onerdm.argtypes = [p, i32, i32, p, i32, i32, i32, p]
    L.thistle_onerdm.restype = i32
    L.thistle_createonerdm.argtypes = [p, i32, i32, p, i32, i32, i32, p, p]
    L.thistle_createonerdm.restype = i32
    L.thistle_quartet_list.argtypes = [p, p, i64, p]
    L.thistle_quartet_list.restype = i64
    L.thistle_quartet_list_screened.argtypes = [i32, p, i64, p, p, i32, p, p, i64, p]
    L.thistle_quartet_list_screened.restype = i64
    L.thistle_unique_total.argtypes = [p, i64, i32, p, p, p, p, p, p]
    L.thistle_unique_total.restype = i64
    L.thistle_unique_total_noalloc.argtypes = [p, i64, i32, p, p, p, p, p, p, p]
    L.thistle_unique_total_noalloc.restype = i32
    L.thistle_pair_tables.argtypes = [p, p, p, p, p]
    L.thistle_pair_tables.restype = i32
    L.thistle_plan_create.argtypes = [i32, p, i64, p, p, i32, p, p, C.POINTER(C.c_void_p)]
    L.thistle_plan_create.restype = i32
    L.thistle_plan_execute.argtypes = [p, p]
    L.thistle_plan_execute.restype = i32
    L.thistle_plan_result.argtypes = [p, p]
    L.thistle_plan_result.restype = i32
    L.thistle_plan_timings.argtypes = [p, p, p, p]
    L.thistle_plan_timings.restype = i32
    L.thistle_plan_destroy.argtypes = [p]
    L.thistle_plan_destroy.restype = None
    L.thistle_plan_kernel_timings.argtypes = [p, p, p]
    L.thistle_plan_kernel_timings.restype = i32
    L.thistle_plan_set_quartets.argtypes = [p, p, i64]
    L.thistle_plan_set_quartets.restype = i32
    L.thistle_plan_mixed_fallback.argtypes = [p, p, p, i32, p]
    L.thistle_plan_mixed_fallback.restype = i32
    L.thistle_batch_create.argtypes = [i32, p, p, p, p, p, C.POINTER(C.c_void_p)]
    L.thistle_batch_create.restype = i32
    L.thistle_batch_execute.argtypes = [p, p]
    L.thistle_batch_execute.restype = i32
    L.thistle_batch_info.argtypes = [p, p, p, p, p]
    L.thistle_batch_info.restype = i32
    L.thistle_batch_tuning.argtypes = [p, p, p]
    L.thistle_batch_tuning.restype = i32
    L.thistle_batch_destroy.argtypes = [p]
    L.thistle_batch_destroy.restype = None
    L.thistle_device_count.argtypes = []
    L.thistle_device_count.restype = i32
    L.thistle_device_synchronize.argtypes = [i32]
    L.thistle_device_synchronize.restype = i32
    L.thistle_version.argtypes = []
    L.thistle_version.restype = C.c_char_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _Keep:
    """Holds the numpy arrays a Basis points into."""

    def __init__(self):
        self.arrs = []

    def i32(self, a, order="C"):
        a = np.require(np.asarray(a), dtype=np.int32, requirements=[order])
        self.arrs.append(a)
        return a

    def f64(self, a, order="C"):
        a = np.require(np.asarray(a), dtype=np.float64, requirements=[order])
        self.arrs.append(a)
        return a


def make_basis(ga, l, m, n, xx, yy, zz, group, mmod, q, cutoffz, cutoffmd, cutoffcentre,
               maxl=None, ng=None):
    k = _Keep()
    ga, xx, yy, zz, q = (k.f64(v) for v in (ga, xx, yy, zz, q))
    l, m, n, group = (k.i32(v) for v in (l, m, n, group))
    mm = k.f64(mmod, "F") if mmod is not None else None
    b = Basis()
    b.ngto = int(l.size)
    b.ng = int(group.max()) if ng is None else int(ng)
    b.maxl = int(l.max()) if maxl is None else int(maxl)
    b.ga, b.l, b.m, b.n = _ptr(ga), _ptr(l), _ptr(m), _ptr(n)
    b.xx, b.yy, b.zz, b.group = _ptr(xx), _ptr(yy), _ptr(zz), _ptr(group)
    b.nmo = int(mm.shape[0]) if mm is not None else 0
    b.mmod = _ptr(mm)
    b.nq = int(q.size)
    b.q = _ptr(q)
    b.cutoffz, b.cutoffmd, b.cutoffcentre = float(cutoffz), float(cutoffmd), float(cutoffcentre)
    b._keep = k
    return b


def make_opts(precision=FP64, rdm_cutoff=RDM_CUTOFF, stream=None, device=0, waves=0):
    o = Opts()
    o.precision = int(precision)
    o.rdm_cutoff = float(rdm_cutoff)
    o.stream = stream
    o.device = int(device)
    o.waves = int(waves)
    return o


def device_count():
    return int(load().thistle_device_count())


def synchronize(device=0):
    check(load().thistle_device_synchronize(int(device)), "device_synchronize")


def check(code, where):
    if code != THISTLE_OK:
        raise ThistleError(code, where)
