"""ctypes front of the C++/OpenMP CPU port of the host contraction route
(oracle/cpu_port.cpp): descriptor replay with transforms, a cache-blocked
DGEMM and the output transform on host cores.

TEST INFRASTRUCTURE ONLY (tests and bench.py's CPU-baseline leg).
"""
import ctypes as C
import os

import numpy as np

from tenet_amd import _native as N

from .build import LIB

_lib = None
_dp = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from .build import build
            build()
        L = C.CDLL(LIB)
        L.tntc_transform.argtypes = [C.POINTER(N.TransformDesc), _dp, _dp, C.c_double, C.c_double, C.c_int]
        L.tntc_gemm_jobs.argtypes = [C.POINTER(N.GemmJob), C.c_size_t, _dp, _dp, _dp, C.c_int, C.c_int,
                                     C.c_double, C.c_double]
        L.tntc_contract.argtypes = [C.POINTER(N.ContractDesc), _dp, _dp, _dp, C.c_double, C.c_double,
                                    _dp, _dp, _dp]
        L.tntc_threads.restype = C.c_int
        L.tntc_set_threads.argtypes = [C.c_int]
        L.tntc_set_threads.restype = None
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"cpu port error {rc}")


def threads():
    return int(lib().tntc_threads())


def set_threads(n):
    lib().tntc_set_threads(int(n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def transform(desc, src, dst=None, alpha=1.0, beta=0.0, overwrite=True):
    d = desc.contents
    src = np.ascontiguousarray(src, dtype=np.float64)
    if dst is None:
        dst = np.zeros(d.dst_len)
    _check(lib().tntc_transform(desc, _p(src), _p(dst), alpha, beta, 1 if overwrite else 0))
    return dst


class Contract:
    """One compiled contraction replayed on the CPU with reusable workspaces."""

    def __init__(self, cdesc, dst_len):
        self.desc = cdesc
        d = cdesc.contents
        self.ws_l = np.empty(d.lhs_len) if d.lhs else None
        self.ws_r = np.empty(d.rhs_len) if d.rhs else None
        self.ws_c = np.empty(d.core_len) if d.out else None
        self.dst_len = dst_len

    def __call__(self, lhs, rhs, dst=None, alpha=1.0, beta=0.0):
        lhs = np.ascontiguousarray(lhs, dtype=np.float64)
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if dst is None:
            dst = np.empty(self.dst_len)
            beta = 0.0
        _check(lib().tntc_contract(self.desc, _p(dst), _p(lhs), _p(rhs), alpha, beta,
                                   _p(self.ws_l), _p(self.ws_r), _p(self.ws_c)))
        return dst


def gemm_jobs(jobs, lhs, rhs, dst, lhs_op=0, rhs_op=0, alpha=1.0, beta=0.0):
    arr = (N.GemmJob * len(jobs))(*[N.GemmJob(*j) for j in jobs])
    _check(lib().tntc_gemm_jobs(arr, len(jobs), _p(dst), _p(lhs), _p(rhs), lhs_op, rhs_op, alpha, beta))
    return dst
