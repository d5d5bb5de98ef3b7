"""ctypes wrapper of oracle/libvpm_cpu.so -- the CPU baseline (test infrastructure).

Used only by tests (to cross-check the NumPy oracle) and by bench.py's
cpu_baseline leg (timed on the host cores of the machine that holds the GPU).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from .build_cpu import LIB, build

_lib = None
_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_int64)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        for name in ("vpm_cpu_selfinduced_velocity", "vpm_cpu_strength_gradients"):
            getattr(lib, name).argtypes = [C.c_int64, _D, _D, _D, C.c_int64, _L, _D]
            getattr(lib, name).restype = None
        lib.vpm_cpu_strength_gradients_pse.argtypes = [C.c_int64, _D, _D, _D, _D, C.c_int64, _L, _D]
        lib.vpm_cpu_strength_gradients_pse.restype = None
        lib.vpm_cpu_max_threads.restype = C.c_int
        lib.vpm_cpu_set_threads.argtypes = [C.c_int]
        lib.vpm_cpu_set_threads.restype = None
        _lib = lib
    return _lib


def _p(a, t=_D):
    return a.ctypes.data_as(t)


def _prep(x, a, s, tidx):
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    if tidx is None:
        tidx = np.arange(len(x), dtype=np.int64)
    tidx = np.ascontiguousarray(tidx, dtype=np.int64)
    return x, a, s, tidx, np.empty((len(tidx), 3))


def selfinduced_velocity(x, a, s, tidx=None):
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    load().vpm_cpu_selfinduced_velocity(len(x), _p(x), _p(a), _p(s), len(tidx), _p(tidx, _L),
                                        _p(out))
    return out


def strength_gradients(x, a, s, tidx=None):
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    load().vpm_cpu_strength_gradients(len(x), _p(x), _p(a), _p(s), len(tidx), _p(tidx, _L),
                                      _p(out))
    return out


def strength_gradients_pse(x, a, s, nu, tidx=None):
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    load().vpm_cpu_strength_gradients_pse(len(x), _p(x), _p(a), _p(s), _p(nu), len(tidx),
                                          _p(tidx, _L), _p(out))
    return out


def threads() -> int:
    return int(load().vpm_cpu_max_threads())


def time_velocity_stretching(x, a, s, m, nthreads=None):
    """Times velocity + stretching (two passes, as the reference runs them) for
    m strided targets against all sources on `nthreads` threads (default: all).
    Returns (pairs_per_s, seconds, (u, S), tidx)."""
    n = len(x)
    step = max(1, n // m)
    tidx = np.arange(0, n, step, dtype=np.int64)[:m]
    lib = load()
    full = threads()
    if nthreads:
        lib.vpm_cpu_set_threads(int(nthreads))
    try:
        t0 = time.perf_counter()
        u = selfinduced_velocity(x, a, s, tidx)
        S = strength_gradients(x, a, s, tidx)
        dt = time.perf_counter() - t0
    finally:
        lib.vpm_cpu_set_threads(full)
    return len(tidx) * n / dt, dt, (u, S), tidx


def host_cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def omp_threads_env() -> int:
    v = os.environ.get("OMP_NUM_THREADS")
    return int(v) if v and v.isdigit() else threads()
