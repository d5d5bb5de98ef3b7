"""ctypes wrapper of oracle/libvpm_cpu.so -- the CPU baseline (test infrastructure).

Used only by tests (to cross-check the NumPy oracle) and by bench.py's
cpu_baseline leg (timed on the host cores of the machine that holds the GPU).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from .build_cpu import LIB, build, runtime_lib

_lib = None
_D = C.POINTER(C.c_double)
_L = C.POINTER(C.c_int64)


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        path, _ = runtime_lib()
        lib = C.CDLL(str(path))
        for name in ("vpm_cpu_selfinduced_velocity", "vpm_cpu_strength_gradients"):
            getattr(lib, name).argtypes = [C.c_int64, _D, _D, _D, C.c_int64, _L, _D]
            getattr(lib, name).restype = None
        lib.vpm_cpu_strength_gradients_pse.argtypes = [C.c_int64, _D, _D, _D, _D, C.c_int64, _L, _D]
        lib.vpm_cpu_strength_gradients_pse.restype = None
        for name in ("vpm_cpu_cuda_selfinduced_velocity", "vpm_cpu_cuda_vorticity"):
            getattr(lib, name).argtypes = [C.c_int64, _D, _D, _D, C.c_int64, _L, _D]
            getattr(lib, name).restype = None
        lib.vpm_cpu_cuda_strength_gradients_pse.argtypes = [C.c_int64, _D, _D, _D, _D, C.c_int64,
                                                            _L, _D]
        lib.vpm_cpu_cuda_strength_gradients_pse.restype = None
        lib.vpm_cpu_cuda_enstrophy_rows.argtypes = [C.c_int64, _D, _D, _D, _D]
        lib.vpm_cpu_cuda_enstrophy_rows.restype = None
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


# ---- the GPU plugin's forms (ParticlesPhysicsCuda.py) -------------------
def cuda_selfinduced_velocity(x, a, s, tidx=None):
    """ParticlesPhysicsCuda.py:133-191 (sigma_ij mean, j != i)."""
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    load().vpm_cpu_cuda_selfinduced_velocity(len(x), _p(x), _p(a), _p(s), len(tidx),
                                             _p(tidx, _L), _p(out))
    return out


def cuda_vorticity(x, a, s, tidx=None):
    """ParticlesPhysicsCuda.py:291-346 at the particles' own positions."""
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    load().vpm_cpu_cuda_vorticity(len(x), _p(x), _p(a), _p(s), len(tidx), _p(tidx, _L), _p(out))
    return out


def cuda_strength_gradients_pse(x, a, s, nu, tidx=None):
    """ParticlesPhysicsCuda.py:478-530 (nu_i, A = rho^2 + 1)."""
    x, a, s, tidx, out = _prep(x, a, s, tidx)
    nu = np.ascontiguousarray(nu, dtype=np.float64)
    load().vpm_cpu_cuda_strength_gradients_pse(len(x), _p(x), _p(a), _p(s), _p(nu), len(tidx),
                                               _p(tidx, _L), _p(out))
    return out


def cuda_total_enstrophy(x, a, s):
    """ParticlesPhysicsCuda.py:756-777,867-920: per-target rows summed in order."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    rows = np.empty(len(x))
    load().vpm_cpu_cuda_enstrophy_rows(len(x), _p(x), _p(a), _p(s), _p(rows))
    return float(np.sum(rows)) / (8.0 * np.pi) / (4.0 * np.pi)


class CudaSums:
    """The pair sums a ReferenceStepper needs (GPU-plugin forms), from this
    library instead of the NumPy oracle: full-resolution steps in seconds."""

    @staticmethod
    def selfinduced_velocity(x, a, s, variant="cuda"):
        assert variant == "cuda"
        return cuda_selfinduced_velocity(x, a, s)

    @staticmethod
    def vorticity_field_at(y, x, a, s, variant="cuda"):
        assert variant == "cuda" and y is x
        return cuda_vorticity(x, a, s)

    @staticmethod
    def total_enstrophy(x, a, s, variant="cuda"):
        assert variant == "cuda"
        return cuda_total_enstrophy(x, a, s)

    @staticmethod
    def strength_gradients(x, a, s):
        return strength_gradients(x, a, s)

    @staticmethod
    def strength_gradients_pse(x, a, s, nu, variant="cuda"):
        assert variant == "cuda"
        return cuda_strength_gradients_pse(x, a, s, nu)


def isa() -> str:
    """ISA level of the library loaded on this host (x86-64-v3 or -v4)."""
    return runtime_lib()[1]


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
