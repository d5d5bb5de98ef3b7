"""ctypes binding of libvpm_hip.so (the C ABI declared in include/vpm_hip.h).

The library is built in-tree by ``python -m openonda_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback in this package.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvpm_hip.so"

VPM_OK = 0
VPM_ERR_ARG = -1
VPM_ERR_DEVICE = -2
VPM_ERR_LAUNCH = -3
VPM_ERR_WORKSPACE = -4
VPM_ERR_UNSUPPORTED = -5
VPM_ERR_NONFINITE = -6

# vpm_scan_particles flag bits
BAD_POSITION = 1
BAD_STRENGTH = 2
BAD_RADIUS = 4
BAD_VISCOSITY = 8
BAD_NAMES = {BAD_POSITION: "positions", BAD_STRENGTH: "strengths", BAD_RADIUS: "radii",
             BAD_VISCOSITY: "viscosities"}

VARIANT_CUDA = 0
VARIANT_CPU = 1

OUT_VEL = 1
OUT_STRETCH = 2
OUT_PSE = 4
OUT_VORT = 8
OUT_ENSTROPHY = 16

PREC_F64 = 0
PREC_F32C = 1

SIGMA_GENERAL = 0
SIGMA_UNIFORM = 1

G64_AUTO = 0
G64_ALWAYS = 1
G64_NEVER = 2


class VpmError(RuntimeError):
    """A libvpm_hip call returned a non-zero status."""

    def __init__(self, func, code, msg):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


class NonFiniteError(ValueError):
    """Particle data the pair sums cannot take: NaN/Inf positions, strengths,
    radii or viscosities, radii <= 0 or viscosities < 0.  ``code`` is
    VPM_ERR_NONFINITE, ``bits`` the VPM_BAD_* flags found."""

    def __init__(self, where, bits):
        names = [v for k, v in BAD_NAMES.items() if bits & k]
        super().__init__(f"{where}: non-finite or out-of-range {', '.join(names)} "
                         f"(status {VPM_ERR_NONFINITE})")
        self.code = VPM_ERR_NONFINITE
        self.bits = bits


class Particles(C.Structure):
    _fields_ = [("n", C.c_int64), ("ld", C.c_int64), ("pos", C.c_void_p), ("str", C.c_void_p),
                ("sigma", C.c_void_p), ("nu", C.c_void_p)]


class Points(C.Structure):
    _fields_ = [("m", C.c_int64), ("ld", C.c_int64), ("pos", C.c_void_p)]


class ForwardOpts(C.Structure):
    _fields_ = [("outputs", C.c_uint32), ("precision", C.c_int32), ("sigma_mode", C.c_int32),
                ("presorted", C.c_int32), ("tile_config", C.c_int32), ("max_splits", C.c_int32),
                ("g64_mode", C.c_int32)]


class ForwardOut(C.Structure):
    _fields_ = [("ld", C.c_int64), ("vel", C.c_void_p), ("stretch", C.c_void_p),
                ("pse", C.c_void_p), ("vort", C.c_void_p), ("enstrophy", C.c_void_p)]


class AdjointIn(C.Structure):
    _fields_ = [("ld", C.c_int64), ("lam", C.c_void_p), ("mu", C.c_void_p), ("xi", C.c_void_p)]


class AdjointOut(C.Structure):
    _fields_ = [("ld", C.c_int64), ("dpos", C.c_void_p), ("dstr", C.c_void_p)]


class BatchDesc(C.Structure):
    _fields_ = [("nsets", C.c_int32), ("n_items", C.c_int32), ("grid", C.c_int32),
                ("set_off", C.c_void_p), ("part_set", C.c_void_p), ("item_set", C.c_void_p),
                ("item_block", C.c_void_p), ("set_item_first", C.c_void_p),
                ("set_item_count", C.c_void_p)]


_P = C.POINTER
_VP = C.c_void_p
_I64 = C.c_int64
_SZ = C.c_size_t

# name -> (restype, argtypes).  Every symbol declared in include/vpm_hip.h.
SIGNATURES = {
    "vpm_abi_version": (C.c_int, []),
    "vpm_last_error": (C.c_char_p, []),
    "vpm_device_count": (C.c_int, []),
    "vpm_get_induced_velocity_at": (C.c_int, [_P(Points), _P(Particles), _VP, _I64, _VP]),
    "vpm_get_selfinduced_velocity": (C.c_int, [_P(Particles), C.c_int, _VP, _I64, _VP]),
    "vpm_get_vorticity_field_at": (C.c_int, [_P(Points), _P(Particles), C.c_int, _VP, _I64, _VP]),
    "vpm_get_strength_gradients": (C.c_int, [_P(Particles), _VP, _I64, _VP]),
    "vpm_get_strength_gradients_pse": (C.c_int, [_P(Particles), C.c_int, _VP, _I64, _VP]),
    "vpm_reduction_workspace_bytes": (_SZ, [_I64]),
    "vpm_get_total_kinetic_energy": (C.c_int, [_P(Particles), C.c_int, _VP, _VP, _SZ, _VP]),
    "vpm_get_total_helicity": (C.c_int, [_P(Particles), C.c_int, _VP, _VP, _SZ, _VP]),
    "vpm_get_total_enstrophy": (C.c_int, [_P(Particles), C.c_int, _VP, _VP, _SZ, _VP]),
    "vpm_forward_workspace_bytes": (_SZ, [_I64, _P(ForwardOpts)]),
    "vpm_forward": (C.c_int, [_P(Particles), _P(ForwardOpts), _P(ForwardOut), _VP, _SZ, _VP]),
    "vpm_forward_prepare": (C.c_int, [_P(Particles), _P(ForwardOpts), _VP, _SZ, _VP]),
    "vpm_forward_run": (C.c_int, [_P(Particles), _P(ForwardOpts), _P(ForwardOut), _VP, _SZ, _VP]),
    "vpm_forward_noalloc": (C.c_int, [_P(Particles), _P(ForwardOpts), _P(ForwardOut), _VP]),
    "vpm_forward_noalloc_capacity": (_SZ, []),
    "vpm_batch_workspace_bytes": (_SZ, [_I64, C.c_int, C.c_int]),
    "vpm_batch_prepare": (C.c_int, [_P(Particles), _P(BatchDesc), C.c_int, _VP, _SZ, _VP]),
    "vpm_batch_run": (C.c_int, [_P(Particles), _P(BatchDesc), _P(ForwardOpts), _P(ForwardOut), _VP, _SZ, _VP]),
    "vpm_batch_les_update": (C.c_int, [_P(BatchDesc), _I64, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _SZ, _I64, _VP]),
    "vpm_adjoint_workspace_bytes": (_SZ, [_I64, _P(ForwardOpts)]),
    "vpm_adjoint": (C.c_int, [_P(Particles), _P(AdjointIn), _P(ForwardOpts), _P(AdjointOut), _VP, _SZ, _VP]),
    "vpm_induced_velocity_at_adjoint": (C.c_int, [_P(Points), _P(Particles), _VP, _I64, _VP, _VP, _I64, _VP]),
    "vpm_axpy3": (C.c_int, [_I64, _I64, _VP, _VP, C.c_double, _VP, _VP]),
    "vpm_rk4_combine": (C.c_int, [_I64, _I64, _VP, _VP, _VP, _VP, _VP, C.c_double, _VP, _VP]),
    "vpm_strength_update": (C.c_int, [_I64, _I64, C.c_int, C.c_double, _VP, _VP, _VP, _VP, _VP, _VP]),
    "vpm_core_spreading": (C.c_int, [_I64, _VP, _VP, _VP, C.c_double, _VP]),
    "vpm_relax_strengths": (C.c_int, [_I64, _I64, _VP, _VP, _VP]),
    "vpm_state_workspace_bytes": (_SZ, [_I64]),
    "vpm_diagnostics": (C.c_int, [_I64, _I64, _VP, _VP, _VP, _VP, _SZ, _VP]),
    "vpm_scan_particles": (C.c_int, [_VP, _VP, _VP]),
    "vpm_les_update": (C.c_int, [_I64, _I64, _VP, _VP, _VP, _VP, _VP, C.c_int, _VP, _SZ, _VP]),
    "vpm_graph_begin": (C.c_int, [_VP]),
    "vpm_graph_end": (C.c_int, [_VP, _P(C.c_void_p)]),
    "vpm_graph_launch": (C.c_int, [_VP, _VP]),
    "vpm_graph_destroy": (C.c_int, [_VP]),
    "vpm_tile_config_count": (C.c_int, []),
    "vpm_tile_config_default": (C.c_int, [C.c_uint32]),
    "vpm_batch_tile_config_default": (C.c_int, [C.c_uint32]),
    "vpm_tile_config_info": (C.c_int, [C.c_int] + [_P(C.c_int)] * 5),
    "vpm_autotune_workspace_bytes": (C.c_size_t, [C.c_int64, _P(ForwardOpts)]),
    "vpm_autotune": (C.c_int, [_P(Particles), _P(ForwardOpts), _P(ForwardOut), _VP, C.c_size_t,
                               _VP, C.c_int, _P(C.c_int), _P(C.c_float)]),
}

_lock = threading.Lock()
_lib = None


def lib_path() -> Path:
    return Path(os.environ.get("VPM_HIP_LIB", LIB_PATH))


def load() -> C.CDLL:
    """Loads libvpm_hip.so once; raises if it is absent or incomplete.

    torch is imported first: the PyTorch-ROCm wheel carries its own HIP
    runtime under the same soname, and the library must bind to that one
    (a second runtime in the process would not share streams or memory)."""
    global _lib
    import torch  # noqa: F401  (binds libamdhip64.so.7 to torch's runtime)

    with _lock:
        if _lib is not None:
            return _lib
        p = lib_path()
        if not p.exists():
            raise RuntimeError(
                f"{p} is missing: build the HIP library first (python -m openonda_amd.build); "
                "this package has no CPU fallback")
        lib = C.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)  # AttributeError = stale library
            fn.restype = res
            fn.argtypes = args
        if lib.vpm_abi_version() != 3:
            raise RuntimeError("libvpm_hip.so ABI version mismatch")
        _lib = lib
        return lib


def call(name, *args):
    """Calls ``name`` and raises VpmError on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != VPM_OK:
        msg = lib.vpm_last_error().decode(errors="replace")
        raise VpmError(name, rc, msg)
    return rc
