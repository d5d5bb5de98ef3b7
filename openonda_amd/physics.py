"""Tensor-level API of the HIP pair-sum kernels.

All tensors are float64 PyTorch-ROCm tensors on the HIP device, in SoA layout:
vector fields have shape ``(3, n)`` (row c = component c), per-particle
scalars shape ``(n,)``.  Work is enqueued on the current torch stream of the
tensors' device; nothing synchronises except the ``*_sync`` helpers.

This is the fast path underneath the NumPy drop-in module
(:mod:`openonda_amd.ParticlesPhysicsHip`) and :class:`ParticleSystem`.
Reference interface being replaced: the eight functions of
``OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py`` (see each function).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L

VARIANTS = {"cuda": L.VARIANT_CUDA, "cpu": L.VARIANT_CPU}
PRECISIONS = {"f64": L.PREC_F64, "f32c": L.PREC_F32C}
G64_MODES = {"auto": L.G64_AUTO, "always": L.G64_ALWAYS, "never": L.G64_NEVER}
OUTPUTS = {"vel": L.OUT_VEL, "stretch": L.OUT_STRETCH, "pse": L.OUT_PSE, "vort": L.OUT_VORT,
           "enstrophy": L.OUT_ENSTROPHY}


def require_gpu():
    """Raises unless a HIP device is visible and the library is loadable."""
    L.load()
    if not torch.cuda.is_available():
        raise RuntimeError("openonda_amd needs a HIP GPU and none is visible "
                           "(there is no CPU fallback)")


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _vec(t: torch.Tensor, name: str, n: int | None = None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on the HIP device")
    if t.dim() != 2 or t.shape[0] != 3:
        raise ValueError(f"{name} must have shape (3, n), got {tuple(t.shape)}")
    if (t.shape[1] > 1 and t.stride(1) != 1) or (t.shape[1] > 0 and t.stride(0) < t.shape[1]):
        raise ValueError(f"{name} must be row-contiguous SoA (stride (ld, 1))")
    if n is not None and t.shape[1] != n:
        raise ValueError(f"{name} has {t.shape[1]} columns, expected {n}")
    return t


def _scal(t: torch.Tensor, name: str, n: int) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64:
        raise TypeError(f"{name} must be float64, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be on the HIP device")
    if t.shape != (n,) or (n > 1 and t.stride(0) != 1):
        raise ValueError(f"{name} must be a contiguous ({n},) tensor, got {tuple(t.shape)}")
    return t


def _ld(t: torch.Tensor) -> int:
    return t.stride(0) if t.shape[1] > 0 else 1


def _particles(pos, alpha, sigma, nu=None) -> L.Particles:
    _vec(pos, "positions")
    n = pos.shape[1]
    _vec(alpha, "strengths", n)
    if _ld(pos) != _ld(alpha):
        raise ValueError("positions and strengths must share the component stride")
    _scal(sigma, "radii", n)
    if nu is not None:
        _scal(nu, "viscosities", n)
    return L.Particles(n, _ld(pos), pos.data_ptr(), alpha.data_ptr(), sigma.data_ptr(),
                       nu.data_ptr() if nu is not None else None)


def _points(tgt) -> L.Points:
    _vec(tgt, "targets")
    return L.Points(tgt.shape[1], _ld(tgt), tgt.data_ptr())


def _out_vec(out, like, m):
    if out is None:
        return torch.empty((3, m), dtype=torch.float64, device=like.device)
    return _vec(out, "out", m)


# ---------------------------------------------------------------- workspaces
_ws = {}


def workspace(device: torch.device, nbytes: int) -> torch.Tensor:
    """Per-(device, stream) scratch buffer, grown on demand and reused.

    Calls on one stream are ordered, so reusing the buffer between them is
    safe; calls that overlap on different streams get different buffers.
    """
    key = (device, torch.cuda.current_stream(device).cuda_stream)
    buf = _ws.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        _ws[key] = buf
    return buf


def _variant(v):
    try:
        return VARIANTS[v]
    except KeyError:
        raise ValueError(f"variant must be one of {sorted(VARIANTS)}, got {v!r}") from None


# ------------------------------------------------------------ plain drop-in
def induced_velocity_at(targets, pos, alpha, sigma, out=None):
    """u at M targets (ParticlesPhysicsCuda.py:11-97).  Returns (3, M)."""
    p = _particles(pos, alpha, sigma)
    t = _points(targets)
    o = _out_vec(out, pos, targets.shape[1])
    L.call("vpm_get_induced_velocity_at", C.byref(t), C.byref(p), o.data_ptr(), _ld(o),
           _stream(pos))
    return o


def selfinduced_velocity(pos, alpha, sigma, variant="cuda", out=None):
    """u_i from all j != i (ParticlesPhysicsCuda.py:103-191).  Returns (3, N)."""
    p = _particles(pos, alpha, sigma)
    o = _out_vec(out, pos, pos.shape[1])
    L.call("vpm_get_selfinduced_velocity", C.byref(p), _variant(variant), o.data_ptr(), _ld(o),
           _stream(pos))
    return o


def vorticity_field_at(targets, pos, alpha, sigma, variant="cuda", out=None):
    """Smoothed vorticity at M targets (ParticlesPhysicsCuda.py:198-219,291-346)."""
    p = _particles(pos, alpha, sigma)
    t = _points(targets)
    o = _out_vec(out, pos, targets.shape[1])
    L.call("vpm_get_vorticity_field_at", C.byref(t), C.byref(p), _variant(variant), o.data_ptr(),
           _ld(o), _stream(pos))
    return o


def strength_gradients(pos, alpha, sigma, out=None):
    """Transpose-scheme stretching (ParticlesPhysicsCuda.py:351-445)."""
    p = _particles(pos, alpha, sigma)
    o = _out_vec(out, pos, pos.shape[1])
    L.call("vpm_get_strength_gradients", C.byref(p), o.data_ptr(), _ld(o), _stream(pos))
    return o


def strength_gradients_pse(pos, alpha, sigma, nu, variant="cuda", out=None):
    """PSE diffusion (ParticlesPhysicsCuda.py:452-530)."""
    p = _particles(pos, alpha, sigma, nu)
    o = _out_vec(out, pos, pos.shape[1])
    L.call("vpm_get_strength_gradients_pse", C.byref(p), _variant(variant), o.data_ptr(), _ld(o),
           _stream(pos))
    return o


def _reduction(fn, pos, alpha, sigma, variant):
    p = _particles(pos, alpha, sigma)
    res = torch.empty((1,), dtype=torch.float64, device=pos.device)
    nb = L.load().vpm_reduction_workspace_bytes(p.n)
    ws = workspace(pos.device, nb)
    L.call(fn, C.byref(p), _variant(variant), res.data_ptr(), ws.data_ptr(), ws.numel(),
           _stream(pos))
    return res[0]


def total_kinetic_energy(pos, alpha, sigma, variant="cuda"):
    """Kinetic energy, deterministic (ParticlesPhysicsCuda.py:537-642).  0-d tensor."""
    return _reduction("vpm_get_total_kinetic_energy", pos, alpha, sigma, variant)


def total_helicity(pos, alpha, sigma, variant="cuda"):
    """Helicity (ParticlesPhysicsCuda.py:647-750).  0-d tensor."""
    return _reduction("vpm_get_total_helicity", pos, alpha, sigma, variant)


def total_enstrophy(pos, alpha, sigma, variant="cuda"):
    """Enstrophy, i = j included (ParticlesPhysicsCuda.py:756-777,867-920).  0-d tensor."""
    return _reduction("vpm_get_total_enstrophy", pos, alpha, sigma, variant)


# ------------------------------------------------------------ fused forward
def scan_particles(pos, alpha, sigma, nu=None) -> int:
    """VPM_BAD_* bits of a device particle set (0: every position and
    strength finite, every radius finite and > 0, every viscosity finite and
    >= 0).  One scan kernel and a 4-byte read back (synchronises the stream)."""
    p = _particles(pos, alpha, sigma, nu)
    flags = torch.zeros(1, dtype=torch.int32, device=pos.device)
    L.call("vpm_scan_particles", C.byref(p), flags.data_ptr(), _stream(pos))
    return int(flags.item())


def check_particles(pos, alpha, sigma, nu=None, where="particle set"):
    """Raises NonFiniteError (status VPM_ERR_NONFINITE) when
    :func:`scan_particles` finds anything."""
    bits = scan_particles(pos, alpha, sigma, nu)
    if bits:
        raise L.NonFiniteError(where, bits)


def sigma_is_uniform(sigma: torch.Tensor) -> bool:
    """True when every radius equals sigma[0] (synchronises the stream)."""
    return sigma.numel() == 0 or bool((sigma == sigma[0]).all())


def forward(pos, alpha, sigma, nu=None, outputs=("vel", "stretch"), precision="f32c",
            sigma_mode="auto", presorted=False, tile_config=-1, max_splits=0, g64="auto",
            out=None):
    """One fused all-pairs pass (GPU-plugin semantics) returning a dict.

    ``outputs`` is any subset of {"vel", "stretch", "pse", "vort",
    "enstrophy"}; vectors come back as (3, N) tensors, "enstrophy" as a 0-d
    tensor.  ``precision`` is "f32c" (FP32 pairs, compensated FP64 totals)
    or "f64".  ``sigma_mode`` "uniform" asserts sigma_i == sigma_0 (not
    checked), "general" uses sigma_ij per pair, "auto" checks on the device.
    ``tile_config``: -1 (default shape), an index of :func:`tile_configs`, or
    "auto" (the autotuner's cached winner for this N).
    ``max_splits``: 0 lets the library split the sources among workgroups
    when the targets alone cannot fill the device (small N); k >= 1 caps
    the number of splits (1: never split).
    ``g64``: how the "f32c" pass sums the stretching's g alpha_j over source
    tiles near the targets -- "auto" (FP64 per pair in 64-lane waves of
    coherent target strengths, FP32 partials elsewhere), "always" or
    "never"; distant tiles of a uniform-radius pass with velocity take FP32
    partials in every mode.
    ``out`` may supply preallocated tensors by name.
    """
    if isinstance(outputs, str):
        outputs = (outputs,)
    mask = 0
    for k in outputs:
        if k not in OUTPUTS:
            raise ValueError(f"unknown output {k!r}; choose from {sorted(OUTPUTS)}")
        mask |= OUTPUTS[k]
    if not mask:
        raise ValueError("no outputs requested")
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    if (mask & L.OUT_PSE) and nu is None:
        raise ValueError("PSE output needs viscosities")
    p = _particles(pos, alpha, sigma, nu)
    n = p.n
    if sigma_mode == "auto":
        sigma_mode = "uniform" if sigma_is_uniform(sigma) else "general"
    if sigma_mode not in ("uniform", "general"):
        raise ValueError("sigma_mode must be 'auto', 'uniform' or 'general'")
    if tile_config == "auto":
        from . import autotune

        tile_config = autotune.best_config(n, outputs, precision, sigma_mode) \
            if precision == "f32c" else -1
    opts = L.ForwardOpts(mask, PRECISIONS[precision],
                         L.SIGMA_UNIFORM if sigma_mode == "uniform" else L.SIGMA_GENERAL,
                         1 if presorted else 0, int(tile_config), int(max_splits),
                         G64_MODES[g64])
    out = dict(out or {})
    res = {}
    ld = None
    for k in ("vel", "stretch", "pse", "vort"):
        if OUTPUTS[k] & mask:
            t = out.get(k)
            t = _out_vec(t, pos, n)
            if ld is None:
                ld = _ld(t)
            elif _ld(t) != ld:
                raise ValueError("all vector outputs must share one component stride")
            res[k] = t
    if mask & L.OUT_ENSTROPHY:
        e = out.get("enstrophy")
        if e is None:
            e = torch.empty((1,), dtype=torch.float64, device=pos.device)
        res["enstrophy"] = e
    fo = L.ForwardOut(ld if ld is not None else max(n, 1),
                      res["vel"].data_ptr() if "vel" in res else None,
                      res["stretch"].data_ptr() if "stretch" in res else None,
                      res["pse"].data_ptr() if "pse" in res else None,
                      res["vort"].data_ptr() if "vort" in res else None,
                      res["enstrophy"].data_ptr() if "enstrophy" in res else None)
    lib = L.load()
    nb = lib.vpm_forward_workspace_bytes(n, C.byref(opts))
    if nb == 0:
        raise L.VpmError("vpm_forward_workspace_bytes", -3, lib.vpm_last_error().decode())
    ws = workspace(pos.device, nb)
    L.call("vpm_forward", C.byref(p), C.byref(opts), C.byref(fo), ws.data_ptr(), ws.numel(),
           _stream(pos))
    if "enstrophy" in res:
        res["enstrophy"] = res["enstrophy"].reshape(())
    return res


def tile_config_default(outputs=("vel", "stretch")):
    """Index of the tile configuration used for ``tile_config=-1`` (uniform radius)."""
    if isinstance(outputs, str):
        outputs = (outputs,)
    mask = 0
    for k in outputs:
        mask |= OUTPUTS[k]
    return int(L.load().vpm_tile_config_default(mask))


def tile_configs():
    """List of (block, targets_per_lane, tile, subtile, comp) per config index."""
    lib = L.load()
    cfgs = []
    for i in range(lib.vpm_tile_config_count()):
        vals = [C.c_int() for _ in range(5)]
        L.call("vpm_tile_config_info", i, *[C.byref(v) for v in vals])
        cfgs.append(tuple(v.value for v in vals))
    return cfgs


# ------------------------------------------------------------ layout helpers
def to_soa(a, device="cuda") -> torch.Tensor:
    """(N,3) array-like -> (3,N) float64 device tensor."""
    t = torch.as_tensor(a, dtype=torch.float64)
    if t.dim() != 2 or t.shape[1] != 3:
        raise ValueError(f"expected shape (N,3), got {tuple(t.shape)}")
    return t.t().contiguous().to(device)


def from_soa(t: torch.Tensor):
    """(3,N) device tensor -> (N,3) contiguous float64 NumPy array."""
    return t.t().contiguous().cpu().numpy()
