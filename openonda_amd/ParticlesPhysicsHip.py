"""Drop-in physics module for the reference ParticleSystem (HIP, MI355X).

Same eight functions, argument order, host float64 NumPy in / fresh NumPy out
semantics as ``OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py`` (the module the
reference loads with ``importlib`` at ``vpmModule.py:141-154``), computed by
the plain FP64 HIP kernels of libvpm_hip.so.  Each call copies its inputs to
the device, runs one kernel and copies the result back, like the reference
wrappers (``ParticlesPhysicsCuda.py:16-34``).

Differences from the reference, all deliberate:
 * ``get_vorticity_field_at`` returns one row per *target* (the reference
   sizes grid and output by the source count, ``ParticlesPhysicsCuda.py:199-207``;
   both agree whenever M == N, the only way ``vpmModule.py`` calls it).
 * The scalar reductions start from zero and sum in a fixed order (the
   reference accumulates with FP64 atomics into an uninitialised device
   scalar, ``ParticlesPhysicsCuda.py:545,642``).
 * Inputs that are not float64 raise ``TypeError`` and bad shapes raise
   ``ValueError`` (the reference's explicit Numba signatures reject the
   former, ``ParticlesPhysicsCuda.py:39``).
 * NaN or Inf in any input, a radius <= 0 or a viscosity < 0 raises
   ``NonFiniteError`` (a ``ValueError``, status ``VPM_ERR_NONFINITE``) before
   any GPU work; the reference returns NaN.  ``CHECK_FINITE = False`` skips
   the check.

``VARIANT`` selects whose arithmetic is reproduced: "cuda" (this module's
default, the GPU plugin) or "cpu" (``ParticlesPhysics.py``, the Numba CPU
plugin), both evaluated on the GPU.

``PRECISION`` selects the kernels of the four sums the fused pass covers
(self-induced velocity, stretching, PSE, enstrophy): "plain" (the default:
the reference arithmetic restated, one kernel per call), "f64" (the fused
FP64 pass: one rsqrt per pair instead of ``pow``, rel-L2 <= 1e-12 against the
plain kernels, about 11x faster at N = 2^20) or "f32c" (the fused
FP32-compensated pass, rel-L2 <= 1e-6, about 25x).  The fused pass computes
the GPU plugin's arithmetic, so "f64" and "f32c" require ``VARIANT ==
"cuda"``; velocity at targets, vorticity, energy and helicity always use the
plain kernels.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from . import physics as P

VARIANT = "cuda"
PRECISION = "plain"
# Inputs are checked on the host before the upload: NaN/Inf in any array, a
# radius <= 0 or a viscosity < 0 raises NonFiniteError (status
# VPM_ERR_NONFINITE) instead of returning NaN as the reference does.
CHECK_FINITE = True


def _arr(a, name, cols=3, bad=0):
    """Validates one host array; ``bad`` is its VPM_BAD_* bit for the
    finiteness check (0: none)."""
    if not isinstance(a, np.ndarray):
        a = np.asarray(a)
    if a.dtype != np.float64:
        raise TypeError(f"{name} must be a float64 array, got {a.dtype}")
    if cols is None:
        if a.ndim != 1:
            raise ValueError(f"{name} must have shape (N,), got {a.shape}")
    elif a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"{name} must have shape (N, {cols}), got {a.shape}")
    if CHECK_FINITE and bad:
        ok = np.isfinite(a).all()
        if ok and bad == L.BAD_RADIUS:
            ok = bool((a > 0).all())
        elif ok and bad == L.BAD_VISCOSITY:
            ok = bool((a >= 0).all())
        if not ok:
            raise L.NonFiniteError(name, bad)
    return a


def _dev():
    P.require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def _soa(a, dev):
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(dev)


def _src(x, a, s, nu=None, targets=None):
    """Validates every array on the host (before any GPU call), then uploads:
    returns ([targets,] x, a, s[, nu]) as device tensors."""
    y = _arr(targets, "target_positions", bad=L.BAD_POSITION) if targets is not None else None
    x = _arr(x, "particles_positions", bad=L.BAD_POSITION)
    a = _arr(a, "particles_strengths", bad=L.BAD_STRENGTH)
    s = _arr(s, "particles_radii", None, bad=L.BAD_RADIUS)
    n = x.shape[0]
    if a.shape[0] != n or s.shape[0] != n:
        raise ValueError("positions, strengths and radii must have the same length")
    if nu is not None:
        nu = _arr(nu, "viscosities", None, bad=L.BAD_VISCOSITY)
        if nu.shape[0] != n:
            raise ValueError("viscosities must have the same length as positions")
    dev = _dev()
    out = [_soa(x, dev), _soa(a, dev), torch.from_numpy(np.ascontiguousarray(s)).to(dev)]
    if nu is not None:
        out.append(torch.from_numpy(np.ascontiguousarray(nu)).to(dev))
    if y is not None:
        out.insert(0, _soa(y, dev))
    return out


def _host(t):
    return P.from_soa(t)


def _fused():
    """The fused-pass precision in use, or None for the plain kernels."""
    if PRECISION == "plain":
        return None
    if PRECISION not in ("f64", "f32c"):
        raise ValueError(f"PRECISION must be 'plain', 'f64' or 'f32c', got {PRECISION!r}")
    if VARIANT != "cuda":
        raise ValueError("PRECISION 'f64' / 'f32c' computes the GPU plugin's arithmetic: "
                         "it needs VARIANT == 'cuda'")
    return PRECISION


def get_induced_velocity_at(target_positions, particles_positions, particles_strengths,
                            particles_radii):
    """Velocity at targets, (M,3).  ParticlesPhysicsCuda.py:11-34."""
    y, x, a, s = _src(particles_positions, particles_strengths, particles_radii,
                      targets=target_positions)
    return _host(P.induced_velocity_at(y, x, a, s))


def get_selfinduced_velocity(particles_positions, particles_strengths, particles_radii):
    """Self-induced velocity, (N,3).  ParticlesPhysicsCuda.py:103-127."""
    x, a, s = _src(particles_positions, particles_strengths, particles_radii)
    prec = _fused()
    if prec:
        return _host(P.forward(x, a, s, outputs=("vel",), precision=prec)["vel"])
    return _host(P.selfinduced_velocity(x, a, s, VARIANT))


def get_vorticity_field_at(target_positions, particle_positions, particle_strengths,
                           particle_radii):
    """Smoothed vorticity at targets, (M,3).  ParticlesPhysicsCuda.py:198-219."""
    y, x, a, s = _src(particle_positions, particle_strengths, particle_radii,
                      targets=target_positions)
    return _host(P.vorticity_field_at(y, x, a, s, VARIANT))


def get_strength_gradients(particle_positions, particle_strengths, particle_radii):
    """Stretching dalpha/dt, (N,3).  ParticlesPhysicsCuda.py:351-370."""
    x, a, s = _src(particle_positions, particle_strengths, particle_radii)
    prec = _fused()
    if prec:
        return _host(P.forward(x, a, s, outputs=("stretch",), precision=prec)["stretch"])
    return _host(P.strength_gradients(x, a, s))


def get_strength_gradients_PSE(positions, strengths, radii, viscosities):
    """PSE dalpha/dt, (N,3).  ParticlesPhysicsCuda.py:452-473."""
    x, a, s, nu = _src(positions, strengths, radii, viscosities)
    prec = _fused()
    if prec:
        return _host(P.forward(x, a, s, nu, outputs=("pse",), precision=prec)["pse"])
    return _host(P.strength_gradients_pse(x, a, s, nu, VARIANT))


def get_total_kinetic_energy(particles_positions, particles_strengths, particles_radii):
    """Kinetic energy (float).  ParticlesPhysicsCuda.py:537-558."""
    x, a, s = _src(particles_positions, particles_strengths, particles_radii)
    return float(P.total_kinetic_energy(x, a, s, VARIANT).item())


def get_total_helicity(particles_positions, particles_strengths, particles_radii):
    """Helicity (float).  ParticlesPhysicsCuda.py:647-668."""
    x, a, s = _src(particles_positions, particles_strengths, particles_radii)
    return float(P.total_helicity(x, a, s, VARIANT).item())


def get_total_enstrophy(particles_positions, particles_strengths, particles_radii):
    """Enstrophy (float).  ParticlesPhysicsCuda.py:756-777."""
    x, a, s = _src(particles_positions, particles_strengths, particles_radii)
    prec = _fused()
    if prec:
        return float(P.forward(x, a, s, outputs=("enstrophy",), precision=prec)["enstrophy"].item())
    return float(P.total_enstrophy(x, a, s, VARIANT).item())
