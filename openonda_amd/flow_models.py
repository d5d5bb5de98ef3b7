"""Initial conditions: the benchmark particle layouts and the reference's
vortex-particle initial-condition generators (host NumPy, float64).

Layouts used by the benchmark workloads:
 * :func:`torus_ring`  -- a thin ring sampled on a torus: ``n_stations``
   azimuthal stations x 64 cross-section points on rings rho = 0, h, 2h, 3h
   (1/10/21/32 points, half-step angular offset on alternate rings).
 * :func:`leapfrog_rings` -- two coaxial torus rings.
 * :func:`random_cloud` -- positions U[0, 2 pi]^3, strengths N(0, 1/N).
 * :func:`batch_instances` -- uneven ring/cloud instances for batching.

Reference generators restated (``OpenONDA/utilities/vpm_flow_models.py``):
:func:`vortex_ring_vpm` (:86-159), :func:`lamb_oseen_vpm` (:13-82), and the
box lattice of ``OpenONDA/utilities/vpm_solver_helper.py:21-68``.
"""
from __future__ import annotations

import math

import numpy as np

from .vpm_flow_models import lamb_oseen_vpm, vortex_ring_vpm  # noqa: F401  (re-exported)

CROSS_SECTION_COUNTS = (1, 10, 21, 32)


def cross_section(h: float):
    """64 points (rho, psi, area) on rings rho = 0, h, 2h, 3h of a disk section."""
    pts = []
    for m, cnt in enumerate(CROSS_SECTION_COUNTS):
        rho = m * h
        area = math.pi * (0.5 * h) ** 2 if m == 0 else 2.0 * math.pi * m * h * h
        off = 0.5 if m % 2 == 1 else 0.0           # half-step offset on alternate rings
        for q in range(cnt):
            psi = 2.0 * math.pi * (q + off) / cnt
            pts.append((rho, psi, area / cnt))
    return np.array(pts)


def torus_ring(n_stations=64, R=1.0, sigma_c=0.1, gamma=1.0, h=0.1, centre=(0.0, 0.0, 0.0),
               smoothing=None, nu=0.0, renormalise=False):
    """Ring of circulation ``gamma`` with axis x in the y-z plane.

    Vorticity omega = gamma/(pi sigma_c^2) exp(-rho^2/sigma_c^2) e_theta, the
    profile form of vortex_ring_vpm (vpm_flow_models.py:136) with e_theta =
    (0, -sin theta, cos theta) (:140-142), theta = atan2(z, y) (:127).
    Strength alpha = omega * vol with vol = (2 pi (R + rho cos psi)/n_st) *
    (annulus area / count).  ``smoothing`` defaults to 1.3 h.

    Returns (x (N,3), alpha (N,3), sigma (N,), nu (N,)).
    """
    cs = cross_section(h)
    th = 2.0 * np.pi * np.arange(n_stations) / n_stations
    rho, psi, area = cs[:, 0], cs[:, 1], cs[:, 2]
    rr = np.broadcast_to(R + rho * np.cos(psi), (n_stations, len(cs)))  # distance to the axis
    xa = np.broadcast_to(rho * np.sin(psi), rr.shape)
    y = rr * np.cos(th)[:, None]
    z = rr * np.sin(th)[:, None]
    x = np.stack([xa + centre[0], y + centre[1], z + centre[2]], -1).reshape(-1, 3)
    w = gamma / (math.pi * sigma_c ** 2) * np.exp(-(rho ** 2) / sigma_c ** 2)
    vol = (2.0 * np.pi * rr / n_stations) * area[None, :]
    mag = (w[None, :] * vol).reshape(-1)
    thp = np.repeat(th, len(cs))
    alpha = np.stack([np.zeros_like(mag), -mag * np.sin(thp), mag * np.cos(thp)], -1)
    if renormalise:                                          # vpm_flow_models.py:150-152
        g = np.sum(np.sqrt(np.sum(alpha ** 2, 1)))
        alpha *= abs(2 * np.pi * R * gamma / g)
    s = np.full(len(x), 1.3 * h if smoothing is None else smoothing)
    return x, alpha, s, np.full(len(x), float(nu))


def leapfrog_rings(n_stations=512, R=1.0, separation=0.5, sigma_c=0.1, h=0.1, smoothing=0.13,
                   nu=1e-3, jitter=0.01, seed=2):
    """Two coaxial rings (axis x) at x = -/+ separation/2, jittered U(+-jitter h)."""
    parts = [torus_ring(n_stations, R, sigma_c, 1.0, h, (c, 0.0, 0.0), smoothing, nu)
             for c in (-0.5 * separation, 0.5 * separation)]
    x = np.concatenate([p[0] for p in parts])
    a = np.concatenate([p[1] for p in parts])
    s = np.concatenate([p[2] for p in parts])
    v = np.concatenate([p[3] for p in parts])
    rng = np.random.default_rng(seed)
    x = x + rng.uniform(-jitter * h, jitter * h, x.shape)
    return x, a, s, v


def random_cloud(n, seed=None, box=2.0 * np.pi):
    """Positions U[0, box]^3, strengths N(0, 1/n) per component, sigma = 1.3 box/n^(1/3).

    Default seed: 3 + log2(n).
    """
    if seed is None:
        seed = 3 + int(round(math.log2(n)))
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, box, (n, 3))
    a = rng.normal(0.0, math.sqrt(1.0 / n), (n, 3))
    s = np.full(n, 1.3 * box / n ** (1.0 / 3.0))
    return x, a, s, np.zeros(n)


def ring_probes(m=4096, seed=4):
    """Evaluation points U([-0.5,0.5] x [-1.5,1.5]^2)."""
    rng = np.random.default_rng(seed)
    lo = np.array([-0.5, -1.5, -1.5])
    hi = np.array([0.5, 1.5, 1.5])
    return lo + (hi - lo) * rng.uniform(0.0, 1.0, (m, 3))


def batch_instances(count=96, nmin=2048, nmax=131072, seed=5):
    """Uneven instances: N_i log-uniform in [nmin, nmax]; even i rings, odd i clouds.

    Each instance has its own uniform smoothing radius 1.3 h_i.  Returns a
    list of (x, alpha, sigma, nu).
    """
    rng = np.random.default_rng(seed)
    sizes = np.round(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(int)
    out = []
    for i, n in enumerate(sizes):
        if i % 2 == 0:
            st = max(1, int(round(n / 64)))
            h = 2.0 * np.pi / st if st >= 64 else 0.1
            out.append(torus_ring(st, h=min(h, 0.1)))
        else:
            out.append(random_cloud(int(n), seed=1000 + i))
    return out


def rectangular_point_distribution(box_limits, spacing):
    """Box lattice with equal volumes (vpm_solver_helper.py:21-68)."""
    b = box_limits
    num = [int(np.ceil((b[2 * k + 1] - b[2 * k]) / spacing)) + 1 for k in range(3)]
    num = [v + 1 if v % 2 == 0 else v for v in num]
    axes = [np.linspace(b[2 * k], b[2 * k + 1], num=num[k]) for k in range(3)]
    xx, yy, zz = np.meshgrid(*axes, indexing="ij")
    pts = np.vstack([xx.ravel(), yy.ravel(), zz.ravel()]).T
    vol = (b[1] - b[0]) * (b[3] - b[2]) * (b[5] - b[4]) / pts.shape[0]
    return pts, np.full(pts.shape[0], vol)
