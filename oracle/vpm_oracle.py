"""FP64 CPU oracle for the all-pairs regularised Biot-Savart path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP kernels are
compared against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``openonda_amd``) never does.

It restates, in float64 NumPy, the arithmetic of the two physics plugins of the
reference vortex-particle solver:

* ``variant="cuda"`` follows ``OpenONDA/solvers/VPM/ParticlesPhysicsCuda.py``
  (the GPU plugin; the parity target of the HIP drop-in),
* ``variant="cpu"`` follows ``OpenONDA/solvers/VPM/ParticlesPhysics.py``
  (the Numba CPU plugin), whose arithmetic differs in a few places.

Every function cites the reference lines it restates.  The kernel is the
Winckelmans-Leonard high-order algebraic kernel:
q(rho)/rho^3 = (rho^2 + 5/2)/(rho^2 + 1)^(5/2), zeta(rho) = 15/(8 pi) (rho^2+1)^(-7/2).

Pinning (see tests/test_oracle_pins.py): every function here is checked
against something other than its own formula -- quadrature of zeta, the
singular Biot-Savart limit, autograd identities (div u = 0, curl u = full
vorticity, stretching = grad(alpha_i . u_i)), PSE moment conditions on a
lattice, symmetry / antisymmetry invariants, and mpmath brute force at tiny N.

Evaluation is vectorised over blocks of targets so that N up to ~65k runs in
seconds to minutes.  Pure-Python double loops are used only by the tests at
tiny N.
"""
from __future__ import annotations

import math

import numpy as np

# Reference constants: ParticlesPhysicsCuda.py:5-6, ParticlesPhysics.py:4,
# vpmModule.py:22 (epsilon of the LES filter), vpmModule.py:92 (Cnu).
EPS_PAIR = 1e-8
EPS_LES = 1e-9
ONE_OVER_FOUR_PI = 1.0 / (4.0 * np.pi)
CNU_INIT = 0.02958
CNU_MAX = 0.0761
CORE_SPREADING_C = 5.34

VARIANTS = ("cuda", "cpu")


def _check_variant(variant):
    if variant not in VARIANTS:
        raise ValueError(f"variant must be one of {VARIANTS}, got {variant!r}")


def _f64(a, shape_tail=None, name="array"):
    a = np.asarray(a, dtype=np.float64)
    if shape_tail is not None and a.shape[1:] != shape_tail:
        raise ValueError(f"{name} must have shape (n,{shape_tail}), got {a.shape}")
    return a


def _rows(rows, n):
    """Target rows of a self-evaluation: all n, or the given indices."""
    return np.arange(n) if rows is None else np.asarray(rows, dtype=np.int64).reshape(-1)


def _chunks(m, n, budget=4_000_000):
    step = max(1, budget // max(n, 1))
    for s in range(0, m, step):
        yield s, min(m, s + step)


# ---------------------------------------------------------------------------
# Velocity
# ---------------------------------------------------------------------------
def induced_velocity_at(targets, x, alpha, sigma):
    """u(y) at M targets from N sources, sigma = sigma_j, skip r^2 <= 1e-8.

    ParticlesPhysicsCuda.py:39-97 (kernel) and ParticlesPhysics.py:9-74 (same
    arithmetic in both plugins).
    """
    y = _f64(targets, (3,), "targets")
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    out = np.zeros((y.shape[0], 3))
    s2 = s * s
    s3 = s * s2
    for b, e in _chunks(y.shape[0], x.shape[0]):
        r = y[b:e, None, :] - x[None, :, :]                       # PPC:74-76
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        rho2 = r2 / s2[None, :]                                   # PPC:83
        q = (rho2 + 2.5) / ((rho2 + 1.0) ** 2.5 * s3[None, :])    # PPC:84
        q = np.where(r2 > EPS_PAIR, q, 0.0)                       # PPC:80
        c = np.cross(r, a[None, :, :])                            # PPC:86-88
        out[b:e] = -np.einsum("mn,mnk->mk", q, c)                 # PPC:90-92
    return out * ONE_OVER_FOUR_PI                                  # PPC:95-97


def selfinduced_velocity(x, alpha, sigma, variant="cuda", rows=None):
    """u_i = -1/(4pi) sum_{j!=i} g(rho^2)/sigma^3 (r_ij x alpha_j).

    cuda: sigma = (sigma_i + sigma_j)/2, ParticlesPhysicsCuda.py:133-191.
    cpu : sigma = sigma_i (target radius), ParticlesPhysics.py:81-143 (:115).
    ``rows``: evaluate only these targets i (default: all; output in that order).
    """
    _check_variant(variant)
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    n = x.shape[0]
    R = _rows(rows, n)
    out = np.zeros((len(R), 3))
    for b, e in _chunks(len(R), n):
        ii = R[b:e]
        r = x[ii, None, :] - x[None, :, :]
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        if variant == "cuda":
            sg = (s[ii, None] + s[None, :]) / 2.0                  # PPC:166
        else:
            sg = np.broadcast_to(s[ii, None], r2.shape)           # PPH:115
        sg2 = sg * sg
        rho2 = r2 / sg2
        q = (rho2 + 2.5) / ((rho2 + 1.0) ** 2.5 * sg * sg2)       # PPC:177
        q[np.arange(e - b), ii] = 0.0                              # j != i, PPC:160
        c = np.cross(r, a[None, :, :])
        out[b:e] = -np.einsum("mn,mnk->mk", q, c)
    return out * ONE_OVER_FOUR_PI


# ---------------------------------------------------------------------------
# Vorticity (LES filter input)
# ---------------------------------------------------------------------------
def vorticity_field_at(targets, x, alpha, sigma, variant="cuda"):
    """Smoothed vorticity at M targets, sigma = sigma_j, skip r^2 <= 1e-8.

    cuda: simplified form  omega = 1/(4pi) sum 7.5/(sigma^3 A^3.5) alpha_j,
          ParticlesPhysicsCuda.py:291-346 (launched by :198-219).  The
          reference wrapper sizes the grid and the output by N sources
          (:199-200, :207); the restatement returns (M,3), which coincides with
          the reference whenever M == N (the only way vpmModule.py calls it).
    cpu : full curl form  (xi-q) alpha + (3q-xi)(r.alpha/r^2) r,
          ParticlesPhysics.py:211-262.
    """
    _check_variant(variant)
    y = _f64(targets, (3,), "targets")
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    out = np.zeros((y.shape[0], 3))
    s3 = s ** 3
    for b, e in _chunks(y.shape[0], x.shape[0]):
        r = y[b:e, None, :] - x[None, :, :]
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        mask = r2 > EPS_PAIR
        rho2 = r2 / (s * s)[None, :]
        xi = 7.5 / ((rho2 + 1.0) ** 3.5 * s3[None, :])             # PPC:337
        if variant == "cuda":
            w = np.where(mask, xi, 0.0)
            out[b:e] = np.einsum("mn,nk->mk", w, a)                 # PPC:339-341
        else:
            q = (rho2 + 2.5) / ((rho2 + 1.0) ** 2.5 * s3[None, :])  # PPH:253
            t1 = np.where(mask, xi - q, 0.0)                         # PPH:255
            ra = np.einsum("mnk,nk->mn", r, a)
            t2 = np.where(mask, (3.0 * q - xi) * ra / np.where(mask, r2, 1.0), 0.0)  # PPH:256
            out[b:e] = np.einsum("mn,nk->mk", t1, a) + np.einsum("mn,mnk->mk", t2, r)  # PPH:258
    return out * ONE_OVER_FOUR_PI


# ---------------------------------------------------------------------------
# Stretching (transpose scheme)
# ---------------------------------------------------------------------------
def strength_gradients(x, alpha, sigma, rows=None):
    """dalpha_i/dt = 1/(4pi) sum_{j!=i} [f1 (a_i x a_j) + f2 (a_i.(r x a_j)) r].

    f1 = (A+1.5)/(s^3 A^2.5), f2 = 3(A+2.5)/(s^5 A^3.5), A = r^2/s^2 + 1,
    s = (s_i+s_j)/2.  ParticlesPhysicsCuda.py:375-445 and
    ParticlesPhysics.py:268-312 (same arithmetic).  ``rows`` as in
    :func:`selfinduced_velocity`.
    """
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    n = x.shape[0]
    R = _rows(rows, n)
    out = np.zeros((len(R), 3))
    for b, e in _chunks(len(R), n):
        ii = R[b:e]
        r = x[ii, None, :] - x[None, :, :]                         # PPC:403-405
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        sg = (s[ii, None] + s[None, :]) * 0.5                       # PPC:409
        sg2 = sg * sg
        sg3 = sg2 * sg
        A = r2 / sg2 + 1.0                                          # PPC:413
        f1 = (A + 1.5) / (sg3 * A ** 2.5)                           # PPC:416
        f2 = 3.0 * (A + 2.5) / (sg3 * sg2 * A ** 3.5)               # PPC:417
        f1[np.arange(e - b), ii] = 0.0
        f2[np.arange(e - b), ii] = 0.0
        ai = a[ii, None, :]
        p1 = np.cross(np.broadcast_to(ai, r.shape), a[None, :, :])  # PPC:420-422
        p2 = np.cross(r, a[None, :, :])                              # PPC:425-427
        p3 = np.einsum("mnk,mnk->mn", np.broadcast_to(ai, r.shape), p2)  # PPC:430-434
        out[b:e] = np.einsum("mn,mnk->mk", f1, p1) + np.einsum("mn,mnk->mk", f2 * p3, r)
    return out * ONE_OVER_FOUR_PI                                   # PPC:438-440


# ---------------------------------------------------------------------------
# PSE viscous diffusion
# ---------------------------------------------------------------------------
def strength_gradients_pse(x, alpha, sigma, nu, variant="cuda", rows=None):
    """PSE: dalpha_i/dt = 1/(4pi) sum_{j!=i} 105 nu s^3/(s^5 A^4.5) (a_j - a_i).

    cuda: nu = nu_i (target), A = r^2/s^2 + 1, ParticlesPhysicsCuda.py:478-530.
    cpu : nu = (nu_i+nu_j)/2, A = r^2/s^2 + 1 + 1e-8, ParticlesPhysics.py:318-382.
    ``rows`` as in :func:`selfinduced_velocity`.
    """
    _check_variant(variant)
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    v = _f64(nu)
    n = x.shape[0]
    R = _rows(rows, n)
    out = np.zeros((len(R), 3))
    for b, e in _chunks(len(R), n):
        ii = R[b:e]
        r = x[ii, None, :] - x[None, :, :]
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        sg = (s[ii, None] + s[None, :]) * 0.5
        sg2 = sg * sg
        sg3 = sg2 * sg
        sg5 = sg2 * sg3
        if variant == "cuda":
            A = r2 / sg2 + 1.0                                      # PPC:518
            vis = np.broadcast_to(v[ii, None], r2.shape)            # PPC:520 (nu_i)
        else:
            A = r2 / sg2 + 1.0 + EPS_PAIR                           # PPH:369
            vis = (v[ii, None] + v[None, :]) * 0.5                  # PPH:360
        f3 = 105.0 * vis / (sg5 * A ** 4.5)                         # PPC:520
        w = f3 * sg3
        w[np.arange(e - b), ii] = 0.0
        da = a[None, :, :] - a[ii, None, :]                         # PPC:523-525
        out[b:e] = np.einsum("mn,mnk->mk", w, da)
    return out * ONE_OVER_FOUR_PI


# ---------------------------------------------------------------------------
# Scalar diagnostics
# ---------------------------------------------------------------------------
def total_kinetic_energy(x, alpha, sigma, variant="cuda"):
    """E = 1/(16pi) sum_{i!=j} [2rho/A^.5 (ai.aj) + rho^3/A^1.5 ((ai.e)(aj.e) - ai.aj)]/|r|.

    cuda: A = rho^2 + 1, ParticlesPhysicsCuda.py:563-642 (wrapper :537-558).
          The reference accumulates into an uninitialised device scalar
          (cuda.device_array, :545); the restatement starts from 0.
    cpu : A = rho^2 + 1 + 1e-8, ParticlesPhysics.py:388-429.
    """
    _check_variant(variant)
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    n = x.shape[0]
    tot = 0.0
    for b, e in _chunks(n, n):
        r = x[b:e, None, :] - x[None, :, :]
        rn = np.sqrt(r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2])
        idx = np.arange(b, e)
        rn_safe = rn.copy()
        rn_safe[idx - b, idx] = 1.0
        eu = r / rn_safe[..., None]                                  # PPC:598-600
        sg = (s[b:e, None] + s[None, :]) * 0.5
        rho = rn_safe / sg
        A = rho ** 2 + 1.0 + (EPS_PAIR if variant == "cpu" else 0.0)  # PPC:604 / PPH:418
        t1 = 2.0 * rho / A ** 0.5                                    # PPC:606
        t2 = rho ** 3 / A ** 1.5                                     # PPC:607
        ai = a[b:e]
        p1 = ai @ a.T                                                # PPC:610-614
        p2 = np.einsum("mnk,mk->mn", eu, ai)                         # PPC:617-621
        p3 = np.einsum("mnk,nk->mn", eu, a)                          # PPC:624-628
        term = (t1 * p1 + t2 * (p2 * p3 - p1)) / rn_safe             # PPC:631
        term[idx - b, idx] = 0.0
        tot += term.sum()
    return tot / (16.0 * np.pi)                                      # PPC:642 / PPH:429


def total_helicity(x, alpha, sigma, variant="cuda"):
    """H = k sum_{i!=j} (rho^2+2.5)/(s^3 A^2.5) r.(a_i x a_j).

    cuda: k = 1/(4pi) in the kernel (ParticlesPhysicsCuda.py:750) and again in
          the wrapper (:668), so 1/(16 pi^2) overall.
    cpu : k = 1/(4pi), ParticlesPhysics.py:432-468.
    """
    _check_variant(variant)
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    n = x.shape[0]
    tot = 0.0
    for b, e in _chunks(n, n):
        r = x[b:e, None, :] - x[None, :, :]
        sg = (s[b:e, None] + s[None, :]) * 0.5
        if variant == "cuda":
            rho = r / sg[..., None]                                   # PPC:721-723
            rho2 = (rho * rho).sum(-1)
        else:
            rm = np.sqrt((r * r).sum(-1))                             # PPH:455
            rho2 = (rm / sg) ** 2                                      # PPH:458-459
        term = (rho2 + 2.5) / (sg ** 3 * (rho2 + 1.0) ** 2.5)         # PPC:726
        ai = np.broadcast_to(a[b:e, None, :], r.shape)
        p1 = np.cross(ai, a[None, :, :])                               # PPC:728-730
        p2 = (r * p1).sum(-1)                                          # PPC:733-737
        idx = np.arange(b, e)
        term[idx - b, idx] = 0.0
        tot += (p2 * term).sum()
    if variant == "cuda":
        return tot / (4.0 * np.pi) / (4.0 * np.pi)
    return tot / (4.0 * np.pi)


def total_enstrophy(x, alpha, sigma, variant="cuda"):
    """Ens = k sum_i sum_j (r^2 + 3.2 s^2)/(r^2 + s^2)^1.5 (a_i.a_j), i = j included.

    cuda: k = 1/(8pi) in the kernel (ParticlesPhysicsCuda.py:920) and 1/(4pi)
          in the wrapper (:777).  The accumulator is uninitialised in the
          reference (:764); the restatement starts from 0.
    cpu : k = 1/(8pi), ParticlesPhysics.py:517-550.
    """
    _check_variant(variant)
    x = _f64(x, (3,), "positions")
    a = _f64(alpha, (3,), "strengths")
    s = _f64(sigma)
    n = x.shape[0]
    tot = 0.0
    for b, e in _chunks(n, n):
        r = x[b:e, None, :] - x[None, :, :]
        r2 = r[..., 0] * r[..., 0] + r[..., 1] * r[..., 1] + r[..., 2] * r[..., 2]
        sg = (s[b:e, None] + s[None, :]) * 0.5
        sg2 = sg * sg
        t1 = (r2 + 3.2 * sg2) / (r2 + sg2) ** 1.5                     # PPC:900
        p1 = a[b:e] @ a.T                                              # PPC:903-907
        tot += (t1 * p1).sum()
    if variant == "cuda":
        return tot / (8.0 * np.pi) / (4.0 * np.pi)
    return tot / (8.0 * np.pi)


# ---------------------------------------------------------------------------
# O(N) diagnostics and the LES filter (vpmModule.py)
# ---------------------------------------------------------------------------
def total_linear_impulse(x, alpha):
    """I = 1/2 sum x x alpha, vpmModule.py:264-267."""
    return 0.5 * np.sum(np.cross(_f64(x), _f64(alpha)), axis=0)


def total_angular_impulse(x, alpha):
    """A = 1/3 sum x x (x x alpha), vpmModule.py:270-274."""
    x = _f64(x)
    return (1.0 / 3.0) * np.sum(np.cross(x, np.cross(x, _f64(alpha))), axis=0)


def les_filter_constant(enstrophy, cnu, enstrophy_ref=None):
    """Dynamic C_nu update, vpmModule.py:887-909.  Returns (cnu, enstrophy_ref)."""
    e = math.sqrt(enstrophy)                                           # :894
    if enstrophy_ref is None:                                          # :897-898
        enstrophy_ref = e
    enstrophy_ref = 0.1 * e + 0.9 * enstrophy_ref                      # :901
    cnu = cnu * e / (enstrophy_ref + EPS_LES)                          # :904
    cnu = max(EPS_LES, min(cnu, CNU_MAX))                              # :907
    return cnu, enstrophy_ref


def les_particle_viscosity(vorticity, radii, cnu):
    """Sub-grid viscosity nu_t per particle, vpmModule.py:916-962.

    Returns (nu_t, write) where ``write`` tells whether the reference would
    store nu_t (only if max nu_t > 1e-9, :954-960).
    """
    w = _f64(vorticity)
    mag = np.sqrt(np.sum(w * w, axis=1))                               # :935
    unit = w / mag[:, None]                                            # :936
    avg = np.mean(w, axis=0)                                           # :939
    avg_unit = avg / np.sqrt(np.sum(avg * avg))                        # :940-941
    cosb = unit @ avg_unit                                             # :944
    beta = np.arccos(np.clip(cosb, -1.0, 1.0))                         # :945
    b0 = np.pi / 4.0                                                   # :932
    psi = np.where((b0 <= beta) & (beta <= np.pi - b0), 1.0, 0.0)      # :948
    nut = psi * (cnu * _f64(radii)) ** 2 * mag                         # :951
    return nut, bool(np.max(nut) > EPS_LES)                            # :954-956


def core_spreading(radii, nu_eff, dt):
    """sigma <- sqrt(sigma^2 + 5.34 nu_eff dt), vpmModule.py:851-862."""
    return np.sqrt(_f64(radii) ** 2 + CORE_SPREADING_C * _f64(nu_eff) * dt)
