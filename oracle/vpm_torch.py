"""Differentiable FP64 restatement of the pair sums in PyTorch (test infrastructure).

TEST INFRASTRUCTURE ONLY: the reference gradient for the hand-written HIP
adjoint.  The pair arithmetic follows the GPU plugin
(ParticlesPhysicsCuda.py:159-191 velocity, :396-440 stretching, :500-525 PSE,
:65-97 velocity at targets) exactly as oracle/vpm_oracle.py does; its values
are checked against that module in tests/test_oracle_torch.py, and its
gradients come from torch.autograd.  Inputs are (N,3) float64 tensors on any
device; targets are processed in chunks under activation checkpointing so
N = 4096 fits in memory.
"""
from __future__ import annotations

import math

import torch
from torch.utils.checkpoint import checkpoint

C = 1.0 / (4.0 * math.pi)
EPS = 1e-8


def _chunk_size(n, budget=2_000_000):
    return max(1, budget // max(n, 1))


def _pairs(xi, ai, si, nui, idx, x, a, s, want):
    """Outputs for a block of targets (rows) against all sources."""
    r = xi[:, None, :] - x[None, :, :]
    r2 = (r * r).sum(-1)
    sg = (si[:, None] + s[None, :]) * 0.5
    sg2 = sg * sg
    A = r2 / sg2 + 1.0
    self_mask = torch.zeros_like(r2, dtype=torch.bool)
    self_mask[torch.arange(len(idx), device=x.device), idx] = True
    out = []
    if "vel" in want or "stretch" in want:
        g = (A + 1.5) / (sg2 * sg * A ** 2.5)
        g = torch.where(self_mask, torch.zeros_like(g), g)
        c = torch.cross(r, a[None, :, :].expand_as(r), dim=-1)
    if "vel" in want:
        out.append(-C * (g[..., None] * c).sum(1))
    if "stretch" in want:
        f = 3.0 * (A + 2.5) / (sg2 * sg2 * sg * A ** 3.5)
        f = torch.where(self_mask, torch.zeros_like(f), f)
        p1 = torch.cross(ai[:, None, :].expand_as(r), a[None, :, :].expand_as(r), dim=-1)
        p3 = (ai[:, None, :] * c).sum(-1)
        out.append(C * ((g[..., None] * p1).sum(1) + ((f * p3)[..., None] * r).sum(1)))
    if "pse" in want:
        w = 105.0 * nui[:, None] / (sg2 * A ** 4.5)
        w = torch.where(self_mask, torch.zeros_like(w), w)
        out.append(C * (w[..., None] * (a[None, :, :] - ai[:, None, :])).sum(1))
    return tuple(out)


def pair_sums(x, a, s, nu=None, want=("vel", "stretch", "pse")):
    """Self-evaluation of the requested outputs, each (N,3)."""
    n = x.shape[0]
    if nu is None:
        nu = torch.zeros_like(s)
    cs = _chunk_size(n)
    outs = [[] for _ in want]
    for b in range(0, n, cs):
        e = min(n, b + cs)
        idx = torch.arange(b, e, device=x.device)
        # idx is bound now: the checkpoint recomputes fn during backward, after
        # the loop has moved on
        fn = lambda xi, ai, si, nui, x_, a_, idx=idx: _pairs(xi, ai, si, nui, idx, x_, a_, s, want)
        res = checkpoint(fn, x[b:e], a[b:e], s[b:e], nu[b:e], x, a, use_reentrant=False)
        for k, t in enumerate(res):
            outs[k].append(t)
    return tuple(torch.cat(o) for o in outs)


def induced_velocity_at(y, x, a, s):
    """u at targets y (sigma_j, r^2 > 1e-8 mask), (M,3)."""
    m = y.shape[0]
    cs = _chunk_size(x.shape[0])
    parts = []
    for b in range(0, m, cs):
        def fn(yb, x_, a_):
            r = yb[:, None, :] - x_[None, :, :]
            r2 = (r * r).sum(-1)
            s2 = (s * s)[None, :]
            rho2 = r2 / s2
            q = (rho2 + 2.5) / ((rho2 + 1.0) ** 2.5 * s2 * s[None, :])
            q = torch.where(r2 > EPS, q, torch.zeros_like(q))
            c = torch.cross(r, a_[None, :, :].expand_as(r), dim=-1)
            return -C * (q[..., None] * c).sum(1)
        parts.append(checkpoint(fn, y[b:b + cs], x, a, use_reentrant=False))
    return torch.cat(parts)


def rk4_coupled_step(x, a, s, nu, dt):
    """Classical RK4 on dx/dt = u(x,a), da/dt = S(x,a) + D(x,a)."""
    def rhs(xx, aa):
        u, S, D = pair_sums(xx, aa, s, nu)
        return u, S + D
    k1x, k1a = rhs(x, a)
    k2x, k2a = rhs(x + 0.5 * dt * k1x, a + 0.5 * dt * k1a)
    k3x, k3a = rhs(x + 0.5 * dt * k2x, a + 0.5 * dt * k2a)
    k4x, k4a = rhs(x + dt * k3x, a + dt * k3a)
    return (x + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x),
            a + dt / 6 * (k1a + 2 * k2a + 2 * k3a + k4a))
