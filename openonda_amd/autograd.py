"""Differentiable pair sums and rollouts.

* :class:`PairSums` -- the fused forward pass (velocity, stretching, PSE)
  with the hand-written adjoint kernel (``vpm_adjoint``) as its backward.
* :class:`ProbeVelocity` -- velocity at fixed evaluation points
  (``vpm_get_induced_velocity_at``) with ``vpm_induced_velocity_at_adjoint``
  as its backward.
* :func:`rk4_coupled_step` / :func:`rollout` -- classical RK4 on
  dx/dt = u(x, alpha), dalpha/dt = S(x, alpha) + D(x, alpha) with step
  checkpointing: every ``checkpoint_every`` steps only the FP64 state is kept
  and the segment is recomputed during the backward pass.

Tensors are (3, N) float64 SoA on the HIP device, as in
:mod:`openonda_amd.physics`.  Gradients flow to positions and strengths;
radii and viscosities are treated as constants.  The reference solver has no
gradients; this layer has no counterpart there.
"""
from __future__ import annotations

import ctypes as C

import torch
from torch.utils.checkpoint import checkpoint

from . import _lib as L
from . import physics as P

_ALL = ("vel", "stretch", "pse")


def _resolve_sigma_mode(sigma, sigma_mode):
    """"auto" -> "uniform" when every radius equals sigma[0], else "general"
    (one device check); "uniform" and "general" pass through."""
    if sigma_mode == "auto":
        return "uniform" if P.sigma_is_uniform(sigma) else "general"
    if sigma_mode not in ("uniform", "general"):
        raise ValueError("sigma_mode must be 'auto', 'uniform' or 'general'")
    return sigma_mode


def _opts(precision, sigma_mode, presorted=False, max_splits=0):
    return L.ForwardOpts(0, P.PRECISIONS[precision],
                         L.SIGMA_UNIFORM if sigma_mode == "uniform" else L.SIGMA_GENERAL,
                         1 if presorted else 0, -1, int(max_splits))


def adjoint(pos, alpha, sigma, nu, lam=None, mu=None, xi=None, precision="f32c",
            sigma_mode="auto", max_splits=0):
    """dL/dx, dL/dalpha for L = <lam, u> + <mu, S> + <xi, D> (each (3,N) or None).

    ``sigma_mode``: "auto" checks the radii on the device, "uniform" asserts
    sigma_i == sigma_0 (unchecked), "general" uses sigma_ij per pair.
    ``max_splits`` caps the source splits used when the targets alone cannot
    fill the device (0: automatic, 1: never split)."""
    sigma_mode = _resolve_sigma_mode(sigma, sigma_mode)
    p = P._particles(pos, alpha, sigma, nu if xi is not None else None)
    n = p.n
    cot = [t.contiguous() if t is not None else None for t in (lam, mu, xi)]
    for t in cot:
        if t is not None:
            P._vec(t, "cotangent", n)
    ld = max(n, 1)
    ain = L.AdjointIn(ld, *[t.data_ptr() if t is not None else None for t in cot])
    dpos = torch.empty((3, n), dtype=torch.float64, device=pos.device)
    dstr = torch.empty((3, n), dtype=torch.float64, device=pos.device)
    aout = L.AdjointOut(ld, dpos.data_ptr(), dstr.data_ptr())
    opts = _opts(precision, sigma_mode, max_splits=max_splits)
    lib = L.load()
    nb = lib.vpm_adjoint_workspace_bytes(n, C.byref(opts))
    if nb == 0:
        raise L.VpmError("vpm_adjoint_workspace_bytes", -3, lib.vpm_last_error().decode())
    ws = P.workspace(pos.device, nb)
    L.call("vpm_adjoint", C.byref(p), C.byref(ain), C.byref(opts), C.byref(aout), ws.data_ptr(),
           ws.numel(), P._stream(pos))
    return dpos, dstr


class PairSums(torch.autograd.Function):
    """(vel, stretch, pse) = pair sums of (pos, alpha); backward = HIP adjoint."""

    @staticmethod
    def forward(ctx, pos, alpha, sigma, nu, outputs, precision, sigma_mode):
        sigma_mode = _resolve_sigma_mode(sigma, sigma_mode)
        res = P.forward(pos.detach(), alpha.detach(), sigma, nu if "pse" in outputs else None,
                        outputs=outputs, precision=precision, sigma_mode=sigma_mode)
        ctx.save_for_backward(pos, alpha, sigma, nu)
        ctx.outputs, ctx.precision, ctx.sigma_mode = outputs, precision, sigma_mode
        return tuple(res[k] for k in outputs)

    @staticmethod
    def backward(ctx, *grads):
        pos, alpha, sigma, nu = ctx.saved_tensors
        g = dict(zip(ctx.outputs, grads))
        dpos, dstr = adjoint(pos.detach(), alpha.detach(), sigma, nu, g.get("vel"),
                             g.get("stretch"), g.get("pse"), ctx.precision, ctx.sigma_mode)
        return dpos, dstr, None, None, None, None, None


def pair_sums(pos, alpha, sigma, nu=None, outputs=_ALL, precision="f32c", sigma_mode="auto"):
    """Differentiable (vel, stretch, pse) of (pos, alpha); ``sigma_mode`` as in
    :func:`adjoint`."""
    if nu is None:
        nu = torch.zeros_like(sigma)
    return PairSums.apply(pos, alpha, sigma, nu, tuple(outputs), precision, sigma_mode)


class ProbeVelocity(torch.autograd.Function):
    """u at fixed targets (3,M) from sources; backward w.r.t. the sources."""

    @staticmethod
    def forward(ctx, targets, pos, alpha, sigma):
        u = P.induced_velocity_at(targets, pos.detach(), alpha.detach(), sigma)
        ctx.save_for_backward(targets, pos, alpha, sigma)
        return u

    @staticmethod
    def backward(ctx, gu):
        targets, pos, alpha, sigma = ctx.saved_tensors
        gu = gu.contiguous()
        pp = P._particles(pos.detach(), alpha.detach(), sigma)
        tp = P._points(targets)
        n = pp.n
        dpos = torch.empty((3, n), dtype=torch.float64, device=pos.device)
        dstr = torch.empty((3, n), dtype=torch.float64, device=pos.device)
        L.call("vpm_induced_velocity_at_adjoint", C.byref(tp), C.byref(pp), gu.data_ptr(),
               P._ld(gu), dpos.data_ptr(), dstr.data_ptr(), max(n, 1), P._stream(pos))
        return None, dpos, dstr, None


def probe_velocity(targets, pos, alpha, sigma):
    return ProbeVelocity.apply(targets, pos, alpha, sigma)


def rk4_coupled_step(pos, alpha, sigma, nu, dt, precision="f32c", sigma_mode="auto",
                     pse=True):
    """One classical RK4 step of the coupled particle ODE (the radii are
    constant through it: "auto" is resolved once)."""
    sigma_mode = _resolve_sigma_mode(sigma, sigma_mode)
    outs = ("vel", "stretch", "pse") if pse else ("vel", "stretch")

    def rhs(x, a):
        r = pair_sums(x, a, sigma, nu, outs, precision, sigma_mode)
        return r[0], (r[1] + r[2]) if pse else r[1]

    k1x, k1a = rhs(pos, alpha)
    k2x, k2a = rhs(pos + 0.5 * dt * k1x, alpha + 0.5 * dt * k1a)
    k3x, k3a = rhs(pos + 0.5 * dt * k2x, alpha + 0.5 * dt * k2a)
    k4x, k4a = rhs(pos + dt * k3x, alpha + dt * k3a)
    return (pos + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x),
            alpha + dt / 6 * (k1a + 2 * k2a + 2 * k3a + k4a))


def rollout(pos, alpha, sigma, nu, dt, steps, checkpoint_every=10, precision="f32c",
            sigma_mode="auto", pse=True):
    """``steps`` RK4 steps; with checkpoint_every > 0 each segment is
    recomputed in the backward pass instead of storing its stage tensors."""
    sigma_mode = _resolve_sigma_mode(sigma, sigma_mode)
    x, a = pos, alpha
    k = 0
    while k < steps:
        m = min(checkpoint_every if checkpoint_every > 0 else steps, steps - k)

        def segment(xx, aa, m=m):
            for _ in range(m):
                xx, aa = rk4_coupled_step(xx, aa, sigma, nu, dt, precision, sigma_mode, pse)
            return xx, aa

        if checkpoint_every > 0 and (x.requires_grad or a.requires_grad):
            x, a = checkpoint(segment, x, a, use_reentrant=False)
        else:
            x, a = segment(x, a)
        k += m
    return x, a
