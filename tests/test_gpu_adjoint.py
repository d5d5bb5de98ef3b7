"""GPU tests of the hand-written adjoint.

* One evaluation: the HIP vector-Jacobian product of (velocity, stretching,
  PSE) equals torch.autograd of the FP64 torch oracle (oracle/vpm_torch.py),
  for every cotangent subset, uniform and per-particle radii, N up to 4096.
* Velocity at evaluation points: adjoint w.r.t. the sources likewise.
* Rollout (differentiation workload, scaled): the gradient of a probe-velocity
  loss after RK4 steps with checkpointing equals central finite differences
  (eps = 1e-4) and torch.autograd of the oracle rollout.
"""
import numpy as np
import pytest
import torch

from oracle import vpm_torch as T
from openonda_amd import autograd as AG
from openonda_amd import flow_models as FM
from openonda_amd import physics as P

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = torch.as_tensor(a).double().cpu(), torch.as_tensor(b).double().cpu()
    return float((a - b).norm() / b.norm())


def make(n, seed, uniform):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    a = rng.normal(0, 1, (n, 3)) / n
    s = np.full(n, 0.3) if uniform else rng.uniform(0.2, 0.4, n)
    nu = rng.uniform(1e-3, 5e-3, n)
    return x, a, s, nu


def oracle_vjp(x, a, s, nu, lam, mu, xi):
    dev = torch.device("cuda")
    tx = torch.from_numpy(x).to(dev).requires_grad_(True)
    ta = torch.from_numpy(a).to(dev).requires_grad_(True)
    u, S, D = T.pair_sums(tx, ta, torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev))
    L = 0.0
    for cot, out in ((lam, u), (mu, S), (xi, D)):
        if cot is not None:
            L = L + (cot.t() * out).sum()
    gx, ga = torch.autograd.grad(L, (tx, ta))
    return gx.t(), ga.t()


SUBSETS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 1, 1)]


@pytest.mark.parametrize("uniform", [True, False])
@pytest.mark.parametrize("subset", SUBSETS)
def test_vjp_matches_autograd(gpu, uniform, subset):
    n = 700
    x, a, s, nu = make(n, 3, uniform)
    print(f"[adjoint] uniform={uniform} subset={subset}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5)
    cots = [torch.randn(3, n, dtype=torch.float64, device="cuda", generator=g) if f else None
            for f in subset]
    ref_x, ref_a = oracle_vjp(x, a, s, nu, *cots)
    d = (P.to_soa(x), P.to_soa(a), torch.from_numpy(s).cuda(), torch.from_numpy(nu).cuda())
    mode = "uniform" if uniform else "general"
    for prec, tol in (("f64", 1e-11), ("f32c", 2e-5)):
        gx, ga = AG.adjoint(*d, *cots, precision=prec, sigma_mode=mode)
        assert rel(gx, ref_x) <= tol, (prec, "dx", rel(gx, ref_x))
        assert rel(ga, ref_a) <= tol, (prec, "da", rel(ga, ref_a))


def test_vjp_ring_4096(gpu):
    x, a, s, nu = FM.torus_ring(64, nu=1e-3)
    g = torch.Generator(device="cuda").manual_seed(6)
    cots = [torch.randn(3, len(x), dtype=torch.float64, device="cuda", generator=g) for _ in range(3)]
    ref_x, ref_a = oracle_vjp(x, a, s, nu, *cots)
    d = (P.to_soa(x), P.to_soa(a), torch.from_numpy(s).cuda(), torch.from_numpy(nu).cuda())
    gx, ga = AG.adjoint(*d, *cots, precision="f64", sigma_mode="uniform")
    assert rel(gx, ref_x) <= 1e-11 and rel(ga, ref_a) <= 1e-11
    gx, ga = AG.adjoint(*d, *cots, precision="f32c", sigma_mode="uniform")
    assert rel(gx, ref_x) <= 1e-4 and rel(ga, ref_a) <= 1e-4


def test_probe_velocity_adjoint(gpu):
    x, a, s, nu = make(500, 8, False)
    y = FM.ring_probes(300, seed=9)
    dev = torch.device("cuda")
    ty = torch.from_numpy(y).to(dev)
    tx = torch.from_numpy(x).to(dev).requires_grad_(True)
    ta = torch.from_numpy(a).to(dev).requires_grad_(True)
    lam = torch.randn(300, 3, dtype=torch.float64, device=dev)
    u = T.induced_velocity_at(ty, tx, ta, torch.from_numpy(s).to(dev))
    rx, ra = torch.autograd.grad((lam * u).sum(), (tx, ta))
    px = P.to_soa(x).requires_grad_(True)
    pa = P.to_soa(a).requires_grad_(True)
    uu = AG.probe_velocity(P.to_soa(y), px, pa, torch.from_numpy(s).cuda())
    assert rel(uu.t(), u.detach()) <= 1e-12
    (lam.t() * uu).sum().backward()
    assert rel(px.grad.t(), rx) <= 1e-11 and rel(pa.grad.t(), ra) <= 1e-11


def probe_loss(x0, a0, sigma, nu, probes, target, dt, steps, ck, prec):
    x, a = AG.rollout(x0, a0, sigma, nu, dt, steps, checkpoint_every=ck, precision=prec)
    u = AG.probe_velocity(probes, x, a, sigma)
    return ((u - target) ** 2).sum(0).mean()


def test_rollout_gradient_vs_finite_differences(gpu):
    """Differentiation workload at small scale: ring (N = 4096, 64 stations),
    10 RK4 steps of dt = 0.005 with checkpoints every 5 steps, loss = mean
    squared probe-velocity mismatch against a ring of radius 1.05."""
    x, a, s, nu = FM.torus_ring(64, nu=1e-3)
    xt, at, st, _ = FM.torus_ring(64, R=1.05, nu=1e-3)
    y = FM.ring_probes(512, seed=4)
    dev = torch.device("cuda")
    sig, nut = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
    probes = P.to_soa(y, dev)
    target = P.induced_velocity_at(probes, P.to_soa(xt, dev), P.to_soa(at, dev),
                                   torch.from_numpy(st).to(dev))
    x0 = P.to_soa(x, dev).requires_grad_(True)
    a0 = P.to_soa(a, dev).requires_grad_(True)
    L = probe_loss(x0, a0, sig, nut, probes, target, 0.005, 10, 5, "f64")
    gx, ga = torch.autograd.grad(L, (x0, a0))
    rng = np.random.default_rng(44)
    coords = [(int(rng.integers(0, 3)), int(rng.integers(0, len(x)))) for _ in range(32)]
    eps = 1e-4
    for which, (c, i) in enumerate(coords):
        base = x0 if which % 2 == 0 else a0
        grad = gx if which % 2 == 0 else ga
        h = eps * (1.0 if which % 2 == 0 else float(a0.detach().abs().max()))
        vals = []
        for sgn in (1, -1):
            pert = base.detach().clone()
            pert[c, i] += sgn * h
            args = (pert, a0.detach()) if which % 2 == 0 else (x0.detach(), pert)
            with torch.no_grad():
                vals.append(float(probe_loss(*args, sig, nut, probes, target, 0.005, 10, 5, "f64")))
        fd = (vals[0] - vals[1]) / (2 * h)
        scale = float(grad.abs().max())
        assert abs(float(grad[c, i]) - fd) <= 1e-4 * scale, (which, c, i, float(grad[c, i]), fd)
    # FP32-compensated rollout gradient agrees with the FP64 one
    L32 = probe_loss(x0, a0, sig, nut, probes, target, 0.005, 10, 5, "f32c")
    gx32, ga32 = torch.autograd.grad(L32, (x0, a0))
    assert rel(gx32, gx) <= 1e-4 and rel(ga32, ga) <= 1e-4


def test_rollout_gradient_vs_torch_autograd(gpu):
    x, a, s, nu = FM.torus_ring(32, nu=1e-3)  # N = 2048
    y = FM.ring_probes(256, seed=4)
    dev = torch.device("cuda")
    sig, nut = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
    lam = torch.randn(256, 3, dtype=torch.float64, device=dev)
    # oracle rollout, (N,3) layout
    tx = torch.from_numpy(x).to(dev).requires_grad_(True)
    ta = torch.from_numpy(a).to(dev).requires_grad_(True)
    xx, aa = tx, ta
    for _ in range(2):
        xx, aa = T.rk4_coupled_step(xx, aa, sig, nut, 0.005)
    u = T.induced_velocity_at(torch.from_numpy(y).to(dev), xx, aa, sig)
    rx, ra = torch.autograd.grad((lam * u).sum(), (tx, ta))
    px = P.to_soa(x, dev).requires_grad_(True)
    pa = P.to_soa(a, dev).requires_grad_(True)
    xs, as_ = AG.rollout(px, pa, sig, nut, 0.005, 2, checkpoint_every=1, precision="f64")
    uu = AG.probe_velocity(P.to_soa(y, dev), xs, as_, sig)
    (lam.t() * uu).sum().backward()
    assert rel(px.grad.t(), rx) <= 1e-10 and rel(pa.grad.t(), ra) <= 1e-10


@pytest.mark.parametrize("precision,tol", [("f64", 1e-13), ("f32c", 2e-7)])
@pytest.mark.parametrize("uniform", [True, False])
def test_adjoint_source_split_matches_single_pass(gpu, precision, tol, uniform):
    """Sources split among workgroups (small N) give the unsplit gradient."""
    x, a, s, nu = make(3000, 21, uniform)
    print(f"[adjoint-split] {precision} uniform={uniform}", flush=True)
    d = [P.to_soa(x), P.to_soa(a), torch.from_numpy(s).cuda(), torch.from_numpy(nu).cuda()]
    g = torch.Generator(device="cuda").manual_seed(5)
    cots = [torch.randn(3, len(x), dtype=torch.float64, device="cuda", generator=g) for _ in range(3)]
    mode = "uniform" if uniform else "general"
    one = AG.adjoint(*d, *cots, precision=precision, sigma_mode=mode, max_splits=1)
    auto = AG.adjoint(*d, *cots, precision=precision, sigma_mode=mode)
    two = AG.adjoint(*d, *cots, precision=precision, sigma_mode=mode, max_splits=2)
    for ref, got_a, got_2 in zip(one, auto, two):
        assert rel(got_a, ref) <= tol and rel(got_2, ref) <= tol
