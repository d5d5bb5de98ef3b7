"""CPU pins of the differentiable torch oracle: its values equal the NumPy
oracle (itself pinned in test_oracle_pins.py) and its autograd gradients equal
central finite differences."""
import numpy as np
import pytest
import torch

from oracle import vpm_oracle as O
from oracle import vpm_torch as T


def data(n, seed, uniform=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    a = rng.normal(0, 1, (n, 3)) / n
    s = np.full(n, 0.3) if uniform else rng.uniform(0.2, 0.4, n)
    nu = rng.uniform(1e-3, 5e-3, n)
    return x, a, s, nu


def test_values_match_numpy_oracle():
    x, a, s, nu = data(300, 1)
    tx, ta, ts, tn = (torch.from_numpy(v) for v in (x, a, s, nu))
    u, S, D = T.pair_sums(tx, ta, ts, tn)
    rel = lambda p, q: np.linalg.norm(p.detach().numpy() - q) / np.linalg.norm(q)
    assert rel(u, O.selfinduced_velocity(x, a, s)) < 1e-13
    assert rel(S, O.strength_gradients(x, a, s)) < 1e-13
    assert rel(D, O.strength_gradients_pse(x, a, s, nu)) < 1e-13
    y = np.random.default_rng(2).uniform(-1, 1, (50, 3))
    assert rel(T.induced_velocity_at(torch.from_numpy(y), tx, ta, ts), O.induced_velocity_at(y, x, a, s)) < 1e-13


def test_chunked_autograd_equals_finite_differences():
    x, a, s, nu = data(40, 3)
    tx = torch.from_numpy(x).requires_grad_(True)
    ta = torch.from_numpy(a).requires_grad_(True)
    ts, tn = torch.from_numpy(s), torch.from_numpy(nu)
    g = torch.Generator().manual_seed(0)
    lam = [torch.randn(40, 3, dtype=torch.float64, generator=g) for _ in range(3)]

    def loss(xx, aa):
        u, S, D = T.pair_sums(xx, aa, ts, tn)
        return (lam[0] * u).sum() + (lam[1] * S).sum() + (lam[2] * D).sum()

    L = loss(tx, ta)
    gx, ga = torch.autograd.grad(L, (tx, ta))
    h = 1e-6
    for (i, k) in [(0, 0), (7, 2), (39, 1)]:
        for base, grad in ((x, gx), (a, ga)):
            p = base.copy(); p[i, k] += h
            m = base.copy(); m[i, k] -= h
            if base is x:
                fd = (loss(torch.from_numpy(p), ta) - loss(torch.from_numpy(m), ta)) / (2 * h)
            else:
                fd = (loss(tx, torch.from_numpy(p)) - loss(tx, torch.from_numpy(m))) / (2 * h)
            assert float(grad[i, k]) == pytest.approx(float(fd), rel=1e-5, abs=1e-9 * float(grad.abs().max()))


def test_multi_chunk_gradient_equals_single_chunk(monkeypatch):
    """Chunks recomputed under checkpointing keep their own self-pair mask."""
    x, a, s, nu = data(60, 4)
    tx, ta = torch.from_numpy(x), torch.from_numpy(a)
    ts, tn = torch.from_numpy(s), torch.from_numpy(nu)
    g = torch.Generator().manual_seed(1)
    lam = [torch.randn(60, 3, dtype=torch.float64, generator=g) for _ in range(3)]

    def grads():
        xx, aa = tx.clone().requires_grad_(True), ta.clone().requires_grad_(True)
        u, S, D = T.pair_sums(xx, aa, ts, tn)
        L = (lam[0] * u).sum() + (lam[1] * S).sum() + (lam[2] * D).sum()
        return torch.autograd.grad(L, (xx, aa))

    one = grads()
    monkeypatch.setattr(T, "_chunk_size", lambda n, budget=0: 7)
    many = grads()
    for p, q in zip(one, many):
        assert torch.allclose(p, q, rtol=1e-12, atol=1e-14)
