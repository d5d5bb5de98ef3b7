"""Symmetry pins of the FP64 oracle's pair sums (CPU only).

The pair sums of the vortex-particle method are Euclidean-equivariant and
homogeneous in the strengths, whatever the smoothing kernel: under a proper
rotation Q and a shift b (x -> Q x + b, alpha -> Q alpha; alpha is a
pseudo-vector, so only det Q = +1 is used) every vector output rotates with
Q and every scalar is unchanged; relabelling the particles permutes the
vector outputs and leaves the scalars; velocity, PSE and vorticity are
linear in alpha, stretching and the energy-like scalars quadratic.  A
mistake in one component of a cross product, a wrong index or a stray
absolute coordinate breaks one of these, independently of the formula each
function restates (ParticlesPhysicsCuda.py, ParticlesPhysics.py).
"""
import numpy as np
import pytest

from oracle import vpm_oracle as O

N = 48
TOL = 1e-12


def rand_set(seed, n=N):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 3))
    a = rng.normal(0.0, 1.0, (n, 3))
    s = rng.uniform(0.15, 0.35, n)
    nu = rng.uniform(1e-3, 1e-2, n)
    return x, a, s, nu


def rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


# (name, function of (x, a, s, nu, targets), kind, degree in alpha)
OPS = [
    ("selfinduced_velocity[cuda]", lambda x, a, s, nu, t: O.selfinduced_velocity(x, a, s), "vec", 1),
    ("selfinduced_velocity[cpu]",
     lambda x, a, s, nu, t: O.selfinduced_velocity(x, a, s, variant="cpu"), "vec", 1),
    ("induced_velocity_at", lambda x, a, s, nu, t: O.induced_velocity_at(t, x, a, s), "tvec", 1),
    ("vorticity_field_at[cuda]", lambda x, a, s, nu, t: O.vorticity_field_at(t, x, a, s), "tvec", 1),
    ("vorticity_field_at[cpu]",
     lambda x, a, s, nu, t: O.vorticity_field_at(t, x, a, s, variant="cpu"), "tvec", 1),
    ("strength_gradients", lambda x, a, s, nu, t: O.strength_gradients(x, a, s), "vec", 2),
    ("strength_gradients_pse[cuda]",
     lambda x, a, s, nu, t: O.strength_gradients_pse(x, a, s, nu), "vec", 1),
    ("strength_gradients_pse[cpu]",
     lambda x, a, s, nu, t: O.strength_gradients_pse(x, a, s, nu, variant="cpu"), "vec", 1),
    ("total_kinetic_energy[cuda]", lambda x, a, s, nu, t: O.total_kinetic_energy(x, a, s), "scal", 2),
    ("total_kinetic_energy[cpu]",
     lambda x, a, s, nu, t: O.total_kinetic_energy(x, a, s, variant="cpu"), "scal", 2),
    ("total_helicity[cuda]", lambda x, a, s, nu, t: O.total_helicity(x, a, s), "scal", 2),
    ("total_helicity[cpu]", lambda x, a, s, nu, t: O.total_helicity(x, a, s, variant="cpu"), "scal", 2),
    ("total_enstrophy[cuda]", lambda x, a, s, nu, t: O.total_enstrophy(x, a, s), "scal", 2),
    ("total_enstrophy[cpu]", lambda x, a, s, nu, t: O.total_enstrophy(x, a, s, variant="cpu"), "scal", 2),
]
IDS = [o[0] for o in OPS]


def targets(seed):
    return np.random.default_rng(seed + 100).uniform(-1.2, 1.2, (17, 3))


@pytest.mark.parametrize("name,fn,kind,deg", OPS, ids=IDS)
def test_rotation_and_translation(name, fn, kind, deg):
    x, a, s, nu = rand_set(1)
    t = targets(1)
    q = rotation(2)
    b = np.array([0.37, -1.2, 2.9])
    base = fn(x, a, s, nu, t)
    moved = fn(x @ q.T + b, a @ q.T, s, nu, t @ q.T + b)
    if kind == "scal":
        assert moved == pytest.approx(base, rel=TOL)
    else:
        assert rel(moved, base @ q.T) <= TOL


@pytest.mark.parametrize("name,fn,kind,deg", OPS, ids=IDS)
def test_relabelling(name, fn, kind, deg):
    x, a, s, nu = rand_set(3)
    t = targets(3)
    p = np.random.default_rng(4).permutation(N)
    base = fn(x, a, s, nu, t)
    perm = fn(x[p], a[p], s[p], nu[p], t)
    if kind == "scal":
        assert perm == pytest.approx(base, rel=TOL)
    elif kind == "vec":
        assert rel(perm, base[p]) <= TOL
    else:  # outputs at fixed targets: the source order only reorders the sum
        assert rel(perm, base) <= TOL


@pytest.mark.parametrize("name,fn,kind,deg", OPS, ids=IDS)
def test_homogeneity_in_the_strengths(name, fn, kind, deg):
    x, a, s, nu = rand_set(5)
    t = targets(5)
    base = np.asarray(fn(x, a, s, nu, t))
    for c in (-1.0, 2.0, 0.3):
        got = np.asarray(fn(x, c * a, s, nu, t))
        want = c ** deg * base
        if kind == "scal":
            assert float(got) == pytest.approx(float(want), rel=TOL)
        else:
            assert rel(got, want) <= TOL


def test_strengths_superpose_in_the_linear_sums():
    """Velocity, PSE and vorticity are linear in alpha: f(a1 + a2) = f(a1) + f(a2)."""
    x, a1, s, nu = rand_set(6)
    a2 = rand_set(7)[1]
    t = targets(6)
    for name, fn, kind, deg in OPS:
        if deg != 1:
            continue
        lhs = fn(x, a1 + a2, s, nu, t)
        rhs = fn(x, a1, s, nu, t) + fn(x, a2, s, nu, t)
        assert rel(lhs, rhs) <= TOL, name
