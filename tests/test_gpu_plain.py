"""GPU parity of the plain FP64 drop-in kernels against the FP64 oracle.

Every reference function, both arithmetic variants, on seeded random inputs
(small N incl. the empty and one-particle cases), the torus-ring layout of the
plain-method workload (N = 4096) and non-uniform radii.  The HIP kernels
restate the reference arithmetic in FP64, so the bar is rel-L2 <= 1e-12 for
vector outputs and a relative error <= 1e-11 for the scalar reductions (only
the summation order differs).
"""
import numpy as np
import pytest

from oracle import vpm_oracle as O
from openonda_amd import ParticlesPhysicsHip as H
from openonda_amd import flow_models as FM

pytestmark = pytest.mark.gpu

VEC_TOL = 1e-12
SCAL_TOL = 1e-11


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def random_set(n, seed, uniform_sigma=False):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 3))
    a = rng.normal(0.0, 1.0, (n, 3)) / max(n, 1)
    s = np.full(n, 0.2) if uniform_sigma else rng.uniform(0.1, 0.3, n)
    nu = rng.uniform(1e-4, 1e-2, n)
    return x, a, s, nu


CASES = [(1, 0), (2, 1), (7, 2), (257, 3), (1000, 4)]


@pytest.fixture(autouse=True)
def _variant_reset():
    yield
    H.VARIANT = "cuda"
    H.PRECISION = "plain"


@pytest.mark.parametrize("variant", ["cuda", "cpu"])
@pytest.mark.parametrize("n,seed", CASES)
def test_vector_functions(gpu, variant, n, seed):
    print(f"[plain] variant={variant} n={n} seed={seed}", flush=True)
    H.VARIANT = variant
    x, a, s, nu = random_set(n, seed)
    y = np.random.default_rng(seed + 100).uniform(-1.2, 1.2, (max(1, n // 2), 3))
    checks = [
        (H.get_selfinduced_velocity(x, a, s), O.selfinduced_velocity(x, a, s, variant)),
        (H.get_induced_velocity_at(y, x, a, s), O.induced_velocity_at(y, x, a, s)),
        (H.get_vorticity_field_at(y, x, a, s), O.vorticity_field_at(y, x, a, s, variant)),
        (H.get_vorticity_field_at(x, x, a, s), O.vorticity_field_at(x, x, a, s, variant)),
        (H.get_strength_gradients(x, a, s), O.strength_gradients(x, a, s)),
        (H.get_strength_gradients_PSE(x, a, s, nu), O.strength_gradients_pse(x, a, s, nu, variant)),
    ]
    for k, (got, ref) in enumerate(checks):
        assert got.shape == ref.shape, k
        assert rel_l2(got, ref) <= VEC_TOL, (k, rel_l2(got, ref))


@pytest.mark.parametrize("variant", ["cuda", "cpu"])
@pytest.mark.parametrize("n,seed", CASES[1:])
def test_scalar_reductions(gpu, variant, n, seed):
    print(f"[plain-scalar] variant={variant} n={n}", flush=True)
    H.VARIANT = variant
    x, a, s, _ = random_set(n, seed)
    for got, ref in [
        (H.get_total_kinetic_energy(x, a, s), O.total_kinetic_energy(x, a, s, variant)),
        (H.get_total_helicity(x, a, s), O.total_helicity(x, a, s, variant)),
        (H.get_total_enstrophy(x, a, s), O.total_enstrophy(x, a, s, variant)),
    ]:
        assert isinstance(got, float)
        assert abs(got - ref) <= SCAL_TOL * max(abs(ref), 1e-300), (got, ref)


def test_one_particle_and_empty(gpu):
    x, a, s, nu = random_set(1, 9)
    assert np.all(H.get_selfinduced_velocity(x, a, s) == 0.0)
    assert np.all(H.get_strength_gradients(x, a, s) == 0.0)
    assert H.get_total_kinetic_energy(x, a, s) == 0.0
    # enstrophy includes i = j: (3.2 s^2)/(s^2)^1.5 |a|^2 /(8 pi)/(4 pi)
    assert H.get_total_enstrophy(x, a, s) == pytest.approx(O.total_enstrophy(x, a, s), rel=1e-13)
    e = np.zeros((0, 3))
    out = H.get_selfinduced_velocity(e, e, np.zeros(0))
    assert out.shape == (0, 3)
    assert H.get_total_enstrophy(e, e, np.zeros(0)) == 0.0


def test_plain_ring_4096(gpu):
    """Plain-method workload: torus ring, N = 4096, velocity + stretching."""
    x, a, s, nu = FM.torus_ring(64)
    assert len(x) == 4096
    assert rel_l2(H.get_selfinduced_velocity(x, a, s), O.selfinduced_velocity(x, a, s)) <= VEC_TOL
    assert rel_l2(H.get_strength_gradients(x, a, s), O.strength_gradients(x, a, s)) <= VEC_TOL


def test_input_validation(gpu):
    x, a, s, nu = random_set(8, 1)
    with pytest.raises(TypeError):
        H.get_selfinduced_velocity(x.astype(np.float32), a, s)
    with pytest.raises(ValueError):
        H.get_selfinduced_velocity(x[:, :2].copy(), a, s)
    with pytest.raises(ValueError):
        H.get_selfinduced_velocity(x, a[:5], s)


def test_determinism(gpu):
    x, a, s, _ = random_set(3000, 5)
    e1 = H.get_total_kinetic_energy(x, a, s)
    e2 = H.get_total_kinetic_energy(x, a, s)
    assert e1 == e2
    u1 = H.get_selfinduced_velocity(x, a, s)
    u2 = H.get_selfinduced_velocity(x, a, s)
    assert np.array_equal(u1, u2)


@pytest.mark.parametrize("precision,tol", [("f64", VEC_TOL), ("f32c", 1e-6)])
@pytest.mark.parametrize("uniform", [True, False])
def test_fused_precision_switch(gpu, precision, tol, uniform):
    """PRECISION = "f64" / "f32c" routes velocity, stretching, PSE and
    enstrophy through the fused pass; the results stay within its bar of the
    oracle, and the plain-only functions are unaffected."""
    x, a, s, nu = random_set(1500, 9, uniform_sigma=uniform)
    H.PRECISION = precision
    assert rel_l2(H.get_selfinduced_velocity(x, a, s), O.selfinduced_velocity(x, a, s)) <= tol
    assert rel_l2(H.get_strength_gradients(x, a, s), O.strength_gradients(x, a, s)) <= tol
    assert rel_l2(H.get_strength_gradients_PSE(x, a, s, nu),
                  O.strength_gradients_pse(x, a, s, nu)) <= tol
    ens = H.get_total_enstrophy(x, a, s)
    assert abs(ens - O.total_enstrophy(x, a, s)) <= 10 * tol * abs(O.total_enstrophy(x, a, s))
    assert rel_l2(H.get_induced_velocity_at(x[:50], x, a, s),
                  O.induced_velocity_at(x[:50], x, a, s)) <= VEC_TOL
