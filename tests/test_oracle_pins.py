"""Pins of the FP64 oracle (CPU only).

No reference test holds numbers for this path (the reference's only VPM test
checks constructor attributes, tests/test1_import_modules.py:23-58), so each
oracle function is pinned here by something other than its own formula:

* velocity  -- q(rho) from quadrature of zeta (closed form independent of the
  oracle's (rho^2+5/2)/(rho^2+1)^(5/2)); the ring-centre velocity Gamma/(2R);
  curl u equals the CPU plugin's full-curl vorticity (finite differences).
* vorticity -- lattice partition of unity (integral of zeta = 1); q' = 4 pi rho^2 zeta.
* stretching -- transpose identity S_i = grad_{x_i}(alpha_i . u_i) with the
  already pinned velocity (finite differences).
* PSE       -- exact conservation of sum alpha with uniform nu, sigma; the
  Laplacian on a lattice (second moment 2 of the 105/(4pi) A^-9/2 kernel).
* energy    -- the singular Biot-Savart energy (Winckelmans-Leonard) as sigma -> 0.
* helicity  -- equals (1/4pi) sum alpha_i . u_i with the pinned velocity.
* enstrophy -- hand-computed single-particle and far-pair values.
* impulse   -- ring impulse Gamma pi R^2 along the axis; angular impulse by
  hand-computed one- and two-particle values and the x(x.a) - a|x|^2 form.
* core spreading -- a hand value; sigma^2 additive in dt and linear in nu.
* LES filter -- angle thresholds 45/135 degrees; C_nu recursion and clamp.
* time step -- the strength-history trace 1, 1, 9/11, 1.033, 0.899.
* all pair sums -- scalar double loops at tiny N, and the independent C port.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import vpm_oracle as O
from oracle.vpm_steps import ReferenceStepper
from openonda_amd import flow_models as FM

C = 1.0 / (4.0 * math.pi)


def zeta(rho):
    return 15.0 / (8.0 * math.pi) * (rho * rho + 1.0) ** -3.5


def q_quad(rho):
    return integrate.quad(lambda s: 4 * math.pi * s * s * zeta(s), 0.0, rho, epsabs=0, epsrel=1e-13)[0]


def rand_set(n, seed, sig=(0.1, 0.3)):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3))
    a = rng.normal(0, 1, (n, 3))
    s = rng.uniform(*sig, n)
    nu = rng.uniform(1e-3, 1e-2, n)
    return x, a, s, nu


# ------------------------------------------------------------------ velocity
@pytest.mark.parametrize("rho", [0.05, 0.3, 1.0, 2.7, 10.0, 40.0])
def test_velocity_kernel_is_the_integral_of_zeta(rho):
    sigma = 0.2
    r = rho * sigma
    x = np.array([[0.0, 0.0, 0.0]])
    a = np.array([[0.0, 0.0, 1.0]])
    y = np.array([[r, 0.0, 0.0]])
    u = O.induced_velocity_at(y, x, a, np.array([sigma]))[0]
    # regularised Biot-Savart: u = -(1/4pi) q(rho)/r^3 (r x alpha); r x e_z = (0,-r,0)
    expect = np.array([0.0, C * q_quad(rho) / r ** 2, 0.0])
    assert np.allclose(u, expect, rtol=1e-11, atol=0)
    assert integrate.quad(lambda s: 4 * math.pi * s * s * zeta(s), 0, np.inf)[0] == pytest.approx(1.0, rel=1e-12)


def test_ring_centre_velocity_is_gamma_over_2R():
    x, a, s, _ = FM.torus_ring(256, renormalise=True)
    u0 = O.induced_velocity_at(np.zeros((1, 3)), x, a, s)[0]
    # thin-core limit: Gamma/(2R) along +x for vorticity along e_theta
    assert u0[0] == pytest.approx(0.5, rel=0.02)
    assert abs(u0[1]) < 1e-12 and abs(u0[2]) < 1e-12


def _curl(f, y, h=1e-5):
    J = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        J[:, k] = (f(y + e) - f(y - e)) / (2 * h)
    return np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])


def test_curl_of_velocity_is_full_vorticity():
    x, a, s, _ = rand_set(40, 3)
    y = np.array([0.31, -0.22, 0.17])
    f = lambda p: O.induced_velocity_at(p[None, :], x, a, s)[0]
    w_fd = _curl(f, y)
    w = O.vorticity_field_at(y[None, :], x, a, s, variant="cpu")[0]
    assert np.allclose(w_fd, w, rtol=1e-6, atol=1e-6 * np.abs(w).max())


def test_velocity_variants_agree_for_uniform_sigma():
    x, a, s, _ = rand_set(60, 4, sig=(0.2, 0.2))
    assert np.allclose(O.selfinduced_velocity(x, a, s, "cuda"),
                       O.selfinduced_velocity(x, a, s, "cpu"), rtol=1e-13, atol=1e-15)
    # self-induced == induced_at the particle positions (the self pair is masked in both)
    assert np.allclose(O.selfinduced_velocity(x, a, s), O.induced_velocity_at(x, x, a, s),
                       rtol=1e-13, atol=1e-15)


# ------------------------------------------------------------------ vorticity
def test_vorticity_partition_of_unity():
    sigma, h = 0.2, 0.05
    g = np.arange(-16, 17) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + 0.013  # target off-lattice
    a = np.tile([0.0, 0.0, h ** 3], (len(x), 1))
    w = O.vorticity_field_at(np.zeros((1, 3)), x, a, np.full(len(x), sigma))[0]
    rmax = 16 * h / sigma  # 4 sigma: compare with the truncated integral of zeta
    expect = integrate.quad(lambda s: 4 * math.pi * s * s * zeta(s), 0, rmax)[0]
    assert w[2] == pytest.approx(expect, rel=2e-2)
    assert abs(w[0]) < 1e-12 and abs(w[1]) < 1e-12


def test_vorticity_kernel_is_derivative_of_velocity_kernel():
    sigma = 0.3
    for rho in (0.2, 1.0, 3.0):
        r = rho * sigma
        src = (np.zeros((1, 3)), np.array([[0.0, 0.0, 1.0]]), np.array([sigma]))
        z = O.vorticity_field_at(np.array([[r, 0, 0]]), *src)[0][2] * sigma ** 3
        def q_of(rh):
            u = O.induced_velocity_at(np.array([[rh * sigma, 0, 0]]), *src)[0][1]
            return u * (rh * sigma) ** 2 / C
        dq = (q_of(rho + 1e-6) - q_of(rho - 1e-6)) / 2e-6
        assert dq == pytest.approx(4 * math.pi * rho ** 2 * z, rel=1e-7)


# ------------------------------------------------------------------ stretching
def test_stretching_transpose_identity():
    x, a, s, _ = rand_set(30, 5)
    S = O.strength_gradients(x, a, s)
    h = 1e-6
    for i in (0, 7, 29):
        g = np.zeros(3)
        for k in range(3):
            vals = []
            for sgn in (1, -1):
                xp = x.copy()
                xp[i, k] += sgn * h
                vals.append(a[i] @ O.selfinduced_velocity(xp, a, s)[i])
            g[k] = (vals[0] - vals[1]) / (2 * h)
        assert np.allclose(g, S[i], rtol=1e-6, atol=1e-7 * np.abs(S[i]).max())


# ------------------------------------------------------------------ PSE
def test_pse_conserves_total_strength():
    x, a, s, nu = rand_set(80, 6, sig=(0.15, 0.15))
    D = O.strength_gradients_pse(x, a, s, np.full(len(x), 2e-3))
    assert np.abs(D.sum(0)).max() < 1e-12 * np.abs(D).sum()


def test_pse_is_a_laplacian():
    sigma, h, nu = 0.1, 0.05, 1e-3
    g = np.arange(-12, 13) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    omega = np.zeros_like(x)
    omega[:, 0] = x[:, 0] ** 2 + 0.5 * x[:, 1] ** 2  # lap = 3
    a = omega * h ** 3
    centre = np.argmin(np.sum(x * x, 1))
    sub = np.sum(x * x, 1) <= (12 * h) ** 2  # sphere of 6 sigma
    xs, as_ = x[sub], a[sub]
    D = O.strength_gradients_pse(xs, as_, np.full(len(xs), sigma), np.full(len(xs), nu))
    c = np.argmin(np.sum(xs * xs, 1))
    # D_i = nu sigma^3 lap(omega) (implicit volume sigma^3), second moment truncated at 6 sigma
    m2 = 35 * integrate.quad(lambda r: r ** 4 * (1 + r * r) ** -4.5, 0, 6.0)[0]
    assert D[c, 0] == pytest.approx(nu * sigma ** 3 * 3.0 * m2 / 2.0, rel=1e-2)
    assert abs(D[c, 1]) < 1e-15 and abs(D[c, 2]) < 1e-15
    assert 35 * integrate.quad(lambda r: r ** 4 * (1 + r * r) ** -4.5, 0, np.inf)[0] == pytest.approx(2.0)


def test_pse_variants_coincide_for_uniform_viscosity():
    x, a, s, _ = rand_set(50, 7)
    nu = np.full(len(x), 3e-3)
    Dg = O.strength_gradients_pse(x, a, s, nu, "cuda")
    Dc = O.strength_gradients_pse(x, a, s, nu, "cpu")
    assert np.linalg.norm(Dg - Dc) / np.linalg.norm(Dg) < 1e-6  # only the A + 1e-8 differs


# ------------------------------------------------------------------ scalars
def test_energy_singular_limit():
    x, a, s, _ = rand_set(25, 8)
    s = np.full(len(x), 1e-7)
    E = O.total_kinetic_energy(x, a, s)
    sing = 0.0
    for i in range(len(x)):
        for j in range(len(x)):
            if i != j:
                r = x[i] - x[j]
                rn = np.linalg.norm(r)
                e = r / rn
                sing += (a[i] @ a[j] + (a[i] @ e) * (a[j] @ e)) / rn
    assert E == pytest.approx(sing / (16 * math.pi), rel=1e-9)


def test_helicity_is_alpha_dot_velocity():
    x, a, s, _ = rand_set(50, 9)
    u = O.selfinduced_velocity(x, a, s, "cuda")
    hsum = float(np.sum(a * u))
    assert O.total_helicity(x, a, s, "cpu") == pytest.approx(hsum, rel=1e-11)
    assert O.total_helicity(x, a, s, "cuda") == pytest.approx(C * hsum, rel=1e-11)


def test_enstrophy_known_values():
    a1 = np.array([[1.0, 0.0, 0.0]])
    # one particle, sigma = 0.1: (3.2 s^2)/(s^2)^1.5 |a|^2 = 32, times 1/(8pi) [x 1/(4pi)]
    assert O.total_enstrophy(np.zeros((1, 3)), a1, np.array([0.1]), "cpu") == pytest.approx(32 / (8 * math.pi), rel=1e-14)
    assert O.total_enstrophy(np.zeros((1, 3)), a1, np.array([0.1]), "cuda") == pytest.approx(32 / (32 * math.pi ** 2), rel=1e-14)
    # two far particles: self terms + 2 (a1.a2)/r
    x = np.array([[0.0, 0, 0], [1e6, 0, 0]])
    a = np.array([[1.0, 0, 0], [0.5, 0.5, 0]])
    e = O.total_enstrophy(x, a, np.array([0.1, 0.1]), "cpu")
    assert e == pytest.approx((32 * 1.0 + 32 * 0.5 + 2 * 0.5 / 1e6) / (8 * math.pi), rel=1e-12)


def test_ring_linear_impulse():
    x, a, s, _ = FM.torus_ring(128, renormalise=True)
    I = O.total_linear_impulse(x, a)
    assert I[0] == pytest.approx(math.pi, rel=0.02)  # Gamma pi R^2
    assert abs(I[1]) < 1e-12 and abs(I[2]) < 1e-12


def test_angular_impulse_hand_values():
    # one particle at (0,1,0) with alpha = (0,0,1):
    # x x alpha = (1,0,0), x x (1,0,0) = (0,0,-1), A = (0,0,-1)/3
    A = O.total_angular_impulse(np.array([[0.0, 1.0, 0.0]]), np.array([[0.0, 0.0, 1.0]]))
    assert np.array_equal(A, np.array([0.0, 0.0, -1.0]) / 3.0)
    # two particles: (1,0,0) with (0,1,0) gives (0,-1,0); (0,0,2) with (1,0,0)
    # gives (-4,0,0); A = (-4,-1,0)/3
    A2 = O.total_angular_impulse(np.array([[1.0, 0, 0], [0, 0, 2.0]]),
                                 np.array([[0.0, 1, 0], [1.0, 0, 0]]))
    assert np.allclose(A2, np.array([-4.0, -1.0, 0.0]) / 3.0, rtol=0, atol=1e-15)


def test_angular_impulse_is_the_bac_cab_form():
    # x x (x x a) = x (x.a) - a |x|^2, summed independently of any cross product
    rng = np.random.default_rng(31)
    x = rng.normal(0, 1, (200, 3)) + [0.3, -1.2, 0.7]
    a = rng.normal(0, 1, (200, 3))
    ref = sum(xi * float(xi @ ai) - ai * float(xi @ xi) for xi, ai in zip(x, a)) / 3.0
    assert np.allclose(O.total_angular_impulse(x, a), ref, rtol=1e-13, atol=1e-12)
    # and it is not degenerate here: every component is O(1)
    assert np.abs(ref).min() > 1e-2


def test_core_spreading_hand_value_and_additivity():
    # sigma = 0.1, nu = 1e-3, dt = 0.01: sqrt(0.01 + 5.34e-5) = 0.1002666445...
    s = O.core_spreading(np.array([0.1]), np.array([1e-3]), 0.01)
    assert s[0] == pytest.approx(0.10026664450354367, rel=1e-15)
    # sigma^2 grows linearly in nu t: two steps of dt equal one step of 2 dt,
    # and the growth is linear in nu
    r = np.array([0.05, 0.2, 1.3])
    nu = np.array([1e-3, 4e-2, 0.0])
    two = O.core_spreading(O.core_spreading(r, nu, 0.02), nu, 0.02)
    one = O.core_spreading(r, nu, 0.04)
    assert np.allclose(two, one, rtol=1e-15, atol=0)
    d1 = O.core_spreading(r, nu, 0.1) ** 2 - r ** 2
    d2 = O.core_spreading(r, 2 * nu, 0.1) ** 2 - r ** 2
    assert np.allclose(d2, 2 * d1, rtol=1e-12, atol=1e-18)
    assert d1[2] == 0.0


# ------------------------------------------------------------------ LES filter
def test_les_angle_threshold():
    angles = np.radians([0, 44, 46, 90, 134, 136, 180])
    w = np.stack([np.sin(angles), np.zeros_like(angles), np.cos(angles)], 1) * 2.0
    # mirror copies (x -> -x) cancel the mean's x part; a last particle makes
    # the mean vorticity point exactly along +z
    w = np.vstack([w, w * [-1, 1, 1], [0, 0, 100.0]])
    nut, write = O.les_particle_viscosity(w, np.full(len(w), 0.1), 0.05)
    assert list((nut[:-1] > 0).astype(int)) == [0, 0, 1, 1, 1, 0, 0] * 2
    assert nut[3] == pytest.approx((0.05 * 0.1) ** 2 * 2.0)
    assert write


def test_les_constant_recursion():
    c, ref = O.les_filter_constant(4.0, 0.02958, None)  # first call: e_ref = e
    assert ref == pytest.approx(2.0) and c == pytest.approx(0.02958 * 2.0 / (2.0 + 1e-9))
    c2, ref2 = O.les_filter_constant(16.0, c, ref)      # e = 4, e_ref = 0.4 + 1.8 = 2.2
    assert ref2 == pytest.approx(2.2) and c2 == pytest.approx(c * 4.0 / (2.2 + 1e-9))
    c3, _ = O.les_filter_constant(1e8, 0.07, 1.0)       # clamp at 0.0761
    assert c3 == 0.0761


# ------------------------------------------------------------------ time step
def test_strength_history_trace():
    st = ReferenceStepper(np.zeros((1, 3)), np.array([[0.0, 0.0, 1.0]]), np.array([0.1]),
                          np.zeros(1), dt=0.01, method="Euler", flow_model="DNS",
                          viscous_scheme="CoreSpreading")
    mags = []
    for _ in range(5):
        st.step()
        mags.append(float(np.linalg.norm(st.alpha[0])))
    assert mags[0] == pytest.approx(1.0) and mags[1] == pytest.approx(1.0)
    assert mags[2] == pytest.approx(9 / 11)
    assert mags[3] == pytest.approx((12 * 9 / 11 + 16) / 25)
    assert mags[4] == pytest.approx((12 * mags[3] + 16 * 9 / 11 - 3) / 25)


def test_rk4_step_of_a_ring_moves_along_axis():
    x, a, s, nu = FM.torus_ring(32, renormalise=True)
    st = ReferenceStepper(x, a, s, nu, dt=0.05, method="RK4", flow_model="Potential",
                          viscous_scheme=None)
    st.step()
    dx = st.x - x
    # the core swirls, so only the strength-weighted centroid translates
    w = np.linalg.norm(a, axis=1)
    assert np.sum(w * dx[:, 0]) / np.sum(w) > 0
    assert abs(np.sum(w * dx[:, 1])) < 1e-12 and abs(np.sum(w * dx[:, 2])) < 1e-12
    assert st.time_step == 1 and st.flow_time == pytest.approx(0.05)


# ------------------------------------------------------------------ brute force
def _brute(x, a, s, nu):
    n = len(x)
    u = np.zeros((n, 3)); S = np.zeros((n, 3)); D = np.zeros((n, 3)); W = np.zeros((n, 3))
    E = 0.0
    for i in range(n):
        for j in range(n):
            r = x[i] - x[j]
            r2 = float(r @ r)
            sg = 0.5 * (s[i] + s[j])
            E += (r2 + 3.2 * sg * sg) / (r2 + sg * sg) ** 1.5 * float(a[i] @ a[j])
            if r2 > 1e-8:
                rho2 = r2 / s[j] ** 2
                W[i] += 7.5 / ((rho2 + 1) ** 3.5 * s[j] ** 3) * a[j]
            if i == j:
                continue
            A = r2 / sg ** 2 + 1
            u[i] -= (A + 1.5) / (A ** 2.5 * sg ** 3) * np.cross(r, a[j])
            S[i] += (A + 1.5) / (sg ** 3 * A ** 2.5) * np.cross(a[i], a[j]) \
                + 3 * (A + 2.5) / (sg ** 5 * A ** 3.5) * float(a[i] @ np.cross(r, a[j])) * r
            D[i] += 105 * nu[i] / (sg ** 2 * A ** 4.5) * (a[j] - a[i])
    return C * u, C * S, C * D, C * W, E / (8 * math.pi) / (4 * math.pi)


def test_brute_force_tiny():
    x, a, s, nu = rand_set(6, 10)
    x[5] = x[4]  # a coincident pair exercises the r^2 > 1e-8 mask
    u, S, D, W, E = _brute(x, a, s, nu)
    assert np.allclose(O.selfinduced_velocity(x, a, s), u, rtol=1e-12, atol=1e-14)
    assert np.allclose(O.strength_gradients(x, a, s), S, rtol=1e-12, atol=1e-14)
    assert np.allclose(O.strength_gradients_pse(x, a, s, nu), D, rtol=1e-12, atol=1e-16)
    assert np.allclose(O.vorticity_field_at(x, x, a, s), W, rtol=1e-12, atol=1e-14)
    assert O.total_enstrophy(x, a, s) == pytest.approx(E, rel=1e-12)


def test_c_port_matches_numpy_oracle():
    from oracle import cpu_baseline as CB

    x, a, s, nu = rand_set(700, 11)
    rel = lambda p, q: np.linalg.norm(p - q) / np.linalg.norm(q)
    assert rel(CB.selfinduced_velocity(x, a, s), O.selfinduced_velocity(x, a, s, "cpu")) < 1e-13
    assert rel(CB.strength_gradients(x, a, s), O.strength_gradients(x, a, s)) < 1e-13
    assert rel(CB.strength_gradients_pse(x, a, s, nu), O.strength_gradients_pse(x, a, s, nu, "cpu")) < 1e-13
    tidx = np.array([3, 500, 699])
    assert rel(CB.strength_gradients(x, a, s, tidx), O.strength_gradients(x, a, s)[tidx]) < 1e-13


def test_c_port_gpu_plugin_forms_match_numpy_oracle():
    """The C port's GPU-plugin sums (used for full-resolution step checks)
    against the NumPy oracle's "cuda" variant, with non-uniform radii and
    viscosities and a coincident pair; and a ReferenceStepper step driven by
    them against one driven by the NumPy oracle (LES, core spreading, RK2)."""
    from oracle import cpu_baseline as CB
    from oracle.vpm_steps import ReferenceStepper

    x, a, s, nu = rand_set(500, 12)
    x[9] = x[3]
    rel = lambda p, q: np.linalg.norm(p - q) / np.linalg.norm(q)
    assert rel(CB.cuda_selfinduced_velocity(x, a, s), O.selfinduced_velocity(x, a, s, "cuda")) < 1e-13
    assert rel(CB.cuda_vorticity(x, a, s), O.vorticity_field_at(x, x, a, s, "cuda")) < 1e-13
    assert rel(CB.cuda_strength_gradients_pse(x, a, s, nu),
               O.strength_gradients_pse(x, a, s, nu, "cuda")) < 1e-13
    assert CB.cuda_total_enstrophy(x, a, s) == pytest.approx(O.total_enstrophy(x, a, s, "cuda"),
                                                             rel=1e-13)
    tidx = np.array([9, 0, 499])
    assert rel(CB.cuda_selfinduced_velocity(x, a, s, tidx),
               O.selfinduced_velocity(x, a, s, "cuda")[tidx]) < 1e-13
    kw = dict(dt=1e-3, method="RK2", flow_model="LES", viscous_scheme="CoreSpreading")
    r0 = ReferenceStepper(x, a, s, nu, **kw)
    r1 = ReferenceStepper(x, a, s, nu, sums=CB.CudaSums, **kw)
    for _ in range(2):
        r0.step()
        r1.step()
    assert r1.cnu == pytest.approx(r0.cnu, rel=1e-13)
    for k in ("x", "alpha", "sigma", "nu_t"):
        assert rel(getattr(r1, k), getattr(r0, k)) < 1e-12, k


def test_row_subsets_equal_the_full_evaluation():
    """The rows= option of the self-evaluations returns exactly those rows."""
    x, a, s, nu = rand_set(300, 21)
    rows = np.array([0, 7, 299, 150, 7])
    for v in ("cuda", "cpu"):
        assert np.array_equal(O.selfinduced_velocity(x, a, s, v, rows=rows),
                              O.selfinduced_velocity(x, a, s, v)[rows])
        assert np.array_equal(O.strength_gradients_pse(x, a, s, nu, v, rows=rows),
                              O.strength_gradients_pse(x, a, s, nu, v)[rows])
    assert np.array_equal(O.strength_gradients(x, a, s, rows=rows),
                          O.strength_gradients(x, a, s)[rows])
