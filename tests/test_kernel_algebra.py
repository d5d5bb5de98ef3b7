"""The algebraic rewrites the fused kernel relies on, checked in FP64 on the
CPU against the reference's own forms (no GPU).

* Powers of t = A^-1/2 with A = rho^2 + 1 (one rsqrt per pair): the
  reference's (rho^2 + 2.5)/A^2.5, 3 (rho^2 + 3.5)/A^3.5, (rho^2 + 3.2)/A^1.5
  and A^-4.5 (ParticlesPhysicsCuda.py:159-191, :396-440, :500-525, :886-909)
  equal t^3 + 1.5 t^5, t^5 (3 + 7.5 t^2), t + 2.2 t^3 and t^5 t^2 t^2 -- the
  forms and constants (1.5, 7.5, 3, 2.2) of the pair loop.
* The far-field moment form: with d = x_i - x_j relative to any origin,
  sum_j g d x a_j = x_i x G + sum_j g (a_j x x_j) with G = sum_j g a_j, and
  a_i . (d x a_j) = a_j . (a_i x x_i) + a_i . (a_j x x_j) -- the signs of the
  staged moments m_j = a_j x x_j and of p_i = a_i x x_i.
* Splitting the sources into near tiles (direct form) and far tiles (moment
  form, folded per sub-tile) and adding the parts reproduces the direct sums
  for velocity and stretching, whatever the order of near and far tiles.
"""
import numpy as np


def test_power_forms_match_the_reference_kernels():
    rho2 = np.concatenate([[0.0], np.geomspace(1e-6, 1e4, 400)])
    A = rho2 + 1.0
    t = A ** -0.5
    t2, t3 = t * t, t * t * t
    t5 = t3 * t2
    np.testing.assert_allclose(t3 + 1.5 * t5, (rho2 + 2.5) / A ** 2.5, rtol=1e-13)
    np.testing.assert_allclose(t5 * (3.0 + 7.5 * t2), 3.0 * (rho2 + 3.5) / A ** 3.5, rtol=1e-13)
    np.testing.assert_allclose(t + 2.2 * t3, (rho2 + 3.2) / A ** 1.5, rtol=1e-13)
    np.testing.assert_allclose(t5 * t2 * t2, A ** -4.5, rtol=1e-13)
    np.testing.assert_allclose(t5 * t2, A ** -3.5, rtol=1e-13)


def test_moment_identities_hold_for_any_origin():
    rng = np.random.default_rng(5)
    for origin in (np.zeros(3), rng.normal(size=3) * 10):
        xi, xj = rng.normal(size=(2, 3)) + origin
        ai, aj = rng.normal(size=(2, 3))
        d = xi - xj
        xi0, xj0 = xi - origin, xj - origin
        # the cross product of the velocity
        np.testing.assert_allclose(np.cross(d, aj), np.cross(xi0, aj) + np.cross(aj, xj0),
                                   atol=1e-12)
        # the triple product of the stretching, from p_i = a_i x x_i, m_j = a_j x x_j
        p, m = np.cross(ai, xi0), np.cross(aj, xj0)
        np.testing.assert_allclose(ai @ np.cross(d, aj), aj @ p + ai @ m, atol=1e-12)


def _g_f2(d):
    r2 = (d * d).sum(-1)
    A = r2 + 1.0
    return (r2 + 2.5) / A ** 2.5, 3.0 * (r2 + 3.5) / A ** 3.5


def test_near_far_split_with_sub_tile_folds_reproduces_the_direct_sums():
    rng = np.random.default_rng(11)
    n, tile, sub = 96, 16, 4
    x = rng.uniform(-6, 6, (n, 3))
    a = rng.normal(size=(n, 3))
    c = rng.normal(size=3)  # the workgroup centre: any origin
    for i in (0, 17, 95):
        d = x[i] - x
        g, f2 = _g_f2(d)
        g[i] = f2[i] = 0.0
        cc = np.cross(d, a)
        u_ref = (g[:, None] * cc).sum(0)
        e = cc @ a[i]
        S_ref = np.cross(a[i], (g[:, None] * a).sum(0)) + ((f2 * e)[:, None] * d).sum(0)
        # tiles alternate near / far in an order that puts near tiles before far ones
        far_tiles = {t for t in range(n // tile) if t % 3 != 1}
        xi0 = x[i] - c
        p = np.cross(a[i], xi0)
        m = np.cross(a, x - c)
        u = np.zeros(3)
        G = np.zeros(3)
        Sd = np.zeros(3)
        for t0 in range(0, n, tile):
            if t0 // tile in far_tiles:
                for s0 in range(t0, t0 + tile, sub):
                    js = slice(s0, s0 + sub)
                    Gs = (g[js, None] * a[js]).sum(0)
                    Hs = (g[js, None] * m[js]).sum(0)
                    u += np.cross(xi0, Gs) + Hs  # the fold sees this sub-tile's G only
                    G += Gs
                    e1 = a[js] @ p + m[js] @ a[i]
                    Sd += ((f2[js] * e1)[:, None] * d[js]).sum(0)
            else:
                js = slice(t0, t0 + tile)
                u += (g[js, None] * cc[js]).sum(0)
                G += (g[js, None] * a[js]).sum(0)
                Sd += ((f2[js] * e[js])[:, None] * d[js]).sum(0)
        np.testing.assert_allclose(u, u_ref, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(np.cross(a[i], G) + Sd, S_ref, rtol=1e-12, atol=1e-14)
