"""GPU parity of the fused forward pass (FP64 and FP32-compensated).

The FP64 fused pass shares one rsqrt per pair instead of the reference's pow
calls, so it matches the oracle to rel-L2 <= 1e-12.  The FP32-compensated pass
is held to rel-L2 <= 1e-6 per output (the accuracy bar of the fast path), on
both the uniform-radius fast kernels and the general sigma_ij kernels.
"""
import numpy as np
import pytest
import torch

from oracle import vpm_oracle as O
from openonda_amd import flow_models as FM
from openonda_amd import physics as P

pytestmark = pytest.mark.gpu

F64_TOL = 1e-12
F32C_TOL = 1e-6


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(np.asarray(a) - b) / (nb if nb > 0 else 1.0)


def to_dev(x, a, s, nu):
    return (P.to_soa(x), P.to_soa(a), torch.from_numpy(np.ascontiguousarray(s)).cuda(),
            torch.from_numpy(np.ascontiguousarray(nu)).cuda())


def oracle_all(x, a, s, nu):
    return {
        "vel": O.selfinduced_velocity(x, a, s),
        "stretch": O.strength_gradients(x, a, s),
        "pse": O.strength_gradients_pse(x, a, s, nu),
        "vort": O.vorticity_field_at(x, x, a, s),
        "enstrophy": O.total_enstrophy(x, a, s),
    }


def check(res, ref, tol, tag):
    for k, v in res.items():
        if k == "enstrophy":
            got = float(v.item())
            assert abs(got - ref[k]) <= 10 * tol * abs(ref[k]), (tag, k, got, ref[k])
        else:
            err = rel_l2(P.from_soa(v), ref[k])
            assert err <= tol, (tag, k, err)


def ring_set(n_st=32, nu=1e-3):
    return FM.torus_ring(n_st, nu=nu)


_cache = {}


def cached(key, fn):
    if key not in _cache:
        _cache[key] = fn()
    return _cache[key]


def nonuniform_set(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.0, 1.0, (n, 3))
    a = rng.normal(0.0, 1.0 / n, (n, 3))
    s = rng.uniform(0.1, 0.3, n)
    nu = rng.uniform(1e-4, 1e-2, n)
    return x, a, s, nu


ALL = ("vel", "stretch", "pse", "vort", "enstrophy")
SUBSETS = [("vel",), ("stretch",), ("vel", "stretch"), ("pse",), ("vel", "stretch", "pse"),
           ("vort",), ("enstrophy",), ("vort", "enstrophy"),
           ("vel", "stretch", "vort", "enstrophy"), ALL, ("vel", "pse"), ("stretch", "vort")]


@pytest.mark.parametrize("precision,tol", [("f64", F64_TOL), ("f32c", F32C_TOL)])
@pytest.mark.parametrize("outputs", SUBSETS)
def test_ring_uniform(gpu, precision, tol, outputs):
    x, a, s, nu = ring_set()
    print(f"[fused-ring] {precision} {outputs} n={len(x)}", flush=True)
    ref = cached("ring", lambda: oracle_all(x, a, s, nu))
    res = P.forward(*to_dev(x, a, s, nu), outputs=outputs, precision=precision,
                    sigma_mode="uniform")
    assert set(res) == set(outputs)
    check(res, ref, tol, ("ring", precision))


@pytest.mark.parametrize("precision,tol", [("f64", F64_TOL), ("f32c", F32C_TOL)])
@pytest.mark.parametrize("n,seed", [(1, 0), (5, 1), (700, 2), (3000, 3)])
def test_nonuniform_sigma(gpu, precision, tol, n, seed):
    x, a, s, nu = nonuniform_set(n, seed)
    print(f"[fused-general] {precision} n={n}", flush=True)
    ref = cached(("general", n, seed), lambda: oracle_all(x, a, s, nu))
    res = P.forward(*to_dev(x, a, s, nu), outputs=ALL, precision=precision, sigma_mode="auto")
    check(res, ref, tol, ("general", precision, n))


@pytest.mark.parametrize("cfg", range(12))
def test_tile_configs(gpu, cfg):
    """Every autotune candidate produces the same answer on a cloud."""
    x, a, s, nu = FM.random_cloud(4096 + 77, seed=11)
    nu = np.full(len(x), 1e-3)
    print(f"[fused-cfg] {cfg}", flush=True)
    ref = cached("cloud4173", lambda: oracle_all(x, a, s, nu))
    for outs in [("vel", "stretch"), ("vel", "stretch", "pse"), ("vel", "stretch", "vort", "enstrophy")]:
        res = P.forward(*to_dev(x, a, s, nu), outputs=outs, precision="f32c",
                        sigma_mode="uniform", tile_config=cfg)
        check(res, ref, F32C_TOL, ("cfg", cfg, outs))


def test_cloud_f32c_65536_subsample(gpu):
    """Random cloud N = 2^16: FP32-compensated vs the oracle on 512 strided targets."""
    x, a, s, nu = FM.random_cloud(1 << 16)
    print("[fused-cloud-65536]", flush=True)
    res = P.forward(*to_dev(x, a, s, nu), outputs=("vel", "stretch"), precision="f32c",
                    sigma_mode="uniform")
    idx = np.arange(0, len(x), 128)
    # The oracle's self-interaction exclusion is index-based: evaluate the
    # targets as an external set against all sources (the excluded self pair
    # contributes exactly zero to both outputs anyway).
    u_ref = O.induced_velocity_at(x[idx], x, a, s)
    u = P.from_soa(res["vel"])[idx]
    assert rel_l2(u, u_ref) <= F32C_TOL


@pytest.mark.parametrize("precision,tol", [("f64", 1e-13), ("f32c", 2e-6)])
def test_presorted_equals_sorted(gpu, precision, tol):
    """presorted=True skips the Morton sort (the caller's order: station by
    station along the ring, compact blocks, near and far tiles)."""
    x, a, s, nu = ring_set(16)
    d = to_dev(x, a, s, nu)
    outs = ("vel", "stretch", "pse")
    r1 = P.forward(*d, outputs=outs, precision=precision, sigma_mode="uniform")
    r2 = P.forward(*d, outputs=outs, precision=precision, sigma_mode="uniform", presorted=True)
    for k in r1:
        assert rel_l2(P.from_soa(r1[k]), P.from_soa(r2[k])) <= tol, k


def test_empty_forward(gpu):
    z = torch.zeros((3, 0), dtype=torch.float64, device="cuda")
    e = torch.zeros((0,), dtype=torch.float64, device="cuda")
    res = P.forward(z, z, e, e, outputs=ALL, precision="f32c", sigma_mode="general")
    assert res["vel"].shape == (3, 0)
    assert float(res["enstrophy"].item()) == 0.0


@pytest.mark.parametrize("precision", ["f64", "f32c"])
@pytest.mark.parametrize("sigma_mode", ["uniform", "general"])
def test_one_particle_forward(gpu, precision, sigma_mode):
    """A lone particle: the reference sums over j != i, so velocity,
    stretching and PSE are exactly zero; vorticity and enstrophy keep the
    i = j term."""
    x = np.array([[0.3, -1.2, 2.5]])
    a = np.array([[0.7, -0.4, 1.1]])
    s, nu = np.array([0.17]), np.array([2e-3])
    print(f"[fused] one particle {precision} {sigma_mode}", flush=True)
    res = P.forward(*to_dev(x, a, s, nu), outputs=ALL, precision=precision,
                    sigma_mode=sigma_mode, max_splits=1)
    for k in ("vel", "stretch", "pse"):
        assert torch.count_nonzero(res[k]) == 0, (k, res[k])
    ref = oracle_all(x, a, s, nu)
    tol = F64_TOL if precision == "f64" else F32C_TOL
    assert rel_l2(P.from_soa(res["vort"]), ref["vort"]) <= tol
    assert float(res["enstrophy"].item()) == pytest.approx(ref["enstrophy"], rel=10 * tol)


@pytest.mark.parametrize("precision,tol", [("f64", F64_TOL), ("f32c", F32C_TOL)])
@pytest.mark.parametrize("sigma_mode", ["uniform", "general"])
def test_coincident_particles(gpu, precision, tol, sigma_mode):
    """Distinct particles at one position: the reference sums over j != i, so
    such pairs count (d = 0: nothing in the velocity or the stretching's
    r-term, but g(0) alpha_i x alpha_j in the stretching and the full PSE
    weight; the vorticity's r^2 > 1e-8 test drops them).  Only the particle's
    own index is left out, not its position."""
    rng = np.random.default_rng(21)
    n0, nd = 300, 40
    x = rng.uniform(-1.0, 1.0, (n0, 3))
    x = np.concatenate([x, x[rng.choice(n0, nd, replace=False)]])
    n = len(x)
    a = rng.normal(0.0, 1.0, (n, 3)) / n
    s = np.full(n, 0.15) if sigma_mode == "uniform" else rng.uniform(0.1, 0.2, n)
    nu = rng.uniform(1e-3, 1e-2, n)
    print(f"[fused] coincident {precision} {sigma_mode}", flush=True)
    res = P.forward(*to_dev(x, a, s, nu), outputs=ALL, precision=precision,
                    sigma_mode=sigma_mode)
    check(res, oracle_all(x, a, s, nu), tol, ("coincident", precision, sigma_mode))


def test_bad_arguments(gpu):
    x, a, s, nu = to_dev(*ring_set(4))
    with pytest.raises(ValueError):
        P.forward(x, a, s, None, outputs=("pse",))
    with pytest.raises(ValueError):
        P.forward(x, a, s, nu, outputs=("bogus",))
    with pytest.raises(TypeError):
        P.forward(x.float(), a, s, nu)
    from openonda_amd import _lib as L
    with pytest.raises(L.VpmError):
        P.forward(x, a, s, nu, tile_config=99)


def test_autotune_records_a_valid_winner(gpu, tmp_path):
    from openonda_amd import autotune as AT

    x, a, s, nu = FM.random_cloud(1 << 14, seed=12)
    d = to_dev(x, a, s, nu)
    path = tmp_path / "cache.json"
    best, report = AT.tune(d[0], d[1], d[2], outputs=("vel", "stretch"), path=path, reps=1)
    assert len(report) == len(P.tile_configs())
    assert all(r["rel_l2"] <= F32C_TOL for r in report)
    assert AT.best_config(1 << 14, ("vel", "stretch"), path=path) == best["config"]
    assert AT.best_config(1 << 20, ("vel", "stretch"), path=path) == -1


SPLIT_CASES = [
    # n, precision, sigma_mode, outputs
    (8192, "f32c", "uniform", ("vel", "stretch")),
    (8192, "f32c", "uniform", ("vel", "stretch", "pse", "vort", "enstrophy")),
    (4096, "f32c", "general", ("vel", "stretch", "pse", "vort", "enstrophy")),
    (4096, "f64", "uniform", ("vel", "stretch", "pse")),
    (3000, "f64", "general", ("vel", "stretch", "pse", "vort", "enstrophy")),
]


@pytest.mark.parametrize("n,precision,mode,outputs", SPLIT_CASES)
def test_source_split_matches_single_pass(gpu, n, precision, mode, outputs):
    """Sources split among workgroups (small N) give the unsplit result."""
    print(f"[split] n={n} {precision} {mode} {outputs}", flush=True)
    x, a, s, nu = FM.random_cloud(n)
    nu = np.full(n, 1e-3)
    if mode == "general":
        s = s * (1.0 + 0.2 * np.random.default_rng(n).random(n))
    d = to_dev(x, a, s, nu)
    auto = P.forward(*d, outputs=outputs, precision=precision, sigma_mode=mode)
    single = P.forward(*d, outputs=outputs, precision=precision, sigma_mode=mode, max_splits=1)
    two = P.forward(*d, outputs=outputs, precision=precision, sigma_mode=mode, max_splits=2)
    tol = 1e-13 if precision == "f64" else 2e-7
    for k in outputs:
        ref = single[k].cpu().numpy()
        for name, res in (("auto", auto), ("two", two)):
            got = res[k].cpu().numpy()
            err = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)
            assert err <= tol, (name, k, err)
    ref = oracle_all(x, a, s, nu)
    check({k: auto[k] for k in outputs}, ref, F64_TOL if precision == "f64" else F32C_TOL, "auto")


@pytest.mark.parametrize("outputs", [("vel", "stretch"), ("vel", "stretch", "pse"),
                                     ("vel", "stretch", "pse", "vort", "enstrophy")])
def test_g64_modes_on_ring_and_cloud(gpu, outputs):
    """The stretching's alpha_j sum: FP64 per pair in coherent waves (a ring),
    FP32 partials in incoherent ones (a random cloud); both within the bar,
    for velocity + stretching alone (with the far-field moment form), with
    PSE, and with PSE and the LES outputs."""
    xr, ar, sr, nur = ring_set()
    xc, ac, sc, _ = FM.random_cloud(8192, seed=17)
    nuc = np.full(len(xc), 1e-3)
    ref_r = {k: v for k, v in cached("ring", lambda: oracle_all(xr, ar, sr, nur)).items()
             if k in outputs}
    ref_c = {k: v for k, v in cached("cloud8192", lambda: oracle_all(xc, ac, sc, nuc)).items()
             if k in outputs}
    runs = {}
    for name, (x, a, s, nu) in (("ring", (xr, ar, sr, nur)), ("cloud", (xc, ac, sc, nuc))):
        d = to_dev(x, a, s, nu)
        for mode in ("auto", "always", "never"):
            print(f"[g64] {name} {mode}", flush=True)
            runs[name, mode] = P.forward(*d, outputs=outputs, precision="f32c",
                                         sigma_mode="uniform", g64=mode)
    # a ring's waves are coherent: automatic = FP64 per pair, bit for bit
    for k in outputs:
        assert torch.equal(runs["ring", "auto"][k], runs["ring", "always"][k])
        assert torch.equal(runs["cloud", "auto"][k], runs["cloud", "never"][k])
    check(runs["ring", "auto"], ref_r, F32C_TOL, "ring auto")
    check(runs["cloud", "auto"], ref_c, F32C_TOL, "cloud auto")
    check(runs["cloud", "always"], ref_c, F32C_TOL, "cloud always")
    # FP32 partials on the ring lose accuracy: the reason for the FP64 path
    err_never = rel_l2(P.from_soa(runs["ring", "never"]["stretch"]), ref_r["stretch"])
    err_auto = rel_l2(P.from_soa(runs["ring", "auto"]["stretch"]), ref_r["stretch"])
    assert err_never > 2 * err_auto
    with pytest.raises(KeyError):
        P.forward(*to_dev(xr, ar, sr, nur), outputs=outputs, g64="sometimes")


def test_pse_accuracy_on_nearly_uniform_strengths(gpu):
    """PSE sums eta (alpha_j - alpha_i); on a nearly uniform field the
    differences are 1e-3 of the strengths.  Taking them from high and low
    FP32 parts keeps the FP32 pass within the bar, where one rounded FP32
    strength per particle would leave ~1e-4."""
    x, _, s, _ = FM.random_cloud(4096, seed=31)
    rng = np.random.default_rng(32)
    a = np.array([1.0, -0.5, 0.25]) / 4096 * (1.0 + 1e-3 * rng.standard_normal((4096, 3)))
    nu = np.full(4096, 1e-3)
    print("[pse-uniform] n=4096", flush=True)
    ref = O.strength_gradients_pse(x, a, s, nu)
    res = P.forward(*to_dev(x, a, s, nu), outputs=("vel", "stretch", "pse"), precision="f32c",
                    sigma_mode="uniform")
    assert rel_l2(P.from_soa(res["pse"]), ref) <= F32C_TOL


def test_library_autotune_entry_point(gpu):
    """vpm_autotune times every configuration in one call and names the fastest."""
    from openonda_amd import autotune

    x, a, s, nu = FM.random_cloud(16384, seed=12)
    d = to_dev(x, a, s, nu)
    print("[vpm_autotune] n=16384", flush=True)
    best, ms = autotune.device_tune(d[0], d[1], d[2], outputs=("vel", "stretch"), reps=1)
    assert len(ms) == len(P.tile_configs()) and all(t > 0 for t in ms)
    assert 0 <= best < len(ms) and ms[best] == min(ms)


def test_thin_ring_stretching_with_fp64_triple_product(gpu):
    """A thin ring (1978 stations, core 3h ~ 0.01, sigma = 1.3h) has
    |S| ~ 2e-3 |alpha||u|: the stretching is a near-cancellation.  Coherent
    waves take the triple product in FP64 on near tiles (4.8e-7 on the GPU;
    FP32 strengths gave ~4e-6, their low parts in FP32 ~1e-6)."""
    x, a, s, nu = FM.torus_ring(1978, h=2 * np.pi / 1978)
    print(f"[thin-ring] n={len(x)}", flush=True)
    d = to_dev(x, a, s, nu)
    res = P.forward(*d, outputs=("vel", "stretch"), precision="f32c", sigma_mode="uniform")
    u = P.selfinduced_velocity(d[0], d[1], d[2])
    st = P.strength_gradients(d[0], d[1], d[2])
    assert float((res["vel"] - u).norm() / u.norm()) <= 1e-6
    assert float((res["stretch"] - st).norm() / st.norm()) <= 1e-6


def test_mixed_wave_modes_share_the_staged_tile(gpu):
    """A ring of 256 stations inside a random cloud with the same radius:
    sorted blocks where the two meet hold waves of both G modes (FP64 per
    pair and FP32 partials with the far-field moment form), which share each
    staged source tile.  Automatic mode must differ from both forced modes
    (so the blocks do mix) and stay within the bar."""
    xr, ar, sr, _ = FM.torus_ring(256, h=2 * np.pi / 256)
    rng = np.random.default_rng(29)
    nc = 3 * len(xr)
    xc = np.stack([rng.uniform(-0.3, 0.3, nc), rng.uniform(-1.3, 1.3, nc),
                   rng.uniform(-1.3, 1.3, nc)], -1)
    ac = rng.normal(0.0, np.abs(ar).max() / 3, (nc, 3))
    x = np.concatenate([xr, xc])
    a = np.concatenate([ar, ac])
    s = np.full(len(x), sr[0])
    print(f"[mixed] n={len(x)}", flush=True)
    d = to_dev(x, a, s, np.zeros(len(x)))
    runs = {m: P.forward(*d, outputs=("vel", "stretch"), precision="f32c", sigma_mode="uniform",
                         g64=m) for m in ("auto", "always", "never")}
    u = P.selfinduced_velocity(d[0], d[1], d[2])
    st = P.strength_gradients(d[0], d[1], d[2])
    assert not torch.equal(runs["auto"]["stretch"], runs["always"]["stretch"])
    assert not torch.equal(runs["auto"]["stretch"], runs["never"]["stretch"])
    for m, res in runs.items():
        assert torch.isfinite(res["vel"]).all() and torch.isfinite(res["stretch"]).all(), m
    assert float((runs["auto"]["vel"] - u).norm() / u.norm()) <= 1e-6
    assert float((runs["auto"]["stretch"] - st).norm() / st.norm()) <= 1e-6


@pytest.mark.parametrize("outputs", [("vel", "stretch"), ("vel", "stretch", "pse"),
                                     ("vel", "stretch", "pse", "vort", "enstrophy")])
def test_sparse_cloud_keeps_the_self_pair_out_of_g(gpu, outputs):
    """Radius 0.05 against a spacing of ~0.3: the alpha_j sum G of each target
    is dominated by its own pair (2.5 alpha_i), which alpha_i x G cancels.
    The FP32-partial waves must keep that term out of their partials (the
    self-pair mask); a few hundred particles span several waves, sources
    tiles and both radius-mode kernels.  Velocity and stretching are checked
    (the outputs G enters); the other outputs ride along so that each output
    set's kernel runs.  (PSE on such a set is a sum of one or two t^9 terms
    and carries their FP32 roundings unaveraged: ~3e-6 with per-particle
    radii.)"""
    rng = np.random.default_rng(31)
    for n in (37, 300):
        x = rng.uniform(-1.0, 1.0, (n, 3))
        a = rng.normal(0.0, 0.03, (n, 3))
        s = np.full(n, 0.05)
        nu = np.full(n, 1e-3)
        ref = oracle_all(x, a, s, nu)
        ref = {k: v for k, v in ref.items() if k in outputs}
        for mode in ("uniform", "general"):
            print(f"[sparse] n={n} {mode} {outputs}", flush=True)
            res = P.forward(*to_dev(x, a, s, nu), outputs=outputs, precision="f32c",
                            sigma_mode=mode, g64="never")
            check({k: res[k] for k in ("vel", "stretch")}, ref, F32C_TOL,
                  f"sparse n={n} {mode}")


@pytest.mark.parametrize("layout", ["ring", "cloud"])
@pytest.mark.parametrize("outputs", [("vel", "stretch"), ("vel", "stretch", "pse"),
                                     ("vel", "stretch", "pse", "vort", "enstrophy")])
@pytest.mark.parametrize("precision,tol", [("f64", F64_TOL), ("f32c", F32C_TOL)])
def test_single_pass_crosses_near_and_far_tiles(gpu, precision, tol, outputs, layout):
    """One pass over all source tiles (no source split): every workgroup meets
    far tiles (moment form) after near ones, so each far sub-tile's fold must
    see only its own alpha_j sum, in the FP64 pass as in the FP32 one, in
    coherent (ring) and unrelated (cloud) waves."""
    if layout == "ring":
        x, a, s, nu = ring_set(64)
    else:
        x, a, s, _ = FM.random_cloud(4096, seed=41)
        nu = np.full(len(x), 1e-3)
    ref = cached(f"{layout}-single", lambda: oracle_all(x, a, s, nu))
    print(f"[single-pass] {layout} n={len(x)} {precision} {outputs}", flush=True)
    res = P.forward(*to_dev(x, a, s, nu), outputs=outputs, precision=precision,
                    sigma_mode="uniform", max_splits=1)
    check(res, {k: ref[k] for k in outputs}, tol, ("single-pass", layout, precision))


@pytest.mark.parametrize("outputs", [("vel", "stretch"), ("vel", "stretch", "pse"), ALL])
def test_far_test_windows_and_partial_chunks(gpu, outputs):
    """The far tests run per window of 64 chunks (one lane per chunk, two
    ballots).  A cloud of 3 * 64 * 128 + 77 sources spans four windows of
    the 128-source chunks (the last chunk partial) and, at the 64-source
    chunks of the PSE shape, seven; split four ways its source ranges start
    inside windows.  Every target of the FP32 pass, unsplit and split,
    against the FP64 pass (rel-L2 <= 1e-6); the output sets with PSE take
    the far weight from t^7 and records holding 3 m_j, all five outputs also
    the vorticity's weight."""
    n = 3 * 64 * 128 + 77
    x, a, s, nu = FM.random_cloud(n, seed=19)
    nu = np.full(n, 1e-3)
    print(f"[far-windows] n={n} {outputs}", flush=True)
    d = to_dev(x, a, s, nu)
    ref = P.forward(*d, outputs=outputs, precision="f64", sigma_mode="uniform")
    for splits in (1, 4):
        res = P.forward(*d, outputs=outputs, precision="f32c", sigma_mode="uniform",
                        max_splits=splits)
        for k in outputs:
            err = rel_l2(res[k].cpu().numpy(), ref[k].cpu().numpy())
            assert err <= F32C_TOL, (splits, k, err)
