"""vpm_forward_noalloc (the allocation-free forward) against vpm_forward.

Both are called through the C ABI on the same particles and options; the
allocation-free entry point runs in the library's static device array, two
calls in a row, the second with fewer particles (the array keeps the first
call's contents).  Sizes where both choose the same source splits give
bit-identical outputs; the largest case (2^18 particles) may split less in
the static array and is held to rel-L2 1e-12 (FP64) / 1e-6 (FP32c).  A set
too large for the array is refused with VPM_ERR_WORKSPACE.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from openonda_amd import _lib as L
from openonda_amd import flow_models as FM
from openonda_amd import physics as P

pytestmark = pytest.mark.gpu

MASKS = {"vs": L.OUT_VEL | L.OUT_STRETCH, "vsp": L.OUT_VEL | L.OUT_STRETCH | L.OUT_PSE,
         "all": 31}


def _outs(n, dev):
    t = {k: torch.full((3, n), np.nan, dtype=torch.float64, device=dev)
         for k in ("vel", "stretch", "pse", "vort")}
    t["enstrophy"] = torch.full((1,), np.nan, dtype=torch.float64, device=dev)
    fo = L.ForwardOut(n, *[t[k].data_ptr() for k in ("vel", "stretch", "pse", "vort", "enstrophy")])
    return t, fo


def _run(entry, x, a, s, nu, mask, prec, sigma_mode):
    dev = torch.device("cuda", 0)
    pos, al = P.to_soa(x, dev), P.to_soa(a, dev)
    sig, vis = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
    part = P._particles(pos, al, sig, vis)
    opts = L.ForwardOpts(mask, prec, sigma_mode, 0, -1)
    t, fo = _outs(len(x), dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    if entry == "noalloc":
        L.call("vpm_forward_noalloc", C.byref(part), C.byref(opts), C.byref(fo), stream)
    else:
        lib = L.load()
        nb = lib.vpm_forward_workspace_bytes(len(x), C.byref(opts))
        ws = torch.empty(max(nb, 256), dtype=torch.uint8, device=dev)
        L.call("vpm_forward", C.byref(part), C.byref(opts), C.byref(fo), ws.data_ptr(), nb, stream)
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in t.items()}


def _viscous(case, nu=1e-3):
    x, a, s, _ = case
    return x, a, s, np.full(len(x), nu)


def _sets(kind):
    """Two particle sets per kind, the second smaller (the clouds get a
    viscosity so that the PSE sum is not identically zero)."""
    if kind == "ring":
        return [FM.torus_ring(40, nu=1e-3), FM.torus_ring(9, nu=1e-3)]  # 2560, then 576
    if kind == "cloud":
        return [_viscous(FM.random_cloud(5000)), _viscous(FM.random_cloud(1200))]
    return [_viscous(FM.random_cloud(1 << 18)), _viscous(FM.random_cloud(1 << 12))]


@pytest.mark.parametrize("kind", ["ring", "cloud", "large"])
@pytest.mark.parametrize("outs", ["vs", "vsp", "all"])
@pytest.mark.parametrize("prec", [L.PREC_F32C, L.PREC_F64])
def test_noalloc_matches_forward_twice(gpu, kind, outs, prec):
    mask = MASKS[outs]
    for k, (x, a, s, nu) in enumerate(_sets(kind)):
        print(f"[noalloc] {kind} n={len(x)} mask={mask} prec={prec} call {k}", flush=True)
        want = _run("forward", x, a, s, nu, mask, prec, L.SIGMA_UNIFORM)
        got = _run("noalloc", x, a, s, nu, mask, prec, L.SIGMA_UNIFORM)
        keys = [q for q, b in (("vel", 1), ("stretch", 2), ("pse", 4), ("vort", 8),
                               ("enstrophy", 16)) if mask & b]
        for q in keys:
            if kind != "large":
                assert np.array_equal(got[q], want[q]), (q, k)
            else:
                tol = 1e-12 if prec == L.PREC_F64 else 1e-6
                ref = np.linalg.norm(want[q])
                assert ref > 0, (q, k)
                e = np.linalg.norm(got[q] - want[q]) / ref
                assert e <= tol, (q, k, e)


def test_noalloc_general_sigma_and_refusal(gpu):
    rng = np.random.default_rng(7)
    x, a, s, nu = FM.random_cloud(3000)
    s = s * (1 + 0.3 * rng.random(len(s)))
    for prec in (L.PREC_F32C, L.PREC_F64):
        want = _run("forward", x, a, s, nu, MASKS["vsp"], prec, L.SIGMA_GENERAL)
        got = _run("noalloc", x, a, s, nu, MASKS["vsp"], prec, L.SIGMA_GENERAL)
        for q in ("vel", "stretch", "pse"):
            assert np.array_equal(got[q], want[q]), q
    cap = L.load().vpm_forward_noalloc_capacity()
    assert cap == 256 << 20
    # a set whose workspace exceeds the static array is refused before any launch
    n = 4 << 20
    dev = torch.device("cuda", 0)
    pos = torch.zeros((3, n), dtype=torch.float64, device=dev)
    sig = torch.full((n,), 0.1, dtype=torch.float64, device=dev)
    part = P._particles(pos, pos, sig)
    opts = L.ForwardOpts(MASKS["vs"], L.PREC_F32C, L.SIGMA_UNIFORM, 0, -1)
    fo = L.ForwardOut(n, pos.data_ptr(), pos.data_ptr(), None, None, None)
    with pytest.raises(L.VpmError) as e:
        L.call("vpm_forward_noalloc", C.byref(part), C.byref(opts), C.byref(fo),
               torch.cuda.current_stream(dev).cuda_stream)
    assert e.value.code == L.VPM_ERR_WORKSPACE
