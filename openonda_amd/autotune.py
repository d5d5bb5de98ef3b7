"""Autotuner for the FP32-compensated fused pass.

The library instantiates the fused kernel for a set of tile shapes
(``vpm_tile_config_info``: workgroup size, targets per lane, source tile,
sub-tile length, compensation mode).  :func:`tune` times each candidate on
the device with HIP events, measures its rel-L2 against the FP64 fused pass on
the same input, and keeps the fastest one within the accuracy bound.  Results
are cached in a JSON file keyed by (kernel, precision, output set, radius
mode, N rounded to a power of two); :func:`best_config` reads the cache and
``physics.forward(..., tile_config="auto")`` uses it.
"""
from __future__ import annotations

import json
import math
import os
import threading
import warnings
from pathlib import Path

import torch

from . import physics as P

CACHE = Path(os.environ.get("VPM_AUTOTUNE_CACHE",
                            Path(__file__).resolve().parent / "autotune_cache.json"))
_lock = threading.Lock()


def _key(kernel, precision, outputs, sigma_mode, n):
    bucket = 0 if n <= 1 else int(round(math.log2(n)))
    return f"{kernel}|{precision}|{'+'.join(sorted(outputs))}|{sigma_mode}|2^{bucket}"


def load_cache(path=None):
    p = Path(path or CACHE)
    if not p.exists():
        return {}
    with open(p) as f:
        return json.load(f)


def save_cache(cache, path=None):
    p = Path(path or CACHE)
    tmp = p.with_suffix(".tmp")
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, p)


def best_config(n, outputs=("vel", "stretch"), precision="f32c", sigma_mode="uniform",
                kernel="fused", path=None, default=-1):
    """Cached winner for the nearest power-of-two N, or ``default``."""
    with _lock:
        c = load_cache(path)
    e = c.get(_key(kernel, precision, outputs, sigma_mode, n))
    return int(e["config"]) if e else default


def _time(fn, reps):
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(pos, alpha, sigma, nu=None, outputs=("vel", "stretch"), sigma_mode="uniform",
         max_rel=1e-6, reps=2, path=None, save=True):
    """Times every tile configuration on this input; returns (best, report)
    and (optionally) records the winner in the cache.  When no configuration
    meets ``max_rel``, returns (None, report), caches nothing and warns."""
    P.require_gpu()
    n = pos.shape[1]
    ref = P.forward(pos, alpha, sigma, nu, outputs=outputs, precision="f64",
                    sigma_mode=sigma_mode)
    report = []
    for cfg, shape in enumerate(P.tile_configs()):
        res = {}

        def run():
            res.update(P.forward(pos, alpha, sigma, nu, outputs=outputs, precision="f32c",
                                 sigma_mode=sigma_mode, tile_config=cfg))

        ms = _time(run, reps)
        rel = 0.0
        for k, v in res.items():
            if k == "enstrophy":
                d = abs(float(v) - float(ref[k])) / max(abs(float(ref[k])), 1e-300)
            else:
                d = float((v - ref[k]).norm() / ref[k].norm())
            rel = max(rel, d)
        report.append({"config": cfg, "shape": shape, "ms": ms, "rel_l2": rel})
    ok = [r for r in report if r["rel_l2"] <= max_rel]
    if not ok:
        # no shape meets the accuracy bound on this input: keep the library's
        # default shape (nothing is cached, so tile_config="auto" falls back to it)
        warnings.warn(f"autotune: no tile configuration reached rel-L2 <= {max_rel:g} "
                      f"(best {min(r['rel_l2'] for r in report):.2e}); nothing cached",
                      RuntimeWarning, stacklevel=2)
        return None, report
    best = min(ok, key=lambda r: r["ms"])
    if save:
        with _lock:
            c = load_cache(path)
            c[_key("fused", "f32c", outputs, sigma_mode, n)] = {
                "config": best["config"], "shape": list(best["shape"]), "ms": best["ms"],
                "rel_l2": best["rel_l2"], "n": n, "candidates": report}
            save_cache(c, path)
    return best, report


def device_tune(pos, alpha, sigma, nu=None, outputs=("vel", "stretch"), sigma_mode="uniform",
                reps=2):
    """Timing-only tuning inside the library (``vpm_autotune``): one sort, every
    tile configuration timed with HIP events.  Returns (best index, ms list)."""
    import ctypes as C

    from . import _lib as L

    P.require_gpu()
    res = P.forward(pos, alpha, sigma, nu, outputs=outputs, precision="f32c",
                    sigma_mode=sigma_mode)  # output buffers of the right shape
    p = P._particles(pos, alpha, sigma, nu if "pse" in outputs else None)
    mask = 0
    for k in outputs:
        mask |= P.OUTPUTS[k]
    opts = L.ForwardOpts(mask, L.PREC_F32C,
                         L.SIGMA_UNIFORM if sigma_mode == "uniform" else L.SIGMA_GENERAL, 0, -1)
    vec = [k for k in ("vel", "stretch", "pse", "vort") if k in res]
    ld = P._ld(res[vec[0]]) if vec else max(p.n, 1)
    fo = L.ForwardOut(ld, *[res[k].data_ptr() if k in res else None
                            for k in ("vel", "stretch", "pse", "vort", "enstrophy")])
    lib = L.load()
    nb = lib.vpm_autotune_workspace_bytes(p.n, C.byref(opts))
    ws = P.workspace(pos.device, nb)
    count = lib.vpm_tile_config_count()
    ms = (C.c_float * count)()
    best = C.c_int(-1)
    L.call("vpm_autotune", C.byref(p), C.byref(opts), C.byref(fo), ws.data_ptr(), ws.numel(),
           P._stream(pos), int(reps), C.byref(best), ms)
    return best.value, [float(v) for v in ms]


def main(argv=None):
    """python -m openonda_amd.autotune --n 1048576 --outputs vel,stretch [--layout cloud|ring]
    [--cache PATH]"""
    import argparse

    import numpy as np

    from . import flow_models as FM

    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--outputs", default="vel,stretch")
    ap.add_argument("--cache", default=None)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--layout", choices=["cloud", "ring"], default="cloud",
                    help="random cloud, or a torus ring of n/64 stations")
    a = ap.parse_args(argv)
    outs = tuple(a.outputs.split(","))
    if a.layout == "ring":
        x, al, s, nu = FM.torus_ring(max(1, a.n // 64))
        a.n = len(x)
    else:
        x, al, s, nu = FM.random_cloud(a.n)
    dev = torch.device("cuda", 0)
    best, report = tune(P.to_soa(x, dev), P.to_soa(al, dev), torch.from_numpy(s).to(dev),
                        torch.from_numpy(np.full(a.n, 1e-3)).to(dev) if "pse" in outs else None,
                        outputs=outs, reps=a.reps, path=a.cache)
    for r in report:
        print(f"config {r['config']} {tuple(r['shape'])}: {r['ms']:.2f} ms rel-L2 {r['rel_l2']:.2e}",
              flush=True)
    if best is None:
        print("no configuration within the accuracy bound; nothing cached", flush=True)
    else:
        print(f"best: config {best['config']} ({best['ms']:.2f} ms)", flush=True)


if __name__ == "__main__":
    main()
