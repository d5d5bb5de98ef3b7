"""Autotuner for the FP32-compensated fused pass.

The library instantiates the fused kernel for a set of tile shapes
(``vpm_tile_config_info``: workgroup size, targets per lane, source tile,
sub-tile length, compensation mode).  :func:`tune` times each candidate on
the device with HIP events, measures its rel-L2 against the FP64 fused pass on
the same input, and keeps the fastest one within the accuracy bound.  Results
are cached in a JSON file keyed by (kernel, precision, output set, radius
mode, N rounded to a power of two); :func:`best_config` reads the cache and
``physics.forward(..., tile_config="auto")`` uses it.

:func:`tune_batch` does the same for the batched persistent launch over
(tile shape x persistent grid size), keyed by output mask and total pair
count; ``BatchedSystem(tile_config="auto")`` uses its choice.
"""
from __future__ import annotations

import json
import math
import os
import threading
import warnings
from pathlib import Path

import numpy as np
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


# ---------------------------------------------------------------- batch
# The sweep over every (tile shape x grid) candidate is limited to batches
# of this many particles: run over the 96-set workload (3.3e6 particles) it
# faulted the GPU twice.  The cause found in the code -- the kernels without
# far records wrote per-wave target boxes past their one-row LDS array,
# where other waves of the batch kernel could read the shared work-item
# index -- is fixed, but the full-size sweep has not been run since.
TUNE_BATCH_MAX_PARTICLES = 1 << 18


def _batch_key(sizes, mask, precision="f32c"):
    pairs = float(np.sum(np.asarray(sizes, dtype=np.float64) ** 2))
    bucket = int(round(math.log2(pairs))) if pairs > 1 else 0
    return f"batch|{precision}|mask{int(mask)}|pairs2^{bucket}|sets{len(sizes)}"


def best_batch_config(sizes, mask, precision="f32c", path=None):
    """Cached (tile_config, grid) of the batched launch for sets of these
    sizes and this output mask (total pair count rounded to a power of two),
    or None."""
    with _lock:
        c = load_cache(path)
    e = c.get(_batch_key(sizes, mask, precision))
    return (int(e["config"]), int(e["grid"])) if e else None


def tune_batch(sets, pse=True, les=True, configs=None, grids=None, reps=2, max_rel=1e-6,
               path=None, save=True, log=None):
    """Times the batched persistent launch (one stage-1 evaluation: velocity,
    stretching and, as requested, PSE and the LES outputs, over every set) for
    each tile configuration x persistent grid size on these sets, checks every
    set's outputs against the FP64 batched pass (rel-L2 <= ``max_rel`` per
    set and output), and caches the fastest accurate (config, grid) under the
    sets' output mask and total pair count.  ``grids``: persistent workgroup
    counts (0: as many as fit; default 0 and 1-3 per CU).  Returns (best,
    report); best is None (nothing cached, a warning) when no candidate meets
    the bound.  ``log``: a callable that receives one line per candidate."""
    from . import batch as B
    from . import _lib as L

    total = sum(len(q[0]) for q in sets)
    if total > TUNE_BATCH_MAX_PARTICLES:
        raise NotImplementedError(
            f"tune_batch: {total} particles exceed {TUNE_BATCH_MAX_PARTICLES}; the sweep over "
            "every launch shape is run on smaller batches only (see the module notes)")
    P.require_gpu()
    ref = B.BatchedSystem(sets, dt=0.0, precision="f64", pse=pse, les=les)
    mask = ref.stage_mask
    s = ref.stream.cuda_stream
    ref.evaluate(ref.x, ref.a, mask, True, s)
    ref.synchronize()
    outs = [("vel", ref.vel), ("stretch", ref.str_)] + ([("pse", ref.pse_)] if pse else []) + \
        ([("vort", ref.vort)] if les else [])
    refv = {k: v.clone() for k, v in outs}
    refe = ref.enstrophy.clone() if les else None
    del ref
    bs = B.BatchedSystem(sets, dt=0.0, precision="f32c", pse=pse, les=les)
    cus = torch.cuda.get_device_properties(bs.device).multi_processor_count
    configs = list(range(len(P.tile_configs()))) if configs is None else list(configs)
    grids = [0, cus, 2 * cus, 3 * cus] if grids is None else list(grids)
    sl = [bs.set_slice(k) for k in range(bs.nsets)]
    report = []
    for cfg in configs:
        for grid in grids:
            bs.set_launch(cfg, grid)
            run = lambda: bs.evaluate(bs.x, bs.a, mask, False, s)
            bs.evaluate(bs.x, bs.a, mask, True, s)  # sorts (first time) and warms
            bs.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(bs.stream)
            for _ in range(reps):
                run()
            e1.record(bs.stream)
            bs.synchronize()
            ms = e0.elapsed_time(e1) / reps
            worst = 0.0
            for k, v in refv.items():
                got = dict(vel=bs.vel, stretch=bs.str_, pse=bs.pse_, vort=bs.vort)[k]
                for q in sl:
                    nr = float(v[:, q].norm())
                    if nr > 0:
                        worst = max(worst, float((got[:, q] - v[:, q]).norm()) / nr)
            if refe is not None:
                worst = max(worst, float(((bs.enstrophy - refe).abs() /
                                          refe.abs().clamp_min(1e-300)).max()))
            report.append({"config": cfg, "grid": grid, "ms": ms, "rel_l2": worst})
            if log is not None:
                log(f"[tune_batch] config {cfg} grid {grid}: {ms:.2f} ms, rel-L2 {worst:.2e}")
    ok = [r for r in report if r["rel_l2"] <= max_rel]
    if not ok:
        warnings.warn(f"tune_batch: no launch reached rel-L2 <= {max_rel:g}; nothing cached",
                      RuntimeWarning, stacklevel=2)
        return None, report
    best = min(ok, key=lambda r: r["ms"])
    if save:
        with _lock:
            c = load_cache(path)
            c[_batch_key(bs.sizes, mask)] = {
                "config": best["config"], "grid": best["grid"], "ms": best["ms"],
                "rel_l2": best["rel_l2"], "sets": int(bs.nsets), "candidates": report}
            save_cache(c, path)
    return best, report


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
