"""Benchmark: all-pairs regularised Biot-Savart, velocity + stretching.

Workload: random vortex-blob cloud, N = 1,048,576 particles (positions
U[0,2pi]^3, strengths N(0,1/N), sigma = 1.3*2pi/N^(1/3), seed 3+log2 N);
one step = one fused FP32-compensated velocity + stretching evaluation over
all N^2 pairs (Morton sort + pack + pair pass), inputs resident in HBM.

Prints one JSON line: pairs/s, the fused kernel's roofline position
(algorithmic FLOP/s against the MI355X FP32 vector peak, from HIP events on
the launch stream), rel-L2 of the FP32 result against FP64 on a strided target
sample, the plain FP64 drop-in kernels' time on the same input, the same pass
called host-to-host (NumPy in and out, PCIe copies included), the adjoint /
forward time ratio and the CPU baseline (C/OpenMP FP64 port of the reference's
Numba CPU plugin) timed on this host.  --skip-plain skips the comparison legs
(plain FP64, fused FP64, host-to-host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N]
       [--skip-cpu] [--skip-plain] [--skip-adjoint] [--json-only]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "pair-interactions/s & % FP32 peak at N=1M (vs FP64 plain); adjoint/fwd ratio"
FP32_VECTOR_PEAK_TFLOPS = 157.3   # MI355X FP32 vector (VALU) peak
FP64_VECTOR_PEAK_TFLOPS = 78.6
# Algorithmic flops per pair of the fused velocity + stretching pass with a
# uniform radius (shared r, rho^2, one rsqrt, t^5/t^7 powers, the velocity
# cross product, the alpha_j sum for alpha_i x G and the f2 (alpha_i.c) r
# term); Kahan and FP64 folds are not counted.
FLOPS_PER_PAIR_VEL_STRETCH = 53
FLOPS_PER_PAIR_VEL_STRETCH_PSE = 61


def traffic_workload(n):
    """Name of the workload in the traffic summaries under profiles/."""
    return f"random cloud N={n}, velocity+stretching, f32c"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_probe(which, dev):
    """One launch of a secondary kernel (after one untimed warm-up launch),
    for counter runs: the ring of 131072 (velocity + stretching + PSE,
    FP32-compensated), its adjoint, or the 96-set batch's launches of stage 1
    (velocity + stretching + PSE + LES outputs) and of a later stage (velocity
    + stretching + PSE)."""
    import torch

    from openonda_amd import autograd as AG
    from openonda_amd import flow_models as FM
    from openonda_amd import physics as P

    if which in ("ring", "adjoint"):
        x, a, s, nu = FM.torus_ring(2048, nu=1e-3)
        rp, ra = P.to_soa(x, dev), P.to_soa(a, dev)
        rs, rn = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
        if which == "ring":
            fn = lambda: P.forward(rp, ra, rs, rn, outputs=("vel", "stretch", "pse"),
                                   precision="f32c", sigma_mode="uniform")
        else:
            g = torch.Generator(device=dev).manual_seed(0)
            cots = [torch.randn(3, len(x), dtype=torch.float64, device=dev, generator=g)
                    for _ in range(3)]
            fn = lambda: AG.adjoint(rp, ra, rs, rn, *cots, precision="f32c", sigma_mode="uniform")
    else:
        from openonda_amd import batch as B

        bs = B.BatchedSystem(FM.batch_instances(), dt=0.005, precision="f32c", pse=True, les=True)

        def fn():  # an RK4 step's stage 1 (with the LES outputs), then a later stage
            bs.evaluate(bs.x, bs.a, bs.stage_mask, True, bs.stream.cuda_stream)
            bs.evaluate(bs.x, bs.a, bs._mask(False), False, bs.stream.cuda_stream)
            bs.stream.synchronize()
    fn()
    torch.cuda.synchronize()
    fn()
    torch.cuda.synchronize()
    log(f"[bench] pmc probe done: {which}, 2 launches")
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--cpu-targets", type=int, default=32768)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--pmc-probe", nargs="?", const="cloud", default=None,
                    choices=["cloud", "ring", "adjoint", "batch"],
                    help="for rocprofv3 --pmc runs: one launch of the kernel named (cloud: "
                         "the headline pass and one 256 MiB copy; ring: velocity + "
                         "stretching + PSE on the ring of 131072; adjoint: its adjoint; "
                         "batch: a stage-1 and a later-stage launch over the 96 sets)")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the pair sum "
              "would shard by targets, each GPU holding all sources)", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from openonda_amd import flow_models as FM
    from openonda_amd import physics as P
    from openonda_amd import _lib as L
    import ctypes as C

    P.require_gpu()
    dev = torch.device("cuda", 0)
    if args.pmc_probe in ("ring", "adjoint", "batch"):
        return pmc_probe(args.pmc_probe, dev)
    n = args.n
    log(f"[bench] building random cloud N={n}")
    x, a, s, nu = FM.random_cloud(n)
    pos, alpha = P.to_soa(x, dev), P.to_soa(a, dev)
    sig = torch.from_numpy(s).to(dev)
    nu_t = torch.from_numpy(nu).to(dev)

    outputs = ("vel", "stretch")
    opts = L.ForwardOpts(L.OUT_VEL | L.OUT_STRETCH, L.PREC_F32C, L.SIGMA_UNIFORM, 0,
                         args.tile_config)
    part = P._particles(pos, alpha, sig)
    vel = torch.empty((3, n), dtype=torch.float64, device=dev)
    st = torch.empty((3, n), dtype=torch.float64, device=dev)
    fo = L.ForwardOut(n, vel.data_ptr(), st.data_ptr(), None, None, None)
    lib = L.load()
    wsb = lib.vpm_forward_workspace_bytes(n, C.byref(opts))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream(dev)
    sh = stream.cuda_stream

    def step(ev=None):
        L.call("vpm_forward_prepare", C.byref(part), C.byref(opts), ws.data_ptr(), wsb, sh)
        if ev is not None:
            ev[0].record(stream)
        L.call("vpm_forward_run", C.byref(part), C.byref(opts), C.byref(fo), ws.data_ptr(), wsb,
               sh)
        if ev is not None:
            ev[1].record(stream)

    if args.pmc_probe:
        step()
        src = torch.ones(1 << 25, dtype=torch.float64, device=dev)  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] pmc probe done: 1 fused step, 1 copy of 268435456 bytes")
        return 0

    log(f"[bench] warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    events = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(events[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    kern_ms = sum(e0.elapsed_time(e1) for e0, e1 in events) / args.steps
    pairs = float(n) * float(n)
    value = pairs / (ms_step * 1e-3)
    achieved = FLOPS_PER_PAIR_VEL_STRETCH * pairs / (kern_ms * 1e-3) / 1e12
    log(f"[bench] f32c: {ms_step:.2f} ms/step, kernel {kern_ms:.2f} ms, {value:.3e} pairs/s, "
        f"{achieved:.1f} TFLOP/s")

    # HBM bytes per launch of this kernel instance on this workload, from the
    # calibrated FETCH_SIZE + WRITE_SIZE runs summarised under profiles/
    from openonda_amd import profiling

    cfg = P.tile_configs()[args.tile_config if args.tile_config >= 0
                           else P.tile_config_default(outputs)]
    kernel_name = "fused_kernel<float, 3u, true, {}, {}, {}, {}, {}>".format(*cfg)
    traffic = profiling.load_traffic(traffic_workload(n), kernel_name)
    # executed FP32 flops per pair of this kernel instance, from the VALU
    # instruction counters summarised under profiles/ (FMA 2, others 1)
    exec_fpp = profiling.load_record(traffic_workload(n), kernel_name, "exec_fp32_flop_per_pair")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"random vortex-blob cloud, N={n}, fused velocity+stretching, "
                        "FP32 pairs with compensated FP64 totals",
            "n": n, "outputs": list(outputs), "precision": "f32c",
            "sigma": float(s[0]), "seed": 3 + int(round(math.log2(n))),
            "tile_config": args.tile_config,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_VECTOR_PEAK_TFLOPS,
            "traffic": traffic,
            "kernel": kernel_name,
            "kernel_ms": kern_ms,
            "flops_per_pair": FLOPS_PER_PAIR_VEL_STRETCH,
        },
    }
    if exec_fpp is not None:
        ex = exec_fpp * pairs / (kern_ms * 1e-3) / 1e12
        result["roofline"]["executed"] = {
            "flops_per_pair": exec_fpp, "achieved": ex, "frac": ex / FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "source": "SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F32 per pair of this kernel instance "
                      "(rocprofv3 --pmc, profiles/), times this run's pairs / kernel time"}

    # ---- plain FP64 drop-in kernels on the same input ---------------------
    if not args.skip_plain:
        log("[bench] plain FP64 velocity + stretching")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        P.selfinduced_velocity(pos, alpha, sig)
        P.strength_gradients(pos, alpha, sig)
        e1.record(stream)
        torch.cuda.synchronize()
        plain_ms = e0.elapsed_time(e1)
        result["plain_f64"] = {"ms": plain_ms, "pairs_per_s": pairs / (plain_ms * 1e-3),
                               "speedup_f32c_vs_plain": plain_ms / ms_step}
        log(f"[bench] plain f64: {plain_ms:.1f} ms ({plain_ms / ms_step:.2f}x slower)")
        # fused FP64 (same sharing of r, rho^2, one rsqrt per pair)
        o64 = L.ForwardOpts(L.OUT_VEL | L.OUT_STRETCH, L.PREC_F64, L.SIGMA_UNIFORM, 0, -1)
        torch.cuda.synchronize()
        e0.record(stream)
        P.forward(pos, alpha, sig, outputs=outputs, precision="f64", sigma_mode="uniform")
        e1.record(stream)
        torch.cuda.synchronize()
        f64_ms = e0.elapsed_time(e1)
        result["fused_f64"] = {"ms": f64_ms, "speedup_f32c_vs_fused_f64": f64_ms / ms_step,
                               "fp64_tflops": FLOPS_PER_PAIR_VEL_STRETCH * pairs / f64_ms / 1e9,
                               "fp64_peak_tflops": FP64_VECTOR_PEAK_TFLOPS}
        log(f"[bench] fused f64: {f64_ms:.1f} ms")

        # The same pass from host arrays: the drop-in module's boundary takes
        # (N,3) float64 NumPy arrays and returns NumPy, so each call copies its
        # inputs over PCIe, runs the pass and copies the result back
        # (ParticlesPhysicsCuda.py:16-34 does the same).  Wall-clock times.
        from openonda_amd import ParticlesPhysicsHip as PPH

        log("[bench] host-to-host: one velocity + stretching pass from NumPy")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hp, ha = P.to_soa(x, dev), P.to_soa(a, dev)
        hs = torch.from_numpy(s).to(dev)
        ho = P.forward(hp, ha, hs, outputs=outputs, precision="f32c", sigma_mode="uniform")
        hu, hS = P.from_soa(ho["vel"]), P.from_soa(ho["stretch"])
        one_s = time.perf_counter() - t0
        del hp, ha, hs, ho
        log("[bench] host-to-host: drop-in get_selfinduced_velocity + get_strength_gradients")
        prec0 = PPH.PRECISION
        PPH.PRECISION = "f32c"
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            du = PPH.get_selfinduced_velocity(x, a, s)
            dS = PPH.get_strength_gradients(x, a, s)
            two_s = time.perf_counter() - t0
        finally:
            PPH.PRECISION = prec0
        agree = max(float(np.linalg.norm(du - hu) / np.linalg.norm(hu)),
                    float(np.linalg.norm(dS - hS) / np.linalg.norm(hS)))
        result["host_to_host"] = {
            "one_pass_ms": one_s * 1e3, "one_pass_pairs_per_s": pairs / one_s,
            "drop_in_two_calls_ms": two_s * 1e3, "drop_in_pairs_per_s": pairs / two_s,
            "drop_in_vs_one_pass_rel_l2": agree,
            "sample": "wall clock of one call each, NumPy in and out (PCIe copies, "
                      "(N,3)<->(3,N) transposes, Morton sort, pass); not the headline value"}
        log(f"[bench] host-to-host: one pass {one_s * 1e3:.1f} ms, drop-in pair "
            f"{two_s * 1e3:.1f} ms (rel {agree:.1e})")

    # ---- CPU baseline + FP64 accuracy sample -------------------------------
    if not args.skip_cpu:
        from oracle import cpu_baseline as CB

        m = min(args.cpu_targets, n)
        log(f"[bench] CPU baseline: {m} strided targets x {n} sources, {CB.threads()} threads")
        pps, secs, (u_ref, S_ref), tidx = CB.time_velocity_stretching(x, a, s, m)
        u = vel[:, torch.from_numpy(tidx).to(dev)].t().cpu().numpy()
        S = st[:, torch.from_numpy(tidx).to(dev)].t().cpu().numpy()
        rel = {"vel": float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref)),
               "stretch": float(np.linalg.norm(S - S_ref) / np.linalg.norm(S_ref))}
        result["rel_l2_vs_fp64"] = rel
        result["rel_l2_sample"] = f"{len(tidx)} strided targets (stride {n // m}) vs all sources"
        result["cpu_baseline"] = {
            "value": pps, "unit": "pair-interactions/s", "cores": CB.threads(), "kind": "port",
            "sample": f"velocity + stretching (two passes, as the reference runs them) for "
                      f"{len(tidx)} strided targets x {n} sources in {secs:.1f} s; "
                      f"C/OpenMP FP64 port of ParticlesPhysics.py, -O3 -ffast-math "
                      f"-march={CB.isa()}",
            "cpu_model": CB.host_cpu_model(), "isa": CB.isa(),
        }
        log(f"[bench] CPU {pps:.3e} pairs/s ({secs:.1f} s); rel-L2 {rel}")
        m1 = max(1, m // 64)
        pps1, secs1, _, _ = CB.time_velocity_stretching(x, a, s, m1, nthreads=1)
        result["cpu_baseline"]["single_core"] = {
            "value": pps1, "sample": f"{m1} strided targets x {n} sources in {secs1:.1f} s"}
        log(f"[bench] CPU single core {pps1:.3e} pairs/s ({secs1:.1f} s)")

    # ---- adjoint / forward ratio (differentiation workload size) -----------
    if not args.skip_adjoint:
        from openonda_amd import autograd as AG

        xr, ar, sr, nur = FM.torus_ring(2048, nu=1e-3)  # N = 131072
        nr = len(xr)
        rp, ra = P.to_soa(xr, dev), P.to_soa(ar, dev)
        rs, rn = torch.from_numpy(sr).to(dev), torch.from_numpy(nur).to(dev)
        g = torch.Generator(device=dev).manual_seed(0)
        cots = [torch.randn(3, nr, dtype=torch.float64, device=dev, generator=g) for _ in range(3)]

        def fwd():
            P.forward(rp, ra, rs, rn, outputs=("vel", "stretch", "pse"), precision="f32c",
                      sigma_mode="uniform")

        def adj():
            AG.adjoint(rp, ra, rs, rn, *cots, precision="f32c", sigma_mode="uniform")

        times = {}
        for name, fn in (("forward", fwd), ("adjoint", adj)):
            fn()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(3):
                fn()
            e1.record(stream)
            torch.cuda.synchronize()
            times[name] = e0.elapsed_time(e1) / 3
        result["adjoint"] = {
            "workload": f"ring, N={nr}, velocity+stretching+PSE, FP32-compensated",
            "forward_ms": times["forward"], "adjoint_ms": times["adjoint"],
            "adjoint_over_forward": times["adjoint"] / times["forward"]}
        log(f"[bench] adjoint/forward = {times['adjoint']:.2f}/{times['forward']:.2f} ms "
            f"= {times['adjoint'] / times['forward']:.2f}")

        # the ring's fused pass alone: prepared once, HIP events around the
        # run on the launch stream (the kernel and, for these 512 target
        # blocks, the split reduction), as for the headline kernel
        router = ("vel", "stretch", "pse")
        ro = L.ForwardOpts(L.OUT_VEL | L.OUT_STRETCH | L.OUT_PSE, L.PREC_F32C, L.SIGMA_UNIFORM,
                           0, -1)
        rpart = P._particles(rp, ra, rs, rn)
        rbuf = [torch.empty((3, nr), dtype=torch.float64, device=dev) for _ in range(3)]
        rfo = L.ForwardOut(nr, *[b.data_ptr() for b in rbuf], None, None)
        rwsb = lib.vpm_forward_workspace_bytes(nr, C.byref(ro))
        rws = torch.empty(rwsb, dtype=torch.uint8, device=dev)
        L.call("vpm_forward_prepare", C.byref(rpart), C.byref(ro), rws.data_ptr(), rwsb, sh)
        L.call("vpm_forward_run", C.byref(rpart), C.byref(ro), C.byref(rfo), rws.data_ptr(), rwsb,
               sh)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(10):
            L.call("vpm_forward_run", C.byref(rpart), C.byref(ro), C.byref(rfo), rws.data_ptr(),
                   rwsb, sh)
        e1.record(stream)
        torch.cuda.synchronize()
        ring_ms = e0.elapsed_time(e1) / 10
        rpairs = float(nr) * float(nr)
        rcfg = P.tile_configs()[P.tile_config_default(router)]
        rkernel = "fused_kernel<float, 7u, true, {}, {}, {}, {}, {}>".format(*rcfg)
        rach = FLOPS_PER_PAIR_VEL_STRETCH_PSE * rpairs / (ring_ms * 1e-3) / 1e12
        rroof = {"bound": "valu", "achieved": rach, "peak": FP32_VECTOR_PEAK_TFLOPS,
                 "unit": "TFLOP/s", "frac": rach / FP32_VECTOR_PEAK_TFLOPS, "kernel": rkernel,
                 "kernel_ms": ring_ms, "flops_per_pair": FLOPS_PER_PAIR_VEL_STRETCH_PSE}
        rexec = profiling.load_record(f"ring N={nr}, velocity+stretching+PSE, f32c", rkernel,
                                      "exec_fp32_flop_per_pair")
        if rexec is not None:
            rx = rexec * rpairs / (ring_ms * 1e-3) / 1e12
            rroof["executed"] = {"flops_per_pair": rexec, "achieved": rx,
                                 "frac": rx / FP32_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s"}
        result["adjoint"]["forward_roofline"] = rroof
        log(f"[bench] ring forward pass alone: {ring_ms:.2f} ms, {rach:.1f} TFLOP/s "
            f"({rach / FP32_VECTOR_PEAK_TFLOPS:.3f} of peak)")

    # ---- batched persistent launch, graph-replayed RK4 steps -----------------
    if not args.skip_batch:
        from openonda_amd import batch as B

        sets = FM.batch_instances()
        # the launch the batch autotuner cached for this workload, else the
        # library's default shape (the sweep itself is limited to smaller
        # batches: autotune.TUNE_BATCH_MAX_PARTICLES)
        bs = B.BatchedSystem(sets, dt=0.005, precision="f32c", pse=True, les=True,
                             tile_config="auto")
        log(f"[bench] batch: {bs.nsets} sets, N_total={bs.n}, {bs.n_items} work items, "
            f"tile config {bs.tile_config}, grid {bs.grid} "
            f"({'autotuned' if bs.autotuned else 'default'})")
        bs.step()          # sorts, eager
        bs.capture()
        bs.replay(1)       # warm
        bs.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(bs.stream)
        bs.replay(args.batch_steps)
        e1.record(bs.stream)
        bs.synchronize()
        bms = e0.elapsed_time(e1)
        # per-set accuracy of one FP32-compensated evaluation against the plain
        # FP64 kernels, in the sorted order the replayed steps used (no re-sort)
        bs.evaluate(bs.x, bs.a, 3, False, bs.stream.cuda_stream)
        bs.synchronize()
        errs, per_set = [], []
        for k in range(bs.nsets):
            sl = bs.set_slice(k)
            xs, as_, ss = bs.x[:, sl], bs.a[:, sl], bs.sigma[sl]
            u = P.selfinduced_velocity(xs, as_, ss)
            sg = P.strength_gradients(xs, as_, ss)
            ev = float((bs.vel[:, sl] - u).norm() / u.norm())
            es = float((bs.str_[:, sl] - sg).norm() / sg.norm())
            errs.append(max(ev, es))
            per_set.append({"set": k, "n": int(bs.sizes[k]), "kind": "ring" if k % 2 == 0 else "cloud",
                            "vel": ev, "stretch": es})
        pairs_step = 4 * bs.pairs_per_evaluation()
        result["batch"] = {
            "workload": f"{bs.nsets} ring/cloud sets, N_i log-uniform in [2048, 131072], "
                        f"N_total={bs.n}; RK4 (velocity+stretching+PSE, LES filter per set), "
                        f"one persistent launch per stage, HIP-graph replay",
            "tile_config": bs.tile_config, "grid": bs.grid,
            "launch": "autotuned" if bs.autotuned else "default",
            "steps": args.batch_steps, "ms_per_step": bms / args.batch_steps,
            "pairs_per_s": pairs_step * args.batch_steps / (bms * 1e-3),
            "rel_l2_max_vs_fp64": max(errs), "rel_l2_median_vs_fp64": float(np.median(errs)),
            "worst_sets": sorted(per_set, key=lambda r: -max(r["vel"], r["stretch"]))[:3]}
        log(f"[bench] batch: {bms / args.batch_steps:.1f} ms/step, "
            f"{result['batch']['pairs_per_s']:.3e} pairs/s, max rel-L2 {max(errs):.2e}")

    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
