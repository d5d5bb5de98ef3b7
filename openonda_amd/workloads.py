"""Drivers of the benchmark workloads (one GPU).

``python -m openonda_amd.workloads <name> [...]`` prints one JSON line per
workload:

* ``ring``      -- thin ring on a torus, N = 4096: one RK4 step of FP64
                   velocity + stretching through :class:`ParticleSystem`.
* ``leapfrog``  -- two coaxial rings, N = 65,536, 200 RK4 steps of the
                   FP32-compensated velocity + stretching + PSE (nu = 1e-3)
                   with FP64 state; every 10 steps: linear/angular impulse
                   and kinetic-energy drift against step 0, and rel-L2 of each
                   output against the FP64 pass on the same state.
* ``sweep``     -- random clouds N = 2^16, 2^18, 2^20: FP32-compensated vs
                   plain FP64 vs fused FP64 time and rel-L2.
* ``rollout``   -- ring N = 131,072, 50 RK4 steps (dt = 0.005), probe-velocity
                   loss against a ring of radius 1.05, gradient by the
                   hand-written adjoint with checkpoints every 10 steps.
* ``examples``  -- the reference's example cases E1-E5 at their scripts'
                   resolution through ParticleSystem (``--precision``: plain
                   FP64 as scripted, f64 or f32c), one untimed step then
                   ``--steps`` timed driver steps (default 10): particles,
                   ms per step, circulation and kinetic-energy drift.

The FP64 references on the GPU are the library's FP64 kernels, which the
parity tests hold to the CPU oracle.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import autograd as AG
from . import flow_models as FM
from . import physics as P
from . import _lib as L

import ctypes as C


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def _rel(a, b):
    return float((a - b).norm() / b.norm())


def _ev():
    return torch.cuda.Event(enable_timing=True)


def ring(dev):
    from .vpmModule import ParticleSystem

    x, a, s, nu = FM.torus_ring(64)
    ps = ParticleSystem(time_step_size=0.01, time_integration_method="RK4", processing_unit="GPU",
                        flow_model="DNS", viscous_scheme="CoreSpreading", write_backups=False)
    ps.add_particle_field(x, np.zeros_like(x), a, s, nu)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ps.update_velocities()
    ps.update_strengths()
    ps.update_state()
    torch.cuda.synchronize()
    return {"workload": "ring N=4096, 1 RK4 step, plain FP64 velocity + stretching",
            "n": len(x), "seconds": time.perf_counter() - t0,
            "centroid_dx": float(np.mean(ps.get_particle_positions()[:, 0] - x[:, 0]))}


def _diag(pos, al, sig):
    out = torch.empty(10, dtype=torch.float64, device=pos.device)
    n = pos.shape[1]
    ws = P.workspace(pos.device, L.load().vpm_state_workspace_bytes(n))
    L.call("vpm_diagnostics", n, P._ld(pos), pos.data_ptr(), al.data_ptr(), out.data_ptr(),
           ws.data_ptr(), ws.numel(), P._stream(pos))
    return out


def leapfrog(dev, steps=200, every=10, dt=0.01):
    x, a, s, nu = FM.leapfrog_rings()
    pos, al = P.to_soa(x, dev), P.to_soa(a, dev)
    sig, nut = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
    outs = ("vel", "stretch", "pse")
    d0 = _diag(pos, al, sig)
    e0 = float(P.total_kinetic_energy(pos, al, sig))
    rows = []
    t_pairs = 0.0
    t_ms = 0.0
    for k in range(steps + 1):
        if k % every == 0:
            r32 = P.forward(pos, al, sig, nut, outputs=outs, precision="f32c", sigma_mode="uniform")
            r64 = P.forward(pos, al, sig, nut, outputs=outs, precision="f64", sigma_mode="uniform")
            d = _diag(pos, al, sig)
            e = float(P.total_kinetic_energy(pos, al, sig))
            rows.append({"step": k,
                         "rel_l2": {o: _rel(r32[o], r64[o]) for o in outs},
                         "linear_impulse_drift": float((d[0:3] - d0[0:3]).norm() / d0[0:3].norm()),
                         "angular_impulse": [float(v) for v in d[3:6]],
                         "energy_drift": (e - e0) / abs(e0)})
            _log(f"[leapfrog] step {k}: {rows[-1]}")
        if k == steps:
            break
        s0, s1 = _ev(), _ev()
        s0.record()
        with torch.no_grad():
            pos, al = AG.rk4_coupled_step(pos, al, sig, nut, dt, precision="f32c",
                                          sigma_mode="uniform", pse=True)
        s1.record()
        torch.cuda.synchronize()
        t_ms += s0.elapsed_time(s1)
        t_pairs += 4.0 * float(len(x)) ** 2
    return {"workload": f"leapfrogging rings N={len(x)}, {steps} RK4 steps dt={dt}, "
                        "FP32-compensated velocity+stretching+PSE, FP64 state",
            "ms_per_step": t_ms / steps, "pairs_per_s": t_pairs / (t_ms * 1e-3), "report": rows}


def sweep(dev, sizes=(1 << 16, 1 << 18, 1 << 20)):
    res = []
    for n in sizes:
        x, a, s, nu = FM.random_cloud(n)
        pos, al, sig = P.to_soa(x, dev), P.to_soa(a, dev), torch.from_numpy(s).to(dev)
        times = {}
        outs = {}
        for name, fn in (
                ("f32c", lambda: P.forward(pos, al, sig, outputs=("vel", "stretch"),
                                           precision="f32c", sigma_mode="uniform")),
                ("fused_f64", lambda: P.forward(pos, al, sig, outputs=("vel", "stretch"),
                                                precision="f64", sigma_mode="uniform")),
                ("plain_f64", lambda: {"vel": P.selfinduced_velocity(pos, al, sig),
                                       "stretch": P.strength_gradients(pos, al, sig)})):
            outs[name] = fn()
            torch.cuda.synchronize()
            e0, e1 = _ev(), _ev()
            e0.record()
            outs[name] = fn()
            e1.record()
            torch.cuda.synchronize()
            times[name] = e0.elapsed_time(e1)
        rel = {k: _rel(outs["f32c"][k], outs["plain_f64"][k]) for k in ("vel", "stretch")}
        row = {"n": n, "ms": times, "pairs_per_s": {k: n * n / (v * 1e-3) for k, v in times.items()},
               "rel_l2_f32c_vs_plain": rel,
               "speedup_f32c_vs_plain": times["plain_f64"] / times["f32c"]}
        _log(f"[sweep] {row}")
        res.append(row)
    return {"workload": "random cloud sweep, velocity + stretching", "rows": res}


def rollout(dev, n_st=2048, steps=50, dt=0.005, ck=10, probes=4096):
    x, a, s, nu = FM.torus_ring(n_st, nu=1e-3)
    xt, at, st, _ = FM.torus_ring(n_st, R=1.05, nu=1e-3)
    y = P.to_soa(FM.ring_probes(probes), dev)
    sig, nut = torch.from_numpy(s).to(dev), torch.from_numpy(nu).to(dev)
    target = P.induced_velocity_at(y, P.to_soa(xt, dev), P.to_soa(at, dev),
                                   torch.from_numpy(st).to(dev))
    x0 = P.to_soa(x, dev).requires_grad_(True)
    a0 = P.to_soa(a, dev).requires_grad_(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, as_ = AG.rollout(x0, a0, sig, nut, dt, steps, checkpoint_every=ck, precision="f32c")
    u = AG.probe_velocity(y, xs, as_, sig)
    loss = ((u - target) ** 2).sum(0).mean()
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t1 = time.perf_counter()
    gx, ga = torch.autograd.grad(loss, (x0, a0))
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t1
    return {"workload": f"ring N={len(x)}, {steps} RK4 steps dt={dt}, probe loss, "
                        f"adjoint gradient with checkpoints every {ck} steps",
            "loss": float(loss), "forward_s": t_fwd, "backward_s": t_bwd,
            "backward_over_forward": t_bwd / t_fwd,
            "grad_norm_x": float(gx.norm()), "grad_norm_alpha": float(ga.norm()),
            "peak_mem_GB": torch.cuda.max_memory_allocated(dev) / 1e9}


def examples(dev, steps=10, names=("E1", "E2", "E3", "E4", "E5"), precision="plain"):
    from . import examples as EX
    from .vpmModule import ParticleSystem

    rows = []
    for name in names:
        case = EX.CASES[name]()
        ps = ParticleSystem(**case.system, precision=precision, write_backups=False)
        for p, v, a, s, nu, gid in case.groups:
            ps.add_particle_field(p, v, a, s, nu, group_id=gid)
        ps.remove_weak_particles(**case.remove_weak)
        c0, e0 = ps.get_total_strength_magnitude(), ps.get_total_kinetic_energy()

        def step():
            ps.update_velocities()
            ps.update_strengths()
            ps.update_state()

        step()  # untimed: first launches load code objects, workspaces grow
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        row = {"case": name, "n": len(ps), "timed_steps": steps, "steps": steps + 1,
               "of": case.n_time_steps,
               "flow_model": case.system["flow_model"], "ms_per_step": ms,
               "circulation_drift": ps.get_total_strength_magnitude() / c0 - 1.0,
               "kinetic_energy_drift": ps.get_total_kinetic_energy() / e0 - 1.0}
        _log(f"[examples] {row}")
        rows.append(row)
    return {"workload": f"reference examples E1-E5 at script resolution, {steps} steps each, "
                        f"precision {precision}", "rows": rows}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("name", choices=["ring", "leapfrog", "sweep", "rollout", "examples"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--precision", choices=["plain", "f64", "f32c"], default="plain",
                    help="ParticleSystem arithmetic for the examples workload")
    a = ap.parse_args(argv)
    P.require_gpu()
    dev = torch.device("cuda", 0)
    if a.name == "ring":
        r = ring(dev)
    elif a.name == "leapfrog":
        r = leapfrog(dev, steps=a.steps or 200)
    elif a.name == "sweep":
        r = sweep(dev)
    elif a.name == "examples":
        r = examples(dev, steps=a.steps or 10, precision=a.precision)
    else:
        r = rollout(dev, steps=a.steps or 50)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
