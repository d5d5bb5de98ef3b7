"""Summaries of rocprofv3 result databases (SQLite, rocpd schema) as Markdown.

``python -m openonda_amd.profiling stats DB --title T --command C``
    kernel-trace summary: calls, total and average duration per kernel.
``python -m openonda_amd.profiling traffic --fetch DB --write DB --kernel K
        --copy-bytes B --workload W --title T --command C``
    HBM traffic per dispatch of kernel K from a FETCH_SIZE run and a
    WRITE_SIZE run (KiB per dispatch).  Each counter is calibrated on the
    plain device copy of B bytes recorded in the same run: factor =
    B / 1024 / (counter of the copy), applied to the kernel's reading.  The
    result is written as a table and a fenced JSON block that bench.py reads
    (``traffic_bytes`` for the workload named by W).
``python -m openonda_amd.profiling valu DB... --kernel K --workload W --pairs P
        --flops-per-pair F --kernel-ms T --title T --command C``
    VALU counters of kernel K from one or more --pmc passes: instructions per
    pair by class, executed FP32 TFLOP/s against the vector peak beside the
    algorithmic figure, dual-issue share, wait shares; JSON block as above.
"""
from __future__ import annotations

import argparse
import json
import re
import sqlite3
from pathlib import Path

PROFILES = Path(__file__).resolve().parent.parent / "profiles"


def _short(name: str) -> str:
    name = re.sub(r"rocprim::ROCPRIM_\w+_NS::", "rocprim::", name)
    if "rocprim" in name:
        m = re.search(r"(radix_sort_\w+|merge_sort_\w+|[a-z_]+_kernel)", name)
        return "rocprim " + (m.group(1) if m else "kernel")
    return name.replace("vpm::(anonymous namespace)::", "")[:120]


def kernel_stats(db: str):
    """[(kernel, calls, total_us, avg_us, percent)] sorted by total time."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, total_calls, total_duration, percentage from top_kernels").fetchall()
    con.close()
    agg = {}
    for name, calls, total, pct in rows:
        e = agg.setdefault(_short(name), [0, 0.0, 0.0])
        e[0] += calls
        e[1] += total  # the view reports microseconds
        e[2] += pct
    out = [(k, c, t, t / c, p) for k, (c, t, p) in agg.items()]
    return sorted(out, key=lambda r: -r[2])


def counter_values(db: str, counter: str):
    """[(dispatch_id, kernel_name, value)] of one PMC counter."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select dispatch_id, kernel_name, value from counters_collection where counter_name = ? "
        "order by dispatch_id", (counter,)).fetchall()
    con.close()
    return rows


def _calibrated(rows, kernel: str, copy_bytes: int):
    copy = [v for _, k, v in rows if "copyBuffer" in k or "copy" in k.lower()]
    hits = [v for _, k, v in rows if kernel in k]
    if not copy or not hits:
        raise ValueError(f"kernel {kernel!r} or the calibration copy not found")
    factor = copy_bytes / 1024.0 / copy[-1]
    return hits[-1], factor


def traffic(fetch_db: str, write_db: str, kernel: str, copy_bytes: int):
    f_raw, f_fac = _calibrated(counter_values(fetch_db, "FETCH_SIZE"), kernel, copy_bytes)
    w_raw, w_fac = _calibrated(counter_values(write_db, "WRITE_SIZE"), kernel, copy_bytes)
    fetch_b = f_raw * f_fac * 1024.0
    write_b = w_raw * w_fac * 1024.0
    return {"fetch_kib_raw": f_raw, "fetch_factor": f_fac, "write_kib_raw": w_raw,
            "write_factor": w_fac, "fetch_bytes": fetch_b, "write_bytes": write_b,
            "traffic_bytes": fetch_b + write_b}


def stats_markdown(db: str, title: str, command: str) -> str:
    lines = [f"# {title}", "", "Command (from the repository root, on one MI355X):", "",
             "```", command, "```", "",
             "rocprofv3 kernel-trace summary (durations in microseconds):", "",
             "| kernel | calls | total us | avg us | % time |", "|---|---|---|---|---|"]
    for k, c, t, a, p in kernel_stats(db):
        lines.append(f"| `{k}` | {c} | {t:.1f} | {a:.1f} | {p:.3f} |")
    return "\n".join(lines) + "\n"


def traffic_markdown(fetch_db, write_db, kernel, copy_bytes, workload, title, command) -> str:
    t = traffic(fetch_db, write_db, kernel, copy_bytes)
    rec = {"workload": workload, "kernel": kernel, "traffic_bytes": t["traffic_bytes"],
           "fetch_bytes": t["fetch_bytes"], "write_bytes": t["write_bytes"],
           "fetch_factor": t["fetch_factor"], "write_factor": t["write_factor"]}
    lines = [f"# {title}", "", "Commands (from the repository root, on one MI355X):", "",
             "```", command, "```", "",
             f"Calibration: a device copy of {copy_bytes} bytes in the same process.", "",
             "| counter | kernel KiB/dispatch | copy factor | bytes/dispatch |", "|---|---|---|---|",
             f"| FETCH_SIZE | {t['fetch_kib_raw']:.1f} | {t['fetch_factor']:.4f} | "
             f"{t['fetch_bytes']:.4e} |",
             f"| WRITE_SIZE | {t['write_kib_raw']:.1f} | {t['write_factor']:.4f} | "
             f"{t['write_bytes']:.4e} |", "",
             f"HBM traffic per launch: {t['traffic_bytes']:.4e} bytes.", "",
             "```json", json.dumps(rec), "```"]
    return "\n".join(lines) + "\n"


def kernel_counters(dbs, kernel: str):
    """{counter: value} of the last dispatch of ``kernel`` in each database
    (one rocprofv3 --pmc pass per database)."""
    vals = {}
    for db in dbs:
        con = sqlite3.connect(db)
        names = [r[0] for r in con.execute("select distinct counter_name from counters_collection")]
        con.close()
        for c in names:
            hits = [v for _, k, v in counter_values(db, c) if kernel in k]
            if hits:
                vals[c] = hits[-1]
    return vals


# FP32 flops per lane of each counted VALU instruction class (FMA 2)
_FP32 = {"SQ_INSTS_VALU_FMA_F32": 2, "SQ_INSTS_VALU_MUL_F32": 1, "SQ_INSTS_VALU_ADD_F32": 1,
         "SQ_INSTS_VALU_TRANS_F32": 1}
_FP64 = {"SQ_INSTS_VALU_FMA_F64": 2, "SQ_INSTS_VALU_MUL_F64": 1, "SQ_INSTS_VALU_ADD_F64": 1}


def valu_summary(vals, pairs, flops_per_pair, kernel_ms, cus=256, xcds=8, peak_tflops=157.3):
    """Executed-instruction figures of one launch from its counters: VALU
    instructions per pair (wave instructions x 64 lanes / pairs), FP32 and
    FP64 flops per pair by instruction class, executed FP32 TFLOP/s against
    the vector peak beside the algorithmic figure, and issue statistics."""
    lanes = 64.0
    out = {"pairs": pairs, "kernel_ms": kernel_ms}
    per_pair = lambda c: vals[c] * lanes / pairs if c in vals else None
    for c in ("SQ_INSTS_VALU", *_FP32, *_FP64, "SQ_INSTS_VALU_CVT", "SQ_INSTS_LDS",
              "SQ_INSTS_SMEM", "SQ_INSTS_SALU"):
        if c in vals:
            out[c + "_per_pair"] = per_pair(c)
    if all(c in vals for c in _FP32):
        f32 = sum(k * vals[c] for c, k in _FP32.items()) * lanes
        out["exec_fp32_flop_per_pair"] = f32 / pairs
        out["exec_fp32_tflops"] = f32 / (kernel_ms * 1e-3) / 1e12
        out["exec_fp32_frac"] = out["exec_fp32_tflops"] / peak_tflops
    if all(c in vals for c in _FP64):
        f64 = sum(k * vals[c] for c, k in _FP64.items()) * lanes
        out["exec_fp64_flop_per_pair"] = f64 / pairs
    out["alg_tflops"] = flops_per_pair * pairs / (kernel_ms * 1e-3) / 1e12
    out["alg_frac"] = out["alg_tflops"] / peak_tflops
    # GPU cycles of the dispatch: GRBM_GUI_ACTIVE sums the busy cycles of the
    # eight XCDs (it reads 8 x 2.37e9 per second of kernel time here); without
    # it, the kernel time at the nominal 2.4 GHz
    if "GRBM_GUI_ACTIVE" in vals and vals["GRBM_GUI_ACTIVE"]:
        cycles = vals["GRBM_GUI_ACTIVE"] / xcds
        out["clock_ghz"] = cycles / (kernel_ms * 1e-3) / 1e9
    else:
        cycles = kernel_ms * 1e-3 * 2.4e9
    if "SQ_INSTS_VALU" in vals:
        # wave64 VALU instructions issued per clock per CU (the FMA-chain probe
        # peaks at 1.44 with packed FP32 off; 2.0 is the 157.3 TFLOP/s peak)
        out["valu_wave_instr_per_clk_per_cu"] = vals["SQ_INSTS_VALU"] / (cycles * cus)
    if "SQ_ACTIVE_INST_VALU" in vals and "SQ_ACTIVE_INST_VALU2" in vals and vals["SQ_ACTIVE_INST_VALU"]:
        out["dual_issue_share"] = vals["SQ_ACTIVE_INST_VALU2"] / vals["SQ_ACTIVE_INST_VALU"]
    if "SQ_WAIT_ANY" in vals and "SQ_WAVE_CYCLES" in vals and vals["SQ_WAVE_CYCLES"]:
        out["wait_any_share"] = vals["SQ_WAIT_ANY"] / vals["SQ_WAVE_CYCLES"]
    if "SQ_WAIT_INST_ANY" in vals and "SQ_WAVE_CYCLES" in vals and vals["SQ_WAVE_CYCLES"]:
        out["wait_issue_share"] = vals["SQ_WAIT_INST_ANY"] / vals["SQ_WAVE_CYCLES"]
    return out


def valu_markdown(dbs, kernel, workload, pairs, flops_per_pair, kernel_ms, title, command):
    return render_valu(kernel_counters(dbs, kernel), kernel, workload, pairs, flops_per_pair,
                       kernel_ms, title, command)


def render_valu(vals, kernel, workload, pairs, flops_per_pair, kernel_ms, title, command):
    sm = valu_summary(vals, pairs, flops_per_pair, kernel_ms)
    rec = {"workload": workload, "kernel": kernel, **sm}
    lines = [f"# {title}", "", "Commands (from the repository root, on one MI355X):", "",
             "```", command, "```", "",
             f"Kernel `{kernel}`; {pairs:.6g} pairs per launch; kernel time {kernel_ms:.3f} ms "
             f"(kernel-trace average of the same build on the same box).", "",
             "| counter | value (last dispatch) | per pair (x64 lanes) |", "|---|---|---|"]
    for c in sorted(vals):
        pp = vals[c] * 64.0 / pairs if c.startswith("SQ_INSTS") else None
        lines.append(f"| {c} | {vals[c]:.6g} | {'' if pp is None else f'{pp:.3f}'} |")
    lines += ["", "| derived | value |", "|---|---|"]
    for k, v in sm.items():
        if k.endswith("_per_pair") or k in ("pairs", "kernel_ms"):
            continue
        lines.append(f"| {k} | {v:.4g} |")
    lines += ["", "```json", json.dumps(rec), "```"]
    return "\n".join(lines) + "\n"


def load_record(workload: str, kernel: str, field: str):
    """``field`` of the JSON record under profiles/ for (workload, kernel), or None."""
    if not PROFILES.is_dir():
        return None
    for md in sorted(PROFILES.glob("*.md")):
        for block in re.findall(r"```json\n(.*?)\n```", md.read_text(), flags=re.S):
            try:
                rec = json.loads(block)
            except ValueError:
                continue
            if rec.get("workload") == workload and rec.get("kernel") == kernel and field in rec:
                return rec[field]
    return None


def load_traffic(workload: str, kernel: str):
    """traffic_bytes recorded under profiles/ for (workload, kernel), or None."""
    if not PROFILES.is_dir():
        return None
    for md in sorted(PROFILES.glob("*.md")):
        for block in re.findall(r"```json\n(.*?)\n```", md.read_text(), flags=re.S):
            try:
                rec = json.loads(block)
            except ValueError:
                continue
            if rec.get("workload") == workload and rec.get("kernel") == kernel:
                return float(rec["traffic_bytes"])
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("db")
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--workload", required=True)
    t.add_argument("--title", required=True)
    t.add_argument("--command", required=True)
    v = sub.add_parser("valu")
    v.add_argument("dbs", nargs="+")
    v.add_argument("--kernel", required=True)
    v.add_argument("--workload", required=True)
    v.add_argument("--pairs", type=float, required=True)
    v.add_argument("--flops-per-pair", type=float, required=True)
    g = v.add_mutually_exclusive_group(required=True)
    g.add_argument("--kernel-ms", type=float)
    g.add_argument("--kernel-ms-from", metavar="STATS_DB",
                   help="take the kernel's average duration from a --kernel-trace database")
    v.add_argument("--title", required=True)
    v.add_argument("--command", required=True)
    a = ap.parse_args(argv)
    if a.cmd == "valu":
        ms = a.kernel_ms
        if ms is None:
            hits = [avg for k, _, _, avg, _ in kernel_stats(a.kernel_ms_from) if a.kernel in k]
            if not hits:
                raise SystemExit(f"kernel {a.kernel!r} not in {a.kernel_ms_from}")
            ms = hits[-1] / 1000.0
        print(valu_markdown(a.dbs, a.kernel, a.workload, a.pairs, a.flops_per_pair, ms,
                            a.title, a.command), end="")
        return
    if a.cmd == "stats":
        print(stats_markdown(a.db, a.title, a.command), end="")
    else:
        print(traffic_markdown(a.fetch, a.write, a.kernel, a.copy_bytes, a.workload, a.title,
                               a.command), end="")


if __name__ == "__main__":
    main()
