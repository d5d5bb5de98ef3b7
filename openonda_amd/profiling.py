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
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        print(stats_markdown(a.db, a.title, a.command), end="")
    else:
        print(traffic_markdown(a.fetch, a.write, a.kernel, a.copy_bytes, a.workload, a.title,
                               a.command), end="")


if __name__ == "__main__":
    main()
