"""The rocprofv3 summary module on synthetic result databases (no GPU)."""
import json
import sqlite3

import pytest

from openonda_amd import profiling as PR


def _db(path, counter, rows):
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (dispatch_id integer, kernel_name text, "
                "counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)",
                    [(i, k, counter, v) for i, (k, v) in enumerate(rows)])
    con.execute("create table kd (name text, total_calls integer, total_duration real, "
                "percentage real)")
    con.execute("create view top_kernels as select * from kd")
    con.executemany("insert into kd values (?, ?, ?, ?)",
                    [("vpm::(anonymous namespace)::gather_kernel(x)", 3, 300.0, 1.0),
                     ("void fused_kernel<float, 3u, true, 256, 1, 256, 64, 0>(FusedArgs)", 3,
                      2.4e6, 99.0)])
    con.commit()
    con.close()


def test_traffic_calibration_and_lookup(tmp_path, monkeypatch):
    kern = "fused_kernel<float, 3u, true, 256, 1, 256, 64, 0>"
    copy = 1 << 28
    # FETCH_SIZE reads half the true bytes of the copy, WRITE_SIZE all of them
    _db(tmp_path / "f.db", "FETCH_SIZE", [("void " + kern + "(FusedArgs)", 1000.0),
                                          ("__amd_rocclr_copyBuffer", copy / 1024 / 2)])
    _db(tmp_path / "w.db", "WRITE_SIZE", [("void " + kern + "(FusedArgs)", 300.0),
                                          ("__amd_rocclr_copyBuffer", copy / 1024)])
    t = PR.traffic(str(tmp_path / "f.db"), str(tmp_path / "w.db"), kern, copy)
    assert t["fetch_factor"] == pytest.approx(2.0) and t["write_factor"] == pytest.approx(1.0)
    assert t["traffic_bytes"] == pytest.approx((2 * 1000.0 + 300.0) * 1024)
    md = PR.traffic_markdown(str(tmp_path / "f.db"), str(tmp_path / "w.db"), kern, copy,
                             "cloud", "title", "cmd")
    (tmp_path / "t.md").write_text(md)
    monkeypatch.setattr(PR, "PROFILES", tmp_path)
    assert PR.load_traffic("cloud", kern) == pytest.approx(t["traffic_bytes"])
    assert PR.load_traffic("cloud", "other") is None
    rec = json.loads(md.split("```json\n")[1].split("\n```")[0])
    assert rec["kernel"] == kern


def test_kernel_stats(tmp_path):
    _db(tmp_path / "k.db", "FETCH_SIZE", [])
    rows = PR.kernel_stats(str(tmp_path / "k.db"))
    assert rows[0][0].startswith("void fused_kernel") and rows[0][3] == pytest.approx(8e5)
    assert rows[1][0].startswith("gather_kernel") and rows[1][1] == 3
    md = PR.stats_markdown(str(tmp_path / "k.db"), "T", "C")
    assert "| 3 | 2400000.0 | 800000.0 |" in md


def _pmc_db(path, rows):
    """counters_collection rows (kernel, counter, value), one dispatch per row."""
    con = sqlite3.connect(path)
    con.execute("create table counters_collection (dispatch_id integer, kernel_name text, "
                "counter_name text, value real)")
    con.executemany("insert into counters_collection values (?, ?, ?, ?)",
                    [(i, k, c, v) for i, (k, c, v) in enumerate(rows)])
    con.commit()
    con.close()


def test_valu_summary_from_counters(tmp_path):
    """Per-pair instruction counts, executed FP32 flops (FMA 2, others 1), the
    clock from GRBM_GUI_ACTIVE over eight XCDs and the issue rate, by hand."""
    kern = "fused_kernel<float, 3u, true, 256, 1, 256, 64, 0>"
    full = "void " + kern + "(FusedArgs)"
    pairs, ms = 6.4e9, 100.0
    # 29 wave instructions per 64 pairs: 19 FMA, 6 MUL, 3 ADD, 1 TRANS
    w = pairs / 64.0
    _pmc_db(tmp_path / "p1.db", [
        (full, "SQ_INSTS_VALU", 29 * w), (full, "SQ_INSTS_VALU_FMA_F32", 19 * w),
        (full, "SQ_INSTS_VALU_MUL_F32", 6 * w), (full, "SQ_INSTS_VALU_ADD_F32", 3 * w),
        (full, "SQ_INSTS_VALU_TRANS_F32", 1 * w), ("other_kernel", "SQ_INSTS_VALU", 1.0),
        (full, "GRBM_GUI_ACTIVE", 8 * 2.0e9 * ms * 1e-3)])
    _pmc_db(tmp_path / "p2.db", [(full, "SQ_ACTIVE_INST_VALU", 100.0),
                                 (full, "SQ_ACTIVE_INST_VALU2", 25.0),
                                 (full, "SQ_WAIT_ANY", 10.0), (full, "SQ_WAVE_CYCLES", 40.0)])
    vals = PR.kernel_counters([str(tmp_path / "p1.db"), str(tmp_path / "p2.db")], kern)
    assert "other_kernel" not in vals and vals["SQ_INSTS_VALU"] == 29 * w
    sm = PR.valu_summary(vals, pairs, 53, ms)
    assert sm["SQ_INSTS_VALU_per_pair"] == pytest.approx(29.0)
    assert sm["exec_fp32_flop_per_pair"] == pytest.approx(2 * 19 + 6 + 3 + 1)
    assert sm["clock_ghz"] == pytest.approx(2.0)
    assert sm["valu_wave_instr_per_clk_per_cu"] == pytest.approx(29 * w / (2.0e8 * 256))
    assert sm["alg_tflops"] == pytest.approx(53 * pairs / 0.1 / 1e12)
    assert sm["dual_issue_share"] == pytest.approx(0.25)
    assert sm["wait_any_share"] == pytest.approx(0.25)
    # the CLI takes the kernel time from a kernel-trace database
    _db(tmp_path / "st.db", "FETCH_SIZE", [])
    out = tmp_path / "v.md"
    import contextlib
    import io
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        PR.main(["valu", str(tmp_path / "p1.db"), str(tmp_path / "p2.db"), "--kernel", kern,
                 "--workload", "cloud", "--pairs", str(pairs), "--flops-per-pair", "53",
                 "--kernel-ms-from", str(tmp_path / "st.db"), "--title", "T", "--command", "C"])
    out.write_text(buf.getvalue())
    rec = json.loads(out.read_text().split("```json\n")[1].split("\n```")[0])
    assert rec["kernel_ms"] == pytest.approx(800.0)  # 2.4e6 us over 3 calls
    assert rec["exec_fp32_flop_per_pair"] == pytest.approx(48.0)
