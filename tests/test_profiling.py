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
