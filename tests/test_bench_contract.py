"""bench.py's contract on the CPU: the one-GPU guard and the algorithmic
bytes the roofline divides by (hand-computed for the cross-hole batch)."""
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr


def test_algorithmic_bytes_of_a_time_blocked_launch():
    import bench
    # 40 B per cell-update (Ez, Hx, Hy read + write, three coefficients, the
    # snapshot) and 16 B per CPML cell and direction; C2: 400 x 200 cells,
    # 20-cell CPML on all sides, 13 shots, 6 steps per launch.
    cells = 400 * 200
    psi = 2 * 20 * 200 + 2 * 20 * 400
    per_step = 13 * (40 * cells + 16 * psi)
    assert bench.step_bytes(13, 400, 200, 20) == per_step
    assert bench.kernel_bytes(5, 13, 400, 200, 20, 80, 6) == 6 * per_step == 279_552_000
    assert bench.kernel_bytes(6, 13, 400, 200, 20, 80, 6) == 279_552_000


def test_defaults():
    import bench
    a = bench.parse([])
    assert (a.gpus, a.steps, a.warmup, a.shot_streams, a.time_block) == (1, 5, 1, 3, "auto")
    assert a.workload == "c3" and not a.quick


def test_measured_traffic_is_keyed_by_workload_and_kernel(tmp_path):
    """Counter bytes are reported only for the exact workload and kernel the
    counter runs measured."""
    import json
    import bench
    f = tmp_path / "traffic.json"
    f.write_text(json.dumps({"measured": {"c2_crosshole_400x200_13shots|k_tb<64,1024,6,true>":
                                          {"bytes_per_launch": 123}}}))
    assert bench.measured_traffic(f, "c2_crosshole_400x200_13shots|k_tb<64,1024,6,true>") == 123
    assert bench.measured_traffic(f, "c2_crosshole_400x200_13shots|k_tb<64,1024,8,true>") is None
    assert bench.measured_traffic(f, "other|k_tb<64,1024,6,true>") is None
    assert bench.measured_traffic(tmp_path / "missing.json", "x|y") is None


def test_kernel_busy_takes_the_union_inside_the_window(tmp_path):
    import sqlite3
    import bench
    db = tmp_path / "t.db"
    con = sqlite3.connect(db)
    con.execute("create table kernels (name text, start integer, end integer)")
    rows = [("void k_a(int)", 100, 200), ("void k_b(float)", 150, 250),   # overlap: 150 busy
            ("void k_a(int)", 300, 350),                                   # 50 busy
            ("void k_a(int)", 50, 120),                                    # starts before t0
            ("void k_b(float)", 900, 1100)]                                # ends after t1
    con.executemany("insert into kernels values (?, ?, ?)", rows)
    con.commit()
    con.close()
    r = bench.kernel_busy(str(db), 100, 1000)
    assert r["launches"] == 3
    assert abs(r["busy_s"] - 200e-9) < 1e-15 and abs(r["busy_fraction"] - 200 / 900) < 1e-12
    assert r["kernels"]["k_a"] == {"launches": 2, "total_s": round(150e-9, 4), "avg_us": 0.07}
    assert r["kernels"]["k_b"]["launches"] == 1


def test_pmc_summary_scales_counters_by_the_copy_calibration(tmp_path):
    """Counter KiB per dispatch of the probe's kernels, scaled by what the
    same counter read for the copies of known size, keyed by workload; other
    grid sizes of the same kernel are not attributed."""
    import bench
    head = ("Correlation_Id,Dispatch_Id,Agent_Id,Queue_Id,Process_Id,Thread_Id,Grid_Size,"
            "Kernel_Id,Kernel_Name,Workgroup_Size,LDS_Block_Size,Scratch_Size,VGPR_Count,"
            "Accum_VGPR_Count,SGPR_Count,Counter_Name,Counter_Value,Start_Timestamp,End_Timestamp\n")
    g3 = 15 * 24 * 12 * 1024
    copy_kib = bench.COPY_BYTES / 1024 / 2            # counter reads half the true bytes

    def write(d, counter, kv):
        d.mkdir()
        rows = [f'{i},{i},1,1,1,1,{g},0,"{n}",256,0,0,8,0,16,{counter},{v},0,1\n'
                for i, (n, g, v) in enumerate(kv)]
        (d / "1_counter_collection.csv").write_text(head + "".join(rows))
    write(tmp_path / "f", "FETCH_SIZE",
          [("__amd_rocclr_copyBuffer", 1024, copy_kib / 3)] * 3 +
          [("void (anonymous namespace)::k_tb_multi<64, 1024, 6, true>(unsigned char const*, int)", g3, 1000.0)] * 2 +
          [("void (anonymous namespace)::k_tb_multi<64, 1024, 6, true>(unsigned char const*, int)", 999, 5e9)] +
          [("__amd_rocclr_copyBuffer", 1024, 7e9)])       # a later copy: not calibration
    write(tmp_path / "w", "WRITE_SIZE",
          [("__amd_rocclr_copyBuffer", 1024, copy_kib / 3)] * 3 +
          [("(anonymous namespace)::k_tb_multi<64, 1024, 6, true>(unsigned char const*, int)", g3, 500.0)] * 2)
    res = bench.pmc_summary(tmp_path / "f", tmp_path / "w", out=tmp_path / "t.json")
    m = res["measured"]["c3_model3_1200x600_15shots|k_tb_multi<64,1024,6,true>"]
    assert abs(res["calibration"]["fetch"]["factor"] - 2.0) < 1e-12
    assert m["read_bytes"] == 2 * 1000 * 1024 and m["write_bytes"] == 2 * 500 * 1024
    assert m["bytes_per_launch"] == 3 * 1000 * 1024
    assert bench.measured_traffic(tmp_path / "t.json",
                                  "c3_model3_1200x600_15shots|k_tb_multi<64,1024,6,true>") == 3072000


def test_default_line_merges_sections_and_survives_a_failed_one(monkeypatch, capsys):
    """The default run starts one child per section and merges their JSON
    lines: a section that fails leaves an error in its key, one that printed
    its result before failing keeps it with its exit status, and the plain
    path is handed the headline value."""
    import json
    import subprocess as sp
    import bench
    seen = []

    class R:
        def __init__(self, rc, out):
            self.returncode, self.stdout = rc, out

    def fake_run(argv, stdout=None, text=None, timeout=None):
        sec = argv[argv.index("--section") + 1]
        seen.append((sec, argv))
        if sec == "c3":
            return R(0, "noise\n" + json.dumps({"metric": bench.METRIC, "value": 200.0}) + "\n")
        if sec == "one":
            return R(134, json.dumps({"value": 100.0}) + "\n")
        if sec == "c2":
            return R(1, "")
        return R(0, json.dumps({"value": 1.0}) + "\n")

    monkeypatch.setattr(sp, "run", fake_run)
    monkeypatch.setattr(bench.sys, "argv", ["bench.py"])
    status = bench.run_sections(bench.parse([]))
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert status == 0 and out["value"] == 200.0
    assert out["one_model_at_a_time"] == {"value": 100.0, "exit_status": 134,
                                          "batched_speedup": 2.0}
    assert "error" in out["c2_crosshole"]
    assert out["plain_path"] == {"value": 1.0} and out["serving"] == {"value": 1.0}
    assert out["notebook_inversion"] == {"value": 1.0} and out["c4_memory_bound"] == {"value": 1.0}
    plain = dict(seen)["plain"]
    assert plain[plain.index("--ref-value") + 1] == "200.0"
    assert [s for s, _ in seen] == ["c3", "one", "c2", "plain", "serve", "notebook", "c4"]
