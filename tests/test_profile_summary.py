"""profile_summary on synthetic rocprofv3 CSVs (CPU): counter calibration on
the copy probe, per-run and per-launch traffic, time and the algorithmic
bytes, checked against values worked out by hand for the inputs below."""
import csv
import json
import os

import pytest

from indigo2suite_amd import profile_summary as P

ROUND = "void igx::(anonymous namespace)::sssp_do_round(int const*, int const*, int*)"
FIN = "void igx::round_finalize<igx::(anonymous namespace)::FinSsspDo>(igx::Ctl*, int, int, FinSsspDo)"
COPY = "__amd_rocclr_copyBuffer"


def _write_csv(path, fields, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        w.writerows(rows)


def _inputs(tmp_path):
    """Two runs of SSSP (n = 1000, m = 8000): each run three round kernels of
    10 us and three finalize kernels of 2 us, then a 1 MiB copy probe.  The
    probe reads FETCH_SIZE 512 KiB (factor 0.5) and WRITE_SIZE 1024 KiB
    (factor 1); every round kernel reads FETCH 50 KiB and WRITE 10 KiB."""
    trace, fetch, write = [], [], []
    t, d = 1000, 0
    for _ in range(2):
        for k, us in ((ROUND, 10), (FIN, 2)) * 3:
            d += 1
            trace.append({"Kernel_Name": k, "Start_Timestamp": t, "End_Timestamp": t + us * 1000,
                          "Grid_Size_X": 1024})
            fetch.append({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": "FETCH_SIZE",
                          "Counter_Value": 50.0 if k == ROUND else 0.0, "Grid_Size": 1024})
            write.append({"Dispatch_Id": d, "Kernel_Name": k, "Counter_Name": "WRITE_SIZE",
                          "Counter_Value": 10.0 if k == ROUND else 0.0, "Grid_Size": 1024})
            t += 100000
    d += 1
    trace.append({"Kernel_Name": COPY, "Start_Timestamp": t, "End_Timestamp": t + 50000, "Grid_Size_X": 65536})
    fetch.append({"Dispatch_Id": d, "Kernel_Name": COPY, "Counter_Name": "FETCH_SIZE", "Counter_Value": 512.0,
                  "Grid_Size": 65536})
    write.append({"Dispatch_Id": d, "Kernel_Name": COPY, "Counter_Name": "WRITE_SIZE", "Counter_Value": 1024.0,
                  "Grid_Size": 65536})
    _write_csv(str(tmp_path / "trace" / "run_kernel_trace.csv"),
               ["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X"], trace)
    pf = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Grid_Size"]
    _write_csv(str(tmp_path / "fetch" / "run_counter_collection.csv"), pf, fetch)
    _write_csv(str(tmp_path / "write" / "run_counter_collection.csv"), pf, write)
    run = {"algo": "sssp", "graph": "synthetic", "style": "v_do_balanced", "n": 1000, "m": 8000, "runs": 2,
           "seconds": [4e-5, 4e-5], "median_s": 4e-5, "iterations": [3, 3], "copy_probe_mib": 1}
    (tmp_path / "run.json").write_text(json.dumps(run) + "\n")


def _values(text):
    out = {}
    for line in text.splitlines():
        if ": " in line and not line.startswith(("#", "|", "Counter", "Commands")):
            k, v = line.split(": ", 1)
            out[k.strip()] = v.strip()
    return out


def test_summary_figures_from_synthetic_counters(tmp_path):
    _inputs(tmp_path)
    out = tmp_path / "x.md"
    P.main(["--trace", str(tmp_path / "trace"), "--fetch", str(tmp_path / "fetch"), "--write",
            str(tmp_path / "write"), "--run-json", str(tmp_path / "run.json"), "--kernel", "sssp_do_round",
            "-o", str(out)])
    text = out.read_text()
    assert "FETCH_SIZE reads 0.5000 x and WRITE_SIZE 1.0000 x" in text
    v = _values(text)
    alg = 12 * 8000 + 12 * 1000 + 4                   # SSSP: 12m + 12n + 4 per run
    assert int(v["algorithmic_bytes_per_run"]) == alg
    # per run: three rounds of (50 KiB / 0.5 + 10 KiB / 1) = 3 * 112640 bytes
    assert int(v["traffic_bytes_per_run"]) == 3 * 112640
    assert float(v["kernel_seconds_per_run"]) == pytest.approx(36e-6)
    assert v["hbm_fraction"] == f"{alg / 36e-6 / 8e12:.4f}"          # printed to 4 places
    assert float(v["achieved_GBps (algorithmic bytes / kernel time)"]) == pytest.approx(alg / 36e-6 / 1e9, abs=0.05)
    assert int(v["traffic_bytes_per_launch"]) == 112640
    assert int(v["algorithmic_bytes_per_launch"]) == alg * 2 // 6
    assert float(v["avg_launch_seconds"]) == pytest.approx(10e-6)
    # the table: the copy probe is not a row, the round kernel has 3 launches per run
    assert "| `sssp_do_round` | 3.0 | 30.0 | 10.0 | 0.3 | 0.0 |" in text
    assert COPY not in text.split("|---|")[1]


def test_algorithmic_bytes_closed_forms():
    n, m = 10, 20
    assert P.algorithmic_bytes("bfs", n, m) == 8 * m + 12 * n + 4
    assert P.algorithmic_bytes("cc", n, m) == 8 * m + 12 * n + 4
    assert P.algorithmic_bytes("mis", n, m) == 12 * m + 12 * n + 4
    assert P.algorithmic_bytes("tc", n, m) == 4 * m + 4 * n + 4
    # PageRank per iteration; the other algorithms ignore the iteration count
    assert P.algorithmic_bytes("pr", n, m, 5) == 5 * (8 * m + 16 * n + 4)
    assert P.algorithmic_bytes("sssp", n, m, 5) == 12 * m + 12 * n + 4


def test_missing_copy_probe_is_an_error(tmp_path):
    rows = [(1, "sssp_do_round", 50.0, 1024)]
    with pytest.raises(ValueError):
        P.calibration(rows, 1 << 20, 1.0)
