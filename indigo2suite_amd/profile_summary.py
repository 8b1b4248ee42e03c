"""Summarise rocprofv3 output of `profile_run` into a markdown table.

Inputs: the kernel-trace directory (`--kernel-trace --output-format csv`),
the FETCH_SIZE and WRITE_SIZE counter directories (`--pmc FETCH_SIZE`,
`--pmc WRITE_SIZE`, each with `--copy-probe`), and the JSON line that
profile_run printed in the traced run.  The counters are scaled by what they
read for the copy probe (a device-to-device copy of known size reads its
bytes once and writes them once), so the traffic figures are HBM bytes.

Per run of the algorithm it reports every kernel's launches, time and
traffic, the algorithmic bytes of the run (int32 data: BFS and CC
8m + 12n + 4, SSSP and MIS 12m + 12n + 4, PageRank 8m + 16n + 4 per
iteration, TC 4m + 4n + 4), the achieved fraction of the 8 TB/s HBM peak and
the ratio of measured traffic to algorithmic bytes.  `--kernel K` adds the
per-launch figures of one kernel (the "traffic_bytes_per_launch" line that
bench.py reads for its roofline kernel).

  python -m indigo2suite_amd.profile_summary --trace T --fetch F --write W --run-json run.json \\
      [--kernel sssp_do_round] [--command "..."] -o profiles/x.md
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys

HBM_PEAK = 8.0e12


def algorithmic_bytes(algo: str, n: int, m: int, iterations: int = 1) -> int:
    per = {"bfs": 8 * m + 12 * n + 4, "cc": 8 * m + 12 * n + 4, "sssp": 12 * m + 12 * n + 4,
           "mis": 12 * m + 12 * n + 4, "pr": 8 * m + 16 * n + 4, "tc": 4 * m + 4 * n + 4}[algo]
    return per * (iterations if algo == "pr" else 1)


def short_name(k: str) -> str:
    k = k.replace("(anonymous namespace)::", "")
    k = re.sub(r"^void ", "", k)
    k = k.split("(")[0]
    k = re.sub(r"^igx::", "", k)
    return re.sub(r"<.*", "", k)


def _one(path_glob: str) -> str:
    files = sorted(glob.glob(path_glob))
    if not files:
        raise FileNotFoundError(path_glob)
    return files[0]


def dispatches(trace_dir: str):
    rows = []
    with open(_one(os.path.join(trace_dir, "*kernel_trace.csv"))) as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), short_name(r["Kernel_Name"]),
                         (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9, int(r["Grid_Size_X"])))
    rows.sort()
    return rows


def counter(pmc_dir: str, name: str):
    out = []
    with open(_one(os.path.join(pmc_dir, "*counter_collection.csv"))) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == name:
                out.append((int(r["Dispatch_Id"]), short_name(r["Kernel_Name"]), float(r["Counter_Value"]),
                            int(r["Grid_Size"])))
    out.sort()
    return out


def calibration(rows, probe_bytes: int, per_way: float):
    """Scale factor of a counter: its reading (KiB) on the copy probe, the
    largest runtime copy dispatch, over the bytes the probe moves one way."""
    probes = [r for r in rows if r[1].startswith("__amd_rocclr_copyBuffer")]
    if not probes:
        raise ValueError("no copy probe dispatch in the counter run (use --copy-probe)")
    best = max(probes, key=lambda r: r[3])
    return best[2] * 1024.0 / (probe_bytes * per_way)


def summarise(args) -> str:
    run = [json.loads(x) for x in open(args.run_json).read().splitlines() if x.startswith("{")][-1]
    algo, n, m, runs = run["algo"], run["n"], run["m"], run["runs"]
    iters = run["iterations"]
    probe = run.get("copy_probe_mib", 0) << 20
    disp = [d for d in dispatches(args.trace) if not d[1].startswith("__amd_rocclr")]
    fetch = counter(args.fetch, "FETCH_SIZE")
    write = counter(args.write, "WRITE_SIZE")
    probe = (probe or 256 << 20)
    fs = calibration(fetch, probe, 1.0)
    ws = calibration(write, probe, 1.0)
    fetch = [r for r in fetch if not r[1].startswith("__amd_rocclr")]
    write = [r for r in write if not r[1].startswith("__amd_rocclr")]
    t = collections.OrderedDict()
    for _, k, s, _ in disp:
        e = t.setdefault(k, {"calls": 0, "s": 0.0, "fetch": 0.0, "write": 0.0})
        e["calls"] += 1
        e["s"] += s
    for _, k, v, _ in fetch:
        if k in t:
            t[k]["fetch"] += v * 1024.0 / fs
    for _, k, v, _ in write:
        if k in t:
            t[k]["write"] += v * 1024.0 / ws
    it_run = int(round(sum(iters) / len(iters)))
    alg = algorithmic_bytes(algo, n, m, it_run)
    tot_s = sum(e["s"] for e in t.values()) / runs
    tot_b = sum(e["fetch"] + e["write"] for e in t.values()) / runs
    lines = [f"# {algo.upper()} on {run['graph']} (n = {n}, m = {m} directed), style `{run['style']}`: kernel time "
             f"and HBM traffic", ""]
    lines += ["Commands (repo root, one MI355X):", "", "```"]
    lines += args.command or ["(not given)"]
    lines += ["```", ""]
    lines.append(f"Counter calibration on the {probe >> 20} MiB device-to-device copy probe: FETCH_SIZE reads "
                 f"{fs:.4f} x and WRITE_SIZE {ws:.4f} x the bytes moved; readings below are divided by these "
                 f"factors.  {runs} runs; figures per run unless marked per launch; median event time of a run "
                 f"{run['median_s'] * 1e3:.3f} ms, {it_run} iterations.")
    lines += ["", "| kernel | launches / run | time / run (us) | avg us / launch | fetch MB / run | "
              "write MB / run |", "|---|---:|---:|---:|---:|---:|"]
    for k, e in sorted(t.items(), key=lambda kv: -kv[1]["s"]):
        lines.append(f"| `{k}` | {e['calls'] / runs:.1f} | {e['s'] / runs * 1e6:.1f} | "
                     f"{e['s'] / max(e['calls'], 1) * 1e6:.1f} | {e['fetch'] / runs / 1e6:.1f} | "
                     f"{e['write'] / runs / 1e6:.1f} |")
    lines += ["", f"algorithmic_bytes_per_run: {alg}",
              f"kernel_seconds_per_run: {tot_s:.6e}",
              f"traffic_bytes_per_run: {int(tot_b)}",
              f"achieved_GBps (algorithmic bytes / kernel time): {alg / tot_s / 1e9:.1f}" if tot_s > 0 else "",
              f"hbm_fraction: {alg / tot_s / HBM_PEAK:.4f}" if tot_s > 0 else "",
              f"traffic / algorithmic: {tot_b / alg:.3f}"]
    if args.kernel:
        e = t.get(args.kernel)
        if e is None:
            raise KeyError(f"kernel {args.kernel} not in the trace: {list(t)}")
        calls = e["calls"]
        lines += ["", f"Per launch of `{args.kernel}` ({calls / runs:.1f} launches per run):", "",
                  f"traffic_bytes_per_launch: {int((e['fetch'] + e['write']) / calls)}",
                  f"algorithmic_bytes_per_launch: {int(alg * runs / calls)}",
                  f"avg_launch_seconds: {e['s'] / calls:.6e}",
                  f"traffic / algorithmic: {(e['fetch'] + e['write']) / (alg * runs):.3f}"]
    return "\n".join(x for x in lines if x is not None) + "\n"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--trace", required=True)
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--run-json", required=True)
    ap.add_argument("--kernel", default=None)
    ap.add_argument("--command", action="append", default=[], help="command line to quote (repeatable)")
    ap.add_argument("-o", "--output", required=True)
    a = ap.parse_args(argv)
    text = summarise(a)
    with open(a.output, "w") as f:
        f.write(text)
    print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
