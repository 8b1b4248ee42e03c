#!/usr/bin/env python3
"""Benchmark: GTEPS of the Indigo2 kernels on R-MAT scale 22 on one MI355X.

Workload (one "step"): SSSP from the highest-degree vertex and connected
components on an undirected R-MAT graph, scale 22, edge factor 16,
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), symmetrized, deduplicated, no self
loops, int32 CSR, int32 weights uniform in [1, 255] with both directions
equal (seed 1) -- the SSSP + CC configuration.  The graph is resident in HBM
before timing starts.  BFS (from the same source) is measured beside it.

value = whole-step throughput: 2 * m / (time of one SSSP + one CC), in
GTEPS with m the directed edge count (the reference's 1e-9 * edges / runtime
formula, bfs_vertex_cuda.h:163, applied to the two runs of a step).

roofline: the SSSP round kernel (sssp_do_round), the step's dominant kernel,
timed with HIP events on the library stream over the timed steps.  Its
algorithmic bytes are the style-independent SSSP figure, 12 bytes per
directed edge (neighbour id, weight, neighbour distance) and 12 per vertex
(row offsets, distance read and write): 12m + 12n + 4 per run, spread over
the run's launches.  achieved = those bytes / the summed launch time.  The
bytes the kernel's executed work touches (8 per edge scanned: id and weight;
4 per neighbour distance gathered; 12 per row visited; counted on the device)
are reported beside it, as is the copy-calibrated HBM traffic per launch
from rocprofv3.

cpu_baseline: the oracle's OpenMP restatement of the reference's fastest
OpenMP styles (SSSP and CC vertex-based topology-driven pull) on the same
R-MAT 22 graph.  Beside it, the style-matched restatements (SSSP edge-based
data-driven push and CC vertex-based topology-driven push, both serialised by
the reference's global critical section) on a bounded sample: the same
generator at scale 16; and the C++-threads restatement (codes/cpp family) on
the same R-MAT 22 graph.

selector: on the bench graph, for BFS, SSSP and CC, the learned selector's
pick (from features computed on the device, whose time is reported), the
library default and the exhaustive sweep's best style, each timed.

batched: the 512-graph batch (Erdos-Renyi and R-MAT halves, n log-uniform in
[2^10, 2^18], degree U[4, 32]) as one captured persistent launch for BFS, CC
and MIS, against one launch sequence per graph and algorithm.

usage: python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
ROOFLINE_KERNEL = "sssp_do_round"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def verify_bfs(g, src, dist):
    """Size-independent characterisation of BFS levels: dist[src] = 0, no edge
    spans more than one level, every reached vertex but src has a parent one
    level up, and unreached vertices have only unreached neighbours."""
    INF = 2147483647
    s = g.sp()
    d = dist.astype(np.int64)
    du, dv = d[s], d[g.nlist]
    assert dist[src] == 0
    reach_u, reach_v = du != INF, dv != INF
    assert np.all(reach_u == reach_v)
    assert np.all(np.abs(du[reach_u] - dv[reach_u]) <= 1)
    has_parent = np.zeros(g.nodes, bool)
    has_parent[g.nlist[reach_u & (dv == du + 1)]] = True
    assert np.all(has_parent[(d != INF) & (np.arange(g.nodes) != src)])


def verify_sssp(g, src, dist):
    """Shortest-path characterisation: no edge can relax, and every reached
    vertex but src has a tight incoming edge."""
    INF = 2147483647
    s = g.sp()
    d = dist.astype(np.int64)
    du, dv = d[s], d[g.nlist]
    w = g.eweight.astype(np.int64)
    assert dist[src] == 0
    ok = du != INF
    # every edge out of a reached vertex is relaxed; the graph is symmetric,
    # so this also rules out a reached neighbour of an unreached vertex
    assert np.all(dv[ok] <= du[ok] + w[ok])
    tight = np.zeros(g.nodes, bool)
    tight[g.nlist[ok & (dv == du + w)]] = True
    assert np.all(tight[(d != INF) & (np.arange(g.nodes) != src)])


def verify_cc(g, label):
    """Min-label characterisation: labels agree across every edge, label[v] <= v,
    labels are fixed points, and the number of distinct labels equals the
    number of components found by an independent union of the edge list."""
    s = g.sp()
    assert np.array_equal(label[s], label[g.nlist])
    assert np.all(label <= np.arange(g.nodes))
    assert np.array_equal(label[label], label)
    import scipy.sparse as sps
    import scipy.sparse.csgraph as csg
    A = sps.csr_matrix((np.ones(g.edges, np.int8), g.nlist, g.nindex), shape=(g.nodes, g.nodes))
    ncomp, _ = csg.connected_components(A, directed=False)
    assert np.unique(label).size == ncomp


def algorithmic_bytes(algo, n, m):
    """Style-independent bytes per run for int32 data (per PR iteration)."""
    return {"bfs": 8 * m + 12 * n + 4, "cc": 8 * m + 12 * n + 4, "sssp": 12 * m + 12 * n + 4,
            "mis": 12 * m + 12 * n + 4, "pr": 8 * m + 16 * n + 4, "tc": 4 * m + 4 * n + 4}[algo]


def pr_tc_section(args, D, G):
    """PageRank (pull, d = 0.85, stop at L1 residual < 1e-6 or 100 iterations)
    and triangle counting on R-MAT scale 20 and on a 2048 x 2048 grid where 5%
    of the vertices gain one random long-range edge."""
    out = {}
    graphs = (("rmat20", G.rmat(20, 16, seed=1)), ("grid2048", G.grid2d(2048, 0.05, seed=1)))
    D.set_pr_params(1e-6, 100)
    try:
        for tag, g in graphs:
            n, m = g.nodes, g.edges
            dg = D.DeviceGraph(g)
            buf = D.DeviceBuffer(max(4 * n, 8))
            for algo in ("pr", "tc"):
                style = D.styles(algo)[D.style_id(algo, None)]
                infos = [D.run_device(algo, dg, buf, style, 0) for _ in range(3)]
                med = float(np.median([i.seconds for i in infos]))
                it = infos[-1].iterations if algo == "pr" else 1
                b = algorithmic_bytes(algo, n, m) * max(it, 1)
                entry = {"style": style, "n": n, "m_directed": m, "seconds": round(med, 6),
                         "gteps": round(m * max(it, 1) / med * 1e-9, 3), "hbm_frac": round(b / med / (HBM_PEAK_GBS * 1e9), 4)}
                if algo == "pr":
                    entry["iterations"] = it
                    entry["residual"] = infos[-1].residual
                    if not args.no_verify:
                        import oracle as O
                        ref, _, _ = O.pr_fp64(g, fixed_iters=it)
                        got = buf.to_host(np.float32, n).astype(np.float64)
                        entry["max_rel_err_vs_fp64"] = float(np.max(np.abs(got - ref) / np.abs(ref)))
                else:
                    entry["triangles"] = int(infos[-1].count)
                    if not args.no_verify:
                        import oracle as O
                        assert O.tc_forward(g) == entry["triangles"]
                out[f"{algo}_{tag}"] = entry
            buf.free()
            dg.free()
    finally:
        D.set_pr_params(1e-4, 100)
    return out


TRAFFIC_FILE = os.path.join(REPO, "profiles", "sssp_rmat22.md")


def measured_traffic(scale):
    """HBM bytes per launch of the roofline kernel at scale 22, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated);
    None for other scales or when the summary is absent."""
    if scale != 22 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_style_matched(scale=16, reps=3):
    """The oracle's restatements of the style-matched reference OpenMP
    variants ([O-SSSP] edge data-driven push, [O-CC] vertex topology push) on
    R-MAT `scale`: median of `reps` runs each."""
    import oracle as O
    from indigo2suite_amd import graph as G
    g = G.rmat(scale, 16, seed=1).with_weights(1, 255, seed=1)
    src = g.max_degree_vertex()
    ts, tc = [], []
    for _ in range(reps):
        ts.append(O.sssp_omp_edge_data(g, src)[1])
        tc.append(O.cc_omp_push(g)[1])
    t1, t2 = float(np.median(ts)), float(np.median(tc))
    return {"value": round(2 * g.edges / (t1 + t2) * 1e-9, 5), "unit": "GTEPS",
            "sample": (f"R-MAT scale {scale} (n={g.nodes}, m={g.edges}): SSSP_OMP_E_Data_Push restatement "
                       f"{t1:.3f} s + CC_OMP_V_Topo_Push restatement {t2:.3f} s, median of {reps}")}


def cpu_cpp_threads(g, src, reps=3):
    """The oracle's C++-threads restatement of the reference's codes/cpp
    styles (SSSP vertex topology-driven pull, CC vertex topology-driven push,
    std::thread per round over a block partition) on the bench graph, median
    of `reps` runs each, as many threads as OpenMP uses."""
    import oracle as O
    threads = O.num_threads()
    ts, tc = [], []
    for _ in range(reps):
        _, t1, its = O.cpp_sssp(g, src, threads)
        _, t2, itc = O.cpp_cc(g, threads)
        ts.append(t1)
        tc.append(t2)
    t1, t2 = float(np.median(ts)), float(np.median(tc))
    return {"value": round(2 * g.edges / (t1 + t2) * 1e-9, 4), "unit": "GTEPS", "threads": threads,
            "sample": (f"same R-MAT graph: SSSP_CPP_V_Topo_Pull restatement {t1:.3f} s ({its} iterations) + "
                       f"CC_CPP_V_Topo_Push restatement {t2:.3f} s ({itc} iterations), median of {reps}")}


def selector_section(D, dg, g, src, algos=("bfs", "sssp", "cc")):
    """Learned pick vs the library default vs the exhaustive sweep's best, on
    the bench graph, for BFS and the step's algorithms.  The pick uses the
    features computed on the device (the graph is already resident); their
    cost is reported beside it."""
    from indigo2suite_amd import autotune, selector
    out = {}
    deg = g.degrees().astype(np.float64)
    dmax, sum_d2 = int(deg.max()), float(np.sum(deg * deg))
    feat_s = None
    for a in algos:
        pick, feat_s = selector.pick_style_device(dg, a)
        default = D.styles(a)[D.default_entry_style(a, 0)]
        sw = autotune.sweep(dg, a, src, repeats=3, max_degree=dmax, sum_d2=sum_d2)
        best = min(sw, key=sw.get)
        out[a] = {"pick": pick, "pick_s": round(sw[pick], 6), "default": default, "default_s": round(sw[default], 6),
                  "best": best, "best_s": round(sw[best], 6),
                  "regret_vs_best": round(sw[pick] / sw[best] - 1.0, 4),
                  "speedup_vs_default": round(sw[default] / sw[pick], 4), "styles_swept": len(sw)}
    out["features_device_s"] = round(feat_s, 6) if feat_s is not None else None
    return out


def cpu_baseline(scale, g22=None, src22=None, reps=5):
    """The oracle's OpenMP restatement of the fastest reference OpenMP styles
    (SSSP and CC vertex-based topology-driven pull) on the same workload,
    median of `reps` runs on all host threads (a few seconds in all)."""
    import oracle as O
    from indigo2suite_amd import graph as G
    if g22 is not None and scale == 22:
        g, src = g22, src22
    else:
        g = G.rmat(scale, 16, seed=1).with_weights(1, 255, seed=1)
        src = g.max_degree_vertex()
    ts, tc = [], []
    for _ in range(reps):
        _, t1, its = O.sssp_omp_pull(g, src)
        _, t2, itc = O.cc_omp_pull(g)
        ts.append(t1)
        tc.append(t2)
    t1, t2 = float(np.median(ts)), float(np.median(tc))
    value = 2 * g.edges / (t1 + t2) * 1e-9
    return {"value": round(value, 4), "unit": "GTEPS", "cores": O.num_threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": (f"R-MAT scale {scale} (n={g.nodes}, m={g.edges}, same generator, seed 1): oracle OpenMP "
                       f"restatement of SSSP_OMP_V_Topo_Pull ({t1:.3f} s, {its} iterations) + "
                       f"CC_OMP_V_Topo_Pull ({t2:.3f} s, {itc} iterations), median of {reps} runs each")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--cpu-scale", type=int, default=22)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-selector", action="store_true", help="skip the learned-pick vs default vs best section")
    ap.add_argument("--no-batched", action="store_true", help="skip the 512-graph batched-mode section")
    ap.add_argument("--no-extra", action="store_true", help="skip PageRank and TC on R-MAT 20 and the grid")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="also run one device-to-device copy of MIB MiB (the known-size kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding is a design, not part of this build)", file=sys.stderr)
        sys.exit(2)

    from indigo2suite_amd import device as D
    from indigo2suite_amd import graph as G

    D.init(0)
    dinfo = D.device_info()
    log("device", dinfo)
    t = time.time()
    g = G.rmat(args.scale, 16, seed=1).with_weights(1, 255, seed=1)
    src = g.max_degree_vertex()
    n, m = g.nodes, g.edges
    log(f"graph rmat{args.scale}: n={n} m={m} src={src} generated in {time.time() - t:.1f}s")
    dg = D.DeviceGraph(g)
    styles = {a: D.styles(a)[D.style_id(a, None)] for a in ("sssp", "cc", "bfs")}
    log("styles", styles)
    bufs = {a: D.DeviceBuffer(4 * n) for a in ("sssp", "cc", "bfs")}

    def step():
        i1 = D.run_device("sssp", dg, bufs["sssp"], styles["sssp"], src)
        i2 = D.run_device("cc", dg, bufs["cc"], styles["cc"], 0)
        return i1, i2

    for _ in range(args.warmup):
        step()
        D.run_device("bfs", dg, bufs["bfs"], styles["bfs"], src)

    # ---- timed region: K steps, library stream synchronised on both sides
    lib = D.lib()
    lib.igx_sync()
    lib.igx_profile_kernel(ROOFLINE_KERNEL.encode())
    per = {"sssp": [], "cc": []}
    iters = {"sssp": 0, "cc": 0}
    work_e = work_r = work_v = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        i1, i2 = step()
        per["sssp"].append(i1.seconds)
        per["cc"].append(i2.seconds)
        iters["sssp"], iters["cc"] = i1.iterations, i2.iterations
        work_e += i1.edges_processed
        work_r += i1.edges_relaxed
        work_v += i1.vertices_processed
    lib.igx_sync()
    t1 = time.perf_counter()
    import ctypes
    hook_ms = ctypes.c_double()
    hook_n = ctypes.c_int()
    lib.igx_profile_read(ctypes.byref(hook_ms), ctypes.byref(hook_n))
    lib.igx_profile_kernel(None)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 2 * m / ((t1 - t0) / args.steps) * 1e-9

    # BFS beside the step (not part of value)
    bfs_t = []
    for _ in range(args.steps):
        bfs_t.append(D.run_device("bfs", dg, bufs["bfs"], styles["bfs"], src).seconds)
    if not args.no_verify and args.steps > 0:
        # outputs of the last timed step and of the last BFS run
        t = time.time()
        verify_sssp(g, src, bufs["sssp"].to_host(np.int32, n))
        verify_cc(g, bufs["cc"].to_host(np.int32, n))
        verify_bfs(g, src, bufs["bfs"].to_host(np.int32, n))
        log(f"verified sssp/cc/bfs outputs in {time.time() - t:.1f}s")

    per_algo = {}
    for a, ts in (("sssp", per["sssp"]), ("cc", per["cc"]), ("bfs", bfs_t)):
        med = float(np.median(ts))
        b = algorithmic_bytes(a, n, m)
        per_algo[a] = {"style": styles[a], "seconds": round(med, 6), "gteps": round(m / med * 1e-9, 3),
                       "hbm_frac": round(b / med / (HBM_PEAK_GBS * 1e9), 4)}
    per_algo["sssp"]["iterations"] = iters["sssp"]

    # MIS on the same graph
    mis_buf = D.DeviceBuffer(4 * n)
    mis_style = D.styles("mis")[D.style_id("mis", None)]
    mis_t = [D.run_device("mis", dg, mis_buf, mis_style, 0).seconds for _ in range(args.steps)]
    med = float(np.median(mis_t))
    per_algo["mis"] = {"style": mis_style, "seconds": round(med, 6), "gteps": round(m / med * 1e-9, 3),
                       "hbm_frac": round(algorithmic_bytes("mis", n, m) / med / (HBM_PEAK_GBS * 1e9), 4)}
    if not args.no_verify:
        import oracle as O
        assert O.mis_verify(g, mis_buf.to_host(np.int32, n).astype(np.uint8)) == 0
    mis_buf.free()
    if not args.no_extra:
        per_algo.update(pr_tc_section(args, D, G))

    # the configuration's named SSSP style (edge-based, data-driven worklist)
    named = D.run_device("sssp", dg, bufs["sssp"], "e_data_push_thread", src)
    per_algo["sssp_e_data_push_thread"] = {"seconds": round(named.seconds, 6),
                                           "gteps": round(m / named.seconds * 1e-9, 3),
                                           "iterations": named.iterations}

    # roofline of the SSSP round kernel: the style-independent algorithmic
    # bytes of the timed runs (12m + 12n + 4 each) over the summed duration of
    # its launches in those runs
    launches = max(hook_n.value, 1)
    total_s = hook_ms.value * 1e-3
    alg_total = algorithmic_bytes("sssp", n, m) * args.steps
    achieved = alg_total / total_s * 1e-9 if total_s > 0 else 0.0
    # bytes of the work the launches executed, counted on the device
    exec_total = 8 * work_e + 4 * work_r + 12 * work_v + 4 * launches
    traffic = measured_traffic(args.scale)
    alg_per_launch = alg_total / launches
    roofline = {"bound": "hbm", "kernel": ROOFLINE_KERNEL, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "bytes_per_launch": int(alg_per_launch), "avg_launch_us": round(total_s / launches * 1e6, 2),
                "launches": hook_n.value, "launches_per_run": round(hook_n.value / max(args.steps, 1), 2),
                "traffic_over_algorithmic": round(traffic / alg_per_launch, 3) if traffic else None,
                "executed_bytes_per_launch": int(exec_total / launches),
                "executed_achieved": round(exec_total / total_s * 1e-9, 1) if total_s > 0 else 0.0,
                "edges_scanned_per_run": int(work_e / args.steps),
                "edges_relaxed_per_run": int(work_r / args.steps), "rows_visited_per_run": int(work_v / args.steps)}

    if args.copy_probe > 0:
        nb = args.copy_probe << 20
        a, b = D.DeviceBuffer(nb), D.DeviceBuffer(nb)
        D.check(lib.igx_memcpy_d2d(b.ptr, a.ptr, nb), "igx_memcpy_d2d")
        a.free()
        b.free()
        log(f"copy probe: {nb} bytes device to device")

    batched = None
    if not args.no_batched:
        t = time.time()
        from indigo2suite_amd import bench_batched
        batched = bench_batched.measure()
        log(f"batched section in {time.time() - t:.1f}s")

    sel = None
    if not args.no_selector:
        t = time.time()
        sel = selector_section(D, dg, g, src)
        log(f"selector section in {time.time() - t:.1f}s")

    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(args.cpu_scale, g, src)
        cpu["style_matched"] = cpu_style_matched()
        cpu["cpp_threads"] = cpu_cpp_threads(g, src)
        log(f"cpu baseline in {time.time() - t:.1f}s")

    line = {
        "metric": "GTEPS per algo on RMAT-22 & %HBM peak; learned-pick vs default/best style",
        "value": round(value, 3), "unit": "GTEPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded R-MAT, generated on the host)",
        "config": {"workload": f"SSSP (src = max-degree vertex) + CC on R-MAT scale {args.scale}, edge factor 16, "
                               f"a,b,c,d = 0.57,0.19,0.19,0.05, symmetric int32 CSR, weights U[1,255]",
                   "n": n, "m_directed": m, "parallelism": "single GPU"},
        "per_algo": per_algo,
        "roofline": roofline,
        "selector": sel,
        "batched": batched,
        "cpu_baseline": cpu,
        "device": dinfo["name"],
    }
    for b in bufs.values():
        b.free()
    dg.free()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
