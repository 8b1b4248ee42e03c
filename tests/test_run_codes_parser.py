"""The driver's log parser reads every result-line form the reference's
harnesses and variants print (CUDA, OpenMP and C++-threads), and splits a
driver log into one record per (program, input)."""
from indigo2suite_amd import run_codes as R

# the reference's printf formats, filled in (file:line of each form)
SSSP_CUDA = "\n".join([
    "(sssp_edge_cuda.cu)", "Generating weights.", "input: g.egr", "nodes: 3", "edges: 4", "source: 0",
    "iterations: 2", "iterations: 2", "iterations: 3",
    "GPU runtime: %.6f s" % 0.000123,                      # sssp_edge_cuda.h:175
    "GPU Throughput: %.6f gigaedges/s" % 0.032520,         # sssp_edge_cuda.h:176
    "vertex 2 has maximum distance 4", "verification passed", ""])
BFS_CUDA_SLOW = "\n".join([
    "bfs topology-driven CUDA (x.cu)", "iterations: 9",
    "runtime: %.6fs" % 31.5,                               # bfs_vertex_cuda.h:154 (a run of 30 s or more)
    "Throughput: %.6f gigaedges/s" % 0.000001,
    "vertex 7 has maximum distance 8", "turn off verification"])
MIS_EDGE_CUDA = "\n".join([
    "mis edge-based CUDA (x.cu)",
    "iterations: %d,  elements in set: %d (%.1f%%)" % (4, 120, 46.9),
    "iterations: %d,  elements in set: %d (%.1f%%)" % (4, 120, 46.9),
    "iterations: %d,  elements in set: %d (%.1f%%)" % (5, 120, 46.9),
    "runtime: %.6fs" % 0.002,
    "Throughput: %.6f gigabytes/s" % 0.5])                  # mis_edge_cuda.h:171
PR_CUDA = "\n".join([
    "PageRank CUDA v0.1 (x.cu)", "Copyright 2022 Texas State University", "", "graph: g.egr",
    "GPU iterations = 17.", "GPU runtime: %.6fs" % 0.0031, "", "Throughput: %.6f gigaedges/s" % 1.2,
    "All good."])
PR_CPU = "\n".join(["CPU iterations = 17.", "CPU runtime: %.6fs" % 0.5, "", "Total Error: %f" % 0.0002])
TC_OMP = "\n".join(["Triangle counting vertex-centric (x.cpp)", "max degree: 12", "", "Threads: 16",
                    "OMP runtime: %.6f s" % 0.25, "OMP Throughput: %.6f gigaedges/s" % 0.01,
                    "the pattern occurs %d times" % 4235, ""])
CC_CPP = "\n".join(["cc vertex-based CPP (x.cpp)", "iterations: 4", "CPP runtime: %.6fs" % 0.04,
                    "CPP Throughput: %.6f gigaedges/s" % 0.2, "number of connected components: 31",
                    "verification passed"])
TC_MISMATCH = "GPU runtime: 0.000100 s\nGPU Throughput: 0.1 gigaedges/s\nERROR: host 3 device 4"


def test_each_form():
    r = R.parse_program_output(SSSP_CUDA)
    assert r["runtime_s"] == 0.000123 and r["runtime_kind"] == "GPU" and r["throughput"] == 0.03252
    assert r["iterations"] == [2, 2, 3] and r["max_distance"] == 4 and r["verified"] and r["nodes"] == 3
    r = R.parse_program_output(BFS_CUDA_SLOW)
    assert r["runtimes"] == [31.5] and r["verify_off"] and not r["verified"]
    r = R.parse_program_output(MIS_EDGE_CUDA)
    assert r["iterations"] == [4, 4, 5] and r["set_sizes"] == [120] * 3
    assert r["throughput_unit"] == "gigabytes/s" and r["throughput"] == 0.5
    r = R.parse_program_output(PR_CUDA)
    assert r["iterations"] == [17] and r["runtime_s"] == 0.0031 and r["pr_ok"] is True
    r = R.parse_program_output(PR_CPU)
    assert r["runtime_kind"] == "CPU" and r["pr_ok"] is False and r["pr_error"] == 0.0002
    r = R.parse_program_output(TC_OMP)
    assert r["runtime_kind"] == "OMP" and r["triangles"] == 4235 and r["threads"] == 16 and r["max_degree"] == 12
    r = R.parse_program_output(CC_CPP)
    assert r["runtime_kind"] == "CPP" and r["components"] == 31 and r["iterations"] == [4]
    r = R.parse_program_output(TC_MISMATCH)
    assert r["errors"] == ["host 3 device 4"] and "triangles" not in r


def test_driver_log_records_and_csv():
    log = "\n".join([
        "num_codes: 2", "num_inputs: 2", "", "code_path: c/", "input_path: i/", "",
        "compiling SSSP_A.cu", "a.egr: 1 out of 2 inputs", SSSP_CUDA, "b.egr: 2 out of 2 inputs",
        "exit status -11 for c/ on b.egr",
        "compiling MIS_B.cu", "a.egr: 1 out of 2 inputs", MIS_EDGE_CUDA, "b.egr: 2 out of 2 inputs", MIS_EDGE_CUDA])
    recs = R.parse_driver_log(log)
    assert [(r["code"], r["input"]) for r in recs] == [("SSSP_A.cu", "a.egr"), ("SSSP_A.cu", "b.egr"),
                                                       ("MIS_B.cu", "a.egr"), ("MIS_B.cu", "b.egr")]
    assert recs[0]["verified"] and recs[1]["exit_status"] == -11 and "runtime_s" not in recs[1]
    assert recs[3]["set_sizes"] == [120] * 3
    csv = R.records_csv(recs).splitlines()
    assert csv[0].startswith("code,input,exit_status") and len(csv) == 5
    assert csv[3].split(",")[0] == "MIS_B.cu" and "4;4;5" in csv[3]


def test_harness_headers_print_the_parsed_forms():
    """Each HIP harness header prints its reference harness's result-line
    form (checked on the header text: the parser must read what they print)."""
    import os
    inc = R.INCLUDE
    forms = {"sssp_vertex_hip.h": "kGpuSpaced", "sssp_edge_hip.h": "kGpuSpaced", "tc_vertex_hip.h": "kGpuSpaced",
             "tc_edge_hip.h": "kGpuSpaced", "cc_vertex_hip.h": "kCc", "cc_edge_hip.h": "kCc",
             "mis_vertex_hip.h": "kPlain", "mis_edge_hip.h": "kMisEdge"}
    for h, form in forms.items():
        assert f"report({form}," in open(os.path.join(inc, h)).read(), h
    for h in ("bfs_vertex_hip.h", "bfs_edge_hip.h"):
        assert "timed_runs_with_cutoff" in open(os.path.join(inc, h)).read()
    hx = open(os.path.join(inc, "igx_harness.hpp")).read()
    for line in ('"GPU runtime: %.6f s\\nGPU Throughput: %.6f gigaedges/s\\n"',
                 '"runtime: %.6fs\\nThroughput: %.6f gigabytes/s\\n"',
                 '"runtime: %.6fs\\nGPU Throughput: %.6f gigaedges/s\\n"'):
        assert line in hx
