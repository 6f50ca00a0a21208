"""Writes the small fixtures under tests/golden/ from the reference checkout's
result files (data only: CSV headers and per-query trace counters of its
published SIFT1M runs, results/compare_run/).  Run in the development
container: python tests/golden/make_fixtures.py [/root/reference]."""
import csv
import json
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
RUNS = os.path.join(REF, "results", "compare_run", "runs", "repro_gpu")
CPU_RUNS = os.path.join(REF, "results", "compare_run", "runs", "repro_cpu_raw")
FIELDS = ["seed_count", "visited_nodes", "expanded_nodes", "neighbor_slots_scanned", "candidate_nodes",
          "distance_evals", "frontier_pushes", "result_updates", "peak_frontier", "peak_results", "result_count",
          "early_stopped"]


def header(path):
    with open(path) as f:
        return f.readline().strip().split(",")


def main():
    heads = {
        "sweep_summary": header(os.path.join(RUNS, "sweep_summary.csv")),
        "batch_stats": header(os.path.join(RUNS, "repro_gpu_raw_pq0050_batch_stats.csv")),
        "query_stats": header(os.path.join(RUNS, "repro_gpu_raw_pq0050_query_stats.csv")),
        "cpu_sweep_summary": header(os.path.join(CPU_RUNS, "sweep_summary.csv")),
        "cpu_query_stats": header(os.path.join(CPU_RUNS, "pq0050_query_stats.csv")),
        "cpu_json_keys": sorted(json.load(open(os.path.join(REF, "results", "compare_run",
                                                            "repro_cpu_raw.json"))).keys()),
    }
    json.dump(heads, open(os.path.join(HERE, "reference_csv_headers.json"), "w"), indent=1)
    out = {}
    for method in ("repro_gpu_raw", "repro_gpu_hash"):
        for pq in (10, 20, 30, 40, 50):
            rows = list(csv.DictReader(open(os.path.join(RUNS, f"{method}_pq{pq:04d}_query_stats.csv"))))
            out[f"{method}_pq{pq:04d}"] = np.array([[int(r[f]) for f in FIELDS] for r in rows], np.int32)
    for pq in (10, 20, 30, 40, 50):  # the CPU search (bench_cpu) on the same data
        rows = list(csv.DictReader(open(os.path.join(CPU_RUNS, f"pq{pq:04d}_query_stats.csv"))))
        out[f"repro_cpu_raw_pq{pq:04d}"] = np.array([[int(r[f]) for f in FIELDS] for r in rows], np.int32)
    np.savez_compressed(os.path.join(HERE, "reference_query_traces.npz"), **out)


if __name__ == "__main__":
    main()
