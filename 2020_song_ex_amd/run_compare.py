"""run_compare-compatible driver over libsong (the reference's run_compare.py,
minus the official-SONG comparison and the plots).

Same command line as the reference (run_compare.py:1191-1217): it prepares the
base / query subsets (fvecs, cosine -> L2 normalisation), computes ground truth
(here by GPU brute force over the resident vectors instead of numpy,
run_compare.py:199-213), runs ``bench_cpu`` (method repro_cpu_raw, unless
--paper-only; run_compare.py:1316-1331) and ``bench_gpu`` (run_compare.py:1333-1350)
as subprocesses with the reference's flags, and writes metrics_long.csv /
metrics.csv / report.md from the benches' sweep_summary.csv files (the same column
schema the reference loads, run_compare.py:763-803).

    python -m 2020_song_ex_amd.run_compare --base B.fvecs --query Q.fvecs --n 100000 --nq 1000
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import shlex
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

from . import datasets
from .index import Index

HERE = Path(__file__).resolve().parent

DATASET_PRESETS = {  # run_compare.py:23-60 (hdf5 presets need h5py)
    "sift": {"base": "data/sift/sift_base.fvecs", "query": "data/sift/sift_query.fvecs", "dim": 128,
             "n": 1_000_000, "dist": "l2", "fmt": "fvecs"},
    "gist": {"base": "data/gist/gist-960-euclidean.hdf5", "query": "data/gist/gist-960-euclidean.hdf5",
             "dim": 960, "n": 1_000_000, "dist": "l2", "fmt": "hdf5"},
    "nytimes": {"base": "data/nytimes/nytimes-256-angular.hdf5", "query": "data/nytimes/nytimes-256-angular.hdf5",
                "dim": 256, "n": 290_000, "dist": "cos", "fmt": "hdf5"},
    "glove200": {"base": "data/glove200/glove-200-angular.hdf5", "query": "data/glove200/glove-200-angular.hdf5",
                 "dim": 200, "n": 1_183_514, "dist": "cos", "fmt": "hdf5"},
}

SWEEP_REQUIRED = ("method", "pq", "build_ms", "search_ms_mean", "search_ms_std", "qps_mean", "qps_std", "recall")


def run(cmd, cwd=None):
    """Subprocess with the reference's [cmd] logging (run_compare.py:108-146)."""
    print(f"[cmd] run: {' '.join(map(shlex.quote, cmd))}", flush=True)
    t0 = time.perf_counter()
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    dt = (time.perf_counter() - t0) * 1e3
    print(f"[cmd] done: rc={proc.returncode} time_ms={dt:.1f}", flush=True)
    if proc.returncode != 0:
        print(proc.stdout[-4000:], proc.stderr[-4000:], sep="\n", flush=True)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    return dt


def normalize_l2(arr):
    """run_compare.py:181-185."""
    norms = np.linalg.norm(arr, axis=1, keepdims=True)
    norms = np.where(norms == 0, 1.0, norms)
    return (arr / norms).astype(np.float32)


def write_npy_int32(path, arr):
    np.save(path, np.ascontiguousarray(arr, dtype=np.int32))


def recall_at_k(ids, gt, k):
    """run_compare.py:249-255."""
    total = 0.0
    for i in range(min(len(ids), len(gt))):
        gs = set(int(x) for x in gt[i, :k])
        total += sum(1 for x in ids[i, :k] if int(x) in gs) / k
    return total / len(gt)


def read_ids(path, nq, k):
    """run_compare.py:230-246."""
    out = np.full((nq, k), -1, dtype=np.int32)
    row = 0
    with open(path) as f:
        for line in f:
            if row >= nq:
                break
            vals = [int(t) for t in line.split()[:k]]
            if vals:
                out[row, :len(vals)] = vals
                row += 1
    return out


def load_sweep_rows(path, k):
    """Reads sweep_summary.csv with the columns the reference requires (run_compare.py:763-803)."""
    rows = []
    with open(path, newline="") as f:
        for raw in csv.DictReader(f):
            missing = [c for c in SWEEP_REQUIRED if c not in raw]
            if missing:
                raise RuntimeError(f"sweep summary lacks {missing}")
            row = {"method": raw["method"], "pq": int(raw["pq"]), "build_ms": float(raw["build_ms"]),
                   "search_ms": float(raw["search_ms_mean"]), "search_ms_std": float(raw["search_ms_std"]),
                   "qps": float(raw["qps_mean"]), "qps_std": float(raw["qps_std"]), "recall": float(raw["recall"]),
                   "k": k}
            for key in ("upload_ms", "query_projection_ms", "alloc_ms", "h2d_ms", "memset_ms", "kernel_ms",
                        "d2h_ms", "visited_mean", "expanded_mean", "distance_evals_mean", "early_stop_rate",
                        "latency_mean_ms", "latency_p50_ms", "latency_p99_ms"):
                row[key] = float(raw.get(key, 0) or 0)
            rows.append(row)
    if not rows:
        raise RuntimeError(f"empty sweep summary: {path}")
    return rows


def write_csv(path, rows, fields=None):
    fields = fields or list(rows[0].keys())
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields, extrasaction="ignore")
        w.writeheader()
        for r in rows:
            w.writerow(r)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--dataset", default="", choices=list(DATASET_PRESETS) + [""])
    ap.add_argument("--base", default="")
    ap.add_argument("--query", default="")
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--nq", type=int, default=1000)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--pq-size", type=int, default=50)
    ap.add_argument("--pq-list", default="10,20,30,40,50")
    ap.add_argument("--dim", type=int, default=0)
    ap.add_argument("--dist", default="", choices=["l2", "cos", "ip", ""])
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--batch-size", type=int, default=512)
    ap.add_argument("--out-dir", default="results/compare_run")
    ap.add_argument("--save-prepared-arrays", action="store_true")
    ap.add_argument("--paper-only", action="store_true")
    ap.add_argument("--from-metrics", default="")
    ap.add_argument("--threads", type=int, default=0, help="NSW build threads (addition)")
    ap.add_argument("--batch-div", type=int, default=0, help="batched NSW build (addition; 0 = reference)")
    ap.add_argument("--cpu-threads", type=int, default=1,
                    help="search threads of the CPU method (addition; 1 = the reference's serial loop)")
    args = ap.parse_args(argv)

    preset = DATASET_PRESETS.get(args.dataset, {})
    base_path = Path(args.base or preset.get("base", "data/sift/sift_base.fvecs"))
    query_path = Path(args.query or preset.get("query", "data/sift/sift_query.fvecs"))
    args.n = args.n or preset.get("n", 1_000_000)
    args.dist = args.dist or preset.get("dist", "l2")
    out_dir = Path(args.out_dir).resolve()
    prepared, runs = out_dir / "prepared", out_dir / "runs" / "repro_gpu"
    runs_cpu = out_dir / "runs" / "repro_cpu_raw"
    for d in (prepared, runs, runs_cpu):
        d.mkdir(parents=True, exist_ok=True)
    if args.from_metrics:
        rows = list(csv.DictReader(open(args.from_metrics)))
        write_report(out_dir / "report.md", rows, args)
        return 0
    if preset.get("fmt") == "hdf5":
        raise RuntimeError("hdf5 datasets need h5py, which is not installed; convert to fvecs first")
    if not base_path.exists() or not query_path.exists():
        raise RuntimeError(f"missing {base_path} or {query_path} (no download here)")

    print("[1/4] preparing subset and GT...", flush=True)
    base = datasets.read_fvecs(base_path, args.n)
    query = datasets.read_fvecs(query_path, args.nq)
    if args.dist == "cos":
        base, query = normalize_l2(base), normalize_l2(query)
    bench_base, bench_query = base_path, query_path
    if args.dist == "cos":
        bench_base = prepared / f"{args.dataset or 'data'}_base_{args.n}.fvecs"
        bench_query = prepared / f"{args.dataset or 'data'}_query_{args.nq}.fvecs"
        datasets.write_fvecs(bench_base, base)
        datasets.write_fvecs(bench_query, query)
    idx = Index(base, np.full((base.shape[0], 2), -1, np.int32))
    gt, _, _ = idx.bruteforce(query, args.k)
    idx.close()
    write_npy_int32(prepared / "gt.npy", gt)
    if args.save_prepared_arrays:
        np.save(prepared / "base.npy", base)
        np.save(prepared / "query.npy", query)

    threads = args.threads or min(16, os.cpu_count() or 1)
    steps = 4 if args.paper_only else 5
    step = 2
    if not args.paper_only:
        print(f"[{step}/{steps}] running repro CPU raw sweep...", flush=True)
        step += 1
        cpu_exe = HERE / "bench_cpu"
        if not cpu_exe.exists():
            raise RuntimeError(f"{cpu_exe} is not built (make -C 2020_song_ex_amd/csrc)")
        run([str(cpu_exe), "--base", str(bench_base), "--query", str(bench_query), "--n", str(base.shape[0]),
             "--nq", str(query.shape[0]), "--k", str(args.k), "--pq-list", args.pq_list,
             "--repeats", str(args.repeats), "--gt-npy", str(prepared / "gt.npy"), "--out-dir", str(runs_cpu),
             "--ids-out", str(out_dir / "repro_cpu_raw.ids"), "--json-out", str(out_dir / "repro_cpu_raw.json"),
             "--threads", str(args.cpu_threads), "--build-threads", str(threads),
             "--batch-div", str(args.batch_div)])
    print(f"[{step}/{steps}] running repro GPU raw/hash sweep...", flush=True)
    step += 1
    exe = HERE / "bench_gpu"
    if not exe.exists():
        raise RuntimeError(f"{exe} is not built (make -C 2020_song_ex_amd/csrc)")
    cmd = [str(exe), "--base", str(bench_base), "--query", str(bench_query), "--n", str(base.shape[0]),
           "--nq", str(query.shape[0]), "--k", str(args.k), "--gt-npy", str(prepared / "gt.npy"),
           "--pq-list", args.pq_list, "--repeats", str(args.repeats), "--batch-size", str(args.batch_size),
           "--out-dir", str(runs), "--raw-ids-out", str(out_dir / "repro_gpu_raw.ids"),
           "--hash-ids-out", str(out_dir / "repro_gpu_hash.ids"), "--json-out", str(out_dir / "repro_gpu.json"),
           "--threads", str(threads), "--batch-div", str(args.batch_div)]
    if args.paper_only:
        cmd += ["--no-hash", "1"]
    run(cmd)

    print(f"[{step}/{steps}] evaluating...", flush=True)
    step += 1
    rows = [] if args.paper_only else load_sweep_rows(runs_cpu / "sweep_summary.csv", args.k)
    rows += load_sweep_rows(runs / "sweep_summary.csv", args.k)
    for r in rows:  # recall recomputed from the ids files, as the reference does for official runs
        if r["method"] == "repro_cpu_raw":
            ids_path = runs_cpu / f"pq{r['pq']:04d}.ids"
        else:
            ids_path = runs / f"{r['method']}_pq{r['pq']:04d}.ids"
        if ids_path.exists():
            r["recall_from_ids"] = recall_at_k(read_ids(ids_path, query.shape[0], args.k), gt, args.k)
    write_csv(out_dir / "metrics_long.csv", rows)
    primary = max(r["pq"] for r in rows)
    prim = [r for r in rows if r["pq"] == primary]
    write_csv(out_dir / "metrics.csv", prim, ["method", "build_ms", "search_ms", "qps", "recall", "k"])
    print(f"[{step}/{steps}] report...", flush=True)
    write_report(out_dir / "report.md", prim, args)
    (out_dir / "manifest.json").write_text(json.dumps({"subset": {"n": args.n, "nq": args.nq, "k": args.k},
                                                       "metrics_long_csv": str(out_dir / "metrics_long.csv")},
                                                      indent=2))
    print(f"done:\n  metrics : {out_dir / 'metrics.csv'}\n  sweep   : {out_dir / 'metrics_long.csv'}", flush=True)
    return 0


def write_report(path, rows, args):
    lines = ["# SONG on MI355X: comparison report", "",
             f"n={args.n} nq={args.nq} k={args.k} dist={args.dist}", "",
             "| method | build_ms | search_ms | qps | recall |", "|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| {r['method']} | {float(r['build_ms']):.1f} | {float(r['search_ms']):.3f} | "
                     f"{float(r['qps']):.1f} | {float(r['recall']):.4f} |")
    Path(path).write_text("\n".join(lines) + "\n")


if __name__ == "__main__":
    sys.exit(main())
