"""The reference's command-line boundary end to end on the device: the
run_compare driver (run_compare.py:1191-1418) runs bench_cpu and bench_gpu as
subprocesses with the reference's flags, and all of them write the reference's
files."""
import csv
import importlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BENCH_GPU = os.path.join(os.path.dirname(GOLDEN), "..", "2020_song_ex_amd", "bench_gpu")


@pytest.fixture(scope="module")
def compare_run(tmp_path_factory, datasets):
    if not os.path.exists(BENCH_GPU):
        pytest.fail("bench_gpu is not built (make -C 2020_song_ex_amd/csrc)")
    d = tmp_path_factory.mktemp("compare")
    base, q = datasets.gmm_int(3000, 60, dim=32, n_centres=20, sigma=10.0, seed=2)
    datasets.write_fvecs(d / "b.fvecs", base)
    datasets.write_fvecs(d / "q.fvecs", q)
    rc = importlib.import_module("2020_song_ex_amd.run_compare")
    out = d / "out"
    assert rc.main(["--base", str(d / "b.fvecs"), "--query", str(d / "q.fvecs"), "--n", "3000", "--nq", "60",
                    "--k", "10", "--pq-list", "16,64", "--repeats", "2", "--batch-size", "32",
                    "--out-dir", str(out), "--threads", "4"]) == 0
    return base, q, out


def test_driver_outputs(compare_run):
    base, q, out = compare_run
    rows = list(csv.DictReader(open(out / "metrics_long.csv")))
    assert {r["method"] for r in rows} == {"repro_cpu_raw", "repro_gpu_raw", "repro_gpu_hash"}
    assert {int(r["pq"]) for r in rows} == {16, 64}
    for r in rows:
        # the summary's recall is the recall of the ids file it names
        assert abs(float(r["recall"]) - float(r["recall_from_ids"])) < 1e-5, r  # printed to 6 digits
        assert float(r["qps"]) > 0
    raw = {int(r["pq"]): float(r["recall"]) for r in rows if r["method"] == "repro_gpu_raw"}
    assert raw[64] > raw[16] > 0.1  # recall rises with the queue
    cpu = {int(r["pq"]): float(r["recall"]) for r in rows if r["method"] == "repro_cpu_raw"}
    assert cpu[64] >= cpu[16] > 0.1
    assert (out / "repro_cpu_raw.ids").exists() and (out / "repro_cpu_raw.json").exists()
    assert (out / "metrics.csv").exists() and (out / "report.md").exists()


def test_bench_files_have_the_reference_schemas(compare_run):
    base, q, out = compare_run
    heads = json.load(open(os.path.join(GOLDEN, "reference_csv_headers.json")))
    runs = out / "runs" / "repro_gpu"
    with open(runs / "sweep_summary.csv") as f:
        assert next(csv.reader(f)) == heads["sweep_summary"]
    for method in ("repro_gpu_raw", "repro_gpu_hash"):
        with open(runs / f"{method}_pq0016_query_stats.csv") as f:
            rd = csv.reader(f)
            assert next(rd) == heads["query_stats"]
            assert sum(1 for _ in rd) == q.shape[0]
        with open(runs / f"{method}_pq0016_batch_stats.csv") as f:
            assert next(csv.reader(f)) == heads["batch_stats"]
        ids = [line.split() for line in open(runs / f"{method}_pq0064.ids")]
        assert len(ids) == q.shape[0] and all(len(r) == 10 for r in ids)


def test_raw_ids_equal_the_library_search(song, compare_run):
    """bench_gpu's raw method is the library's search on the reference-built graph."""
    base, q, out = compare_run
    runs = out / "runs" / "repro_gpu"
    got = np.array([[int(t) for t in line.split()] for line in open(runs / "repro_gpu_raw_pq0064.ids")])
    adj = song.nsw_build(base, block=32, n_threads=4)  # bench_gpu's default: the reference's sequential build
    idx = song.Index(base, adj)
    try:
        r = idx.search(q, 10, 64, batch=32)
    finally:
        idx.close()
    np.testing.assert_array_equal(got, r.ids)


def test_cpu_ids_equal_the_library_cpu_search(song, compare_run):
    """The driver's CPU method is song_cpu_search (Graph::search) on the same reference-built
    graph, its ids worst first as the reference's bench_cpu writes them."""
    base, q, out = compare_run
    got = np.array([[int(t) for t in line.split()] for line in open(out / "repro_cpu_raw.ids")])
    adj = song.nsw_build(base, block=32, n_threads=4)
    ids, _, _, _, _ = song.cpu_search(base, adj, q, 10, 64, best_first=False)
    np.testing.assert_array_equal(got, ids)


def test_hash_ids_equal_the_library_search(song, compare_run):
    """bench_gpu's hash method (queries hashed on the GPU) is the library's Hamming search over
    the CPU BitHash codes (seed 42, bitHash.h:22) on the reference-built code graph."""
    base, q, out = compare_run
    runs = out / "runs" / "repro_gpu"
    got = np.array([[int(t) for t in line.split()] for line in open(runs / "repro_gpu_hash_pq0064.ids")])
    R = song.bithash_matrix(base.shape[1], 42)
    codes = song.bithash_project(R, base)
    qcodes = song.bithash_project(R, q)
    adj = song.nsw_build(codes, metric="hamming", block=32, n_threads=4)
    idx = song.Index(codes, adj, metric="hamming")
    try:
        r = idx.search(qcodes, 10, 64, batch=32)
    finally:
        idx.close()
    np.testing.assert_array_equal(got, r.ids)
