"""bench_cpu, the reference bench_cpu's command line (bench_cpu.cpp:27-106, 389-542) over
song_cpu_search: its files have the reference's schemas, its ids are the library's CPU search on
the reference's sequential graph, and its summary recall is that of its ids files.  Host only."""
import csv
import json
import os
import subprocess

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BENCH_CPU = os.path.join(os.path.dirname(GOLDEN), "..", "2020_song_ex_amd", "bench_cpu")


@pytest.fixture(scope="module")
def cpu_run(tmp_path_factory, datasets):
    if not os.path.exists(BENCH_CPU):
        pytest.fail("bench_cpu is not built (make -C 2020_song_ex_amd/csrc)")
    d = tmp_path_factory.mktemp("bench_cpu")
    base, q = datasets.gmm_int(2500, 40, dim=24, n_centres=12, sigma=10.0, seed=6)
    datasets.write_fvecs(d / "b.fvecs", base)
    datasets.write_fvecs(d / "q.fvecs", q)
    # ground truth as run_compare writes it (int32 .npy), from an exact float64 brute force
    diff = q[:, None, :].astype(np.float64) - base[None, :, :].astype(np.float64)
    dist = (diff * diff).sum(-1)
    gt = np.array([np.lexsort((np.arange(base.shape[0]), row))[:10] for row in dist], np.int32)
    np.save(d / "gt.npy", gt)
    out = d / "out"
    proc = subprocess.run([BENCH_CPU, "--base", str(d / "b.fvecs"), "--query", str(d / "q.fvecs"),
                           "--n", "2500", "--nq", "40", "--k", "10", "--pq-list", "40,12", "--repeats", "2",
                           "--gt-npy", str(d / "gt.npy"), "--out-dir", str(out), "--ids-out", str(d / "cpu.ids"),
                           "--json-out", str(d / "cpu.json")], capture_output=True, text=True, timeout=300)
    assert proc.returncode == 0, proc.stderr
    return base, q, gt, d, out, proc.stdout


def test_files_have_the_reference_schemas(cpu_run):
    base, q, gt, d, out, stdout = cpu_run
    heads = json.load(open(os.path.join(GOLDEN, "reference_csv_headers.json")))
    with open(out / "sweep_summary.csv") as f:
        rows = list(csv.reader(f))
    assert rows[0] == heads["cpu_sweep_summary"]
    assert [r[0] for r in rows[1:]] == ["repro_cpu_raw", "repro_cpu_raw"]
    assert [int(r[1]) for r in rows[1:]] == [12, 40]  # sorted like the reference's pq list
    with open(out / "pq0012_query_stats.csv") as f:
        rd = csv.reader(f)
        assert next(rd) == heads["cpu_query_stats"]
        assert sum(1 for _ in rd) == q.shape[0]
    js = json.load(open(d / "cpu.json"))
    assert set(heads["cpu_json_keys"]) <= set(js)
    assert js["method"] == "repro_cpu_raw" and js["primary_pq"] == 40
    assert js["degree_stats"]["zero_degree_nodes"] == 0
    assert "[repro_cpu_raw] pq=40" in stdout


def test_ids_are_the_library_cpu_search(song, cpu_run):
    """Worst first, as Graph::search returns them and the reference's bench_cpu writes them,
    on the reference's sequential build (bench_cpu builds with build_graph, graph.h:190-199)."""
    base, q, gt, d, out, _ = cpu_run
    got = np.array([[int(t) for t in line.split()] for line in open(out / "pq0040.ids")])
    np.testing.assert_array_equal(got, np.array([[int(t) for t in line.split()] for line in open(d / "cpu.ids")]))
    adj = song.nsw_build(base, block=32)
    ids, _, tr, _, _ = song.cpu_search(base, adj, q, 10, 40, best_first=False)
    np.testing.assert_array_equal(got, ids)
    # the per-query trace columns are that search's counters
    with open(out / "pq0040_query_stats.csv") as f:
        rows = list(csv.DictReader(f))
    assert [int(r["expanded_nodes"]) for r in rows] == tr[:, 2].tolist()
    assert [int(r["distance_evals"]) for r in rows] == tr[:, 5].tolist()


def test_summary_recall_is_the_recall_of_the_ids(cpu_run):
    base, q, gt, d, out, _ = cpu_run
    rc = __import__("importlib").import_module("2020_song_ex_amd.run_compare")
    rows = {r["pq"]: r for r in rc.load_sweep_rows(out / "sweep_summary.csv", 10)}
    for pq in (12, 40):
        ids = rc.read_ids(out / f"pq{pq:04d}.ids", q.shape[0], 10)
        assert abs(rows[pq]["recall"] - rc.recall_at_k(ids, gt, 10)) < 1e-5
    assert rows[40]["recall"] >= rows[12]["recall"]
    assert rows[40]["qps"] > 0


def test_bad_ground_truth_is_an_error(cpu_run, tmp_path):
    base, q, gt, d, out, _ = cpu_run
    np.save(tmp_path / "gt.npy", gt.astype(np.float32))
    proc = subprocess.run([BENCH_CPU, "--base", str(d / "b.fvecs"), "--query", str(d / "q.fvecs"), "--n", "200",
                           "--nq", "5", "--gt-npy", str(tmp_path / "gt.npy"), "--out-dir", str(tmp_path / "o"),
                           "--json-out", ""], capture_output=True, text=True, timeout=60)
    assert proc.returncode == 1 and "int32 or int64" in proc.stderr
