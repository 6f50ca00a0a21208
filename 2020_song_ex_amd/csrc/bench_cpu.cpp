// bench_cpu.cpp -- command-line driver with the reference bench_cpu's flags and
// output files (bench_cpu.cpp:27-106, 389-542), running the CPU search
// baseline (Graph::search semantics, song_cpu_search) from libsong.so.
//
//   bench_cpu --base B.fvecs --query Q.fvecs [--n N] [--nq NQ] [--k K]
//             [--gt-npy gt.npy] [--pq-list 10,20,...|--pq-size P] [--repeats R]
//             [--out-dir DIR] [--ids-out F] [--json-out F]
//   additions: [--threads T]        search threads (1 = the reference's serial loop)
//              [--build-threads T]  NSW build and brute-force threads
//              [--batch-div D]      batched NSW build (0 = the reference's build_graph)
//
// Outputs, as the reference writes them: <out>/pqNNNN.ids (Graph::search's
// own order, worst first), <out>/pqNNNN_query_stats.csv (per-query latency,
// recall and trace), <out>/sweep_summary.csv (method repro_cpu_raw) and the
// JSON summary.  QPS = nq / wall time of the search loop.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "cli_common.hpp"
#include "song_graph.hpp"

namespace fs = std::filesystem;
using clk = std::chrono::steady_clock;
using namespace song::cli;

namespace {

double ms_between(clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

struct Options {
    std::string base = "data/sift/sift_base.fvecs", query = "data/sift/sift_query.fvecs", gt_npy;
    std::string out_dir = "results/repro_cpu_raw", ids_out, json_out = "results/repro_cpu_raw.json";
    std::string pq_list = std::to_string(song::pq_size);
    int n = 10000, nq = 1000, k = song::TOPK, repeats = 3;
    int threads = 1, build_threads = 1, batch_div = 0;
};

Options parse(int argc, char** argv) {
    std::map<std::string, std::string> kv;  // flags come in pairs, unknown ones are ignored
    for (int i = 1; i + 1 < argc; i += 2) kv[argv[i]] = argv[i + 1];
    Options o;
    auto s = [&](const char* key, std::string& dst) { if (kv.count(key)) dst = kv[key]; };
    auto i = [&](const char* key, int& dst) { if (kv.count(key)) dst = std::stoi(kv[key]); };
    s("--base", o.base); s("--query", o.query); s("--gt-npy", o.gt_npy); s("--out-dir", o.out_dir);
    s("--ids-out", o.ids_out); s("--json-out", o.json_out);
    s("--pq-list", o.pq_list); s("--pq-size", o.pq_list);  // --pq-size wins, as bench_cpu.cpp:105
    i("--n", o.n); i("--nq", o.nq); i("--k", o.k); i("--repeats", o.repeats);
    i("--threads", o.threads); i("--build-threads", o.build_threads); i("--batch-div", o.batch_div);
    o.repeats = std::max(1, o.repeats);
    o.threads = std::max(1, o.threads);
    o.build_threads = std::max(1, o.build_threads);
    return o;
}

// Exact k nearest by squared L2 summed in index order (distance.h:5-12), ties
// to the smaller id; best first.  Queries split over `threads`.
std::vector<std::vector<int>> brute_force(const std::vector<float>& base, int n, const std::vector<float>& q, int nq,
                                          int dim, int k, int threads) {
    std::vector<std::vector<int>> gt(nq);
    std::atomic<int> next{0};
    auto work = [&]() {
        std::vector<std::pair<float, int>> d((size_t)n);
        for (int i; (i = next.fetch_add(1)) < nq;) {
            const float* qi = q.data() + (size_t)i * dim;
            for (int j = 0; j < n; j++) {
                const float* x = base.data() + (size_t)j * dim;
                float s = 0.f;
                for (int t = 0; t < dim; t++) {
                    const float e = qi[t] - x[t];
                    const float p = e * e;
                    s += p;
                }
                d[j] = {s, j};
            }
            const int kk = std::min(k, n);
            std::partial_sort(d.begin(), d.begin() + kk, d.end());
            gt[i].resize(kk);
            for (int r = 0; r < kk; r++) gt[i][r] = d[r].second;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(work);
    for (auto& th : pool) th.join();
    return gt;
}

struct DegreeStats {
    int min_degree = 0, max_degree = 0, zero_degree_nodes = 0;
    double mean_degree = 0, std_degree = 0;
};

DegreeStats degree_stats(const std::vector<int>& adj, int n, int block) {
    std::vector<double> deg((size_t)n);
    DegreeStats s;
    for (int u = 0; u < n; u++) {
        int c = 0;
        for (int j = 0; j < block; j++) c += adj[(size_t)u * block + j] >= 0;
        deg[u] = c;
        s.zero_degree_nodes += c == 0;
    }
    const Stats st = summarize(deg);
    s.min_degree = (int)st.min;
    s.max_degree = (int)st.max;
    s.mean_degree = st.mean;
    s.std_degree = st.stddev;
    return s;
}

struct Row {
    int pq = 0;
    double build_ms = 0, gt_ms = 0, telemetry_ms = 0;
    Stats time, qps, latency, recall, visited, expanded, slots, candidates, devals, fpush, rupd, peakf, peakr;
    double early = 0;
    std::string ids_path, qstats_path;
};

void write_summary(const fs::path& path, const std::vector<Row>& rows) {
    std::ofstream f(path);
    f << "method,pq,build_ms,gt_ms,search_ms_telemetry,search_ms_mean,search_ms_std,search_ms_min,search_ms_max,"
         "qps_mean,qps_std,recall,latency_mean_ms,latency_std_ms,latency_min_ms,latency_p50_ms,latency_p90_ms,"
         "latency_p95_ms,latency_p99_ms,latency_max_ms,recall_mean,recall_p50,recall_p95,visited_mean,visited_p95,"
         "expanded_mean,expanded_p95,neighbor_slots_mean,distance_evals_mean,distance_evals_p95,"
         "frontier_pushes_mean,result_updates_mean,peak_frontier_mean,peak_results_mean,early_stop_rate,"
         "ids_path,query_stats_path\n";
    for (const Row& r : rows)
        f << "repro_cpu_raw," << r.pq << "," << r.build_ms << "," << r.gt_ms << "," << r.telemetry_ms << ","
          << r.time.mean << "," << r.time.stddev << "," << r.time.min << "," << r.time.max << "," << r.qps.mean << ","
          << r.qps.stddev << "," << r.recall.mean << "," << r.latency.mean << "," << r.latency.stddev << ","
          << r.latency.min << "," << r.latency.p50 << "," << r.latency.p90 << "," << r.latency.p95 << ","
          << r.latency.p99 << "," << r.latency.max << "," << r.recall.mean << "," << r.recall.p50 << ","
          << r.recall.p95 << "," << r.visited.mean << "," << r.visited.p95 << "," << r.expanded.mean << ","
          << r.expanded.p95 << "," << r.slots.mean << "," << r.devals.mean << "," << r.devals.p95 << ","
          << r.fpush.mean << "," << r.rupd.mean << "," << r.peakf.mean << "," << r.peakr.mean << "," << r.early << ","
          << r.ids_path << "," << r.qstats_path << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Options o = parse(argc, argv);
        const std::vector<int> pqs = int_list(o.pq_list, song::pq_size);
        const auto t0 = clk::now();
        int nb = 0, dim = 0, nqr = 0, qdim = 0;
        std::vector<float> base = read_fvecs(o.base, o.n, nb, dim);
        const auto t1 = clk::now();
        std::vector<float> query = read_fvecs(o.query, o.nq, nqr, qdim);
        const auto t2 = clk::now();
        if (nb == 0 || nqr == 0) {
            std::cerr << "empty base/query\n";
            return 1;
        }
        if (qdim != dim) throw song::Error("query dim differs from base dim");
        o.n = nb;
        o.nq = nqr;

        const auto tg0 = clk::now();
        std::vector<std::vector<int>> gt = o.gt_npy.empty()
                                               ? brute_force(base, o.n, query, o.nq, dim, o.k, o.build_threads)
                                               : load_gt(o.gt_npy, o.nq, o.k);
        const auto tg1 = clk::now();

        song::BuildConfig bc;
        bc.n_threads = o.build_threads;
        bc.batch_div = o.batch_div;
        bc.exact_order = true;
        const auto tb0 = clk::now();
        song::Graph g = song::build_graph(base, o.n, dim, bc);
        const auto tb1 = clk::now();
        const double build_ms = ms_between(tb0, tb1), gt_ms = ms_between(tg0, tg1);
        const DegreeStats ds = degree_stats(g.adj, o.n, g.block);

        fs::create_directories(o.out_dir);
        std::vector<Row> rows;
        std::vector<int> ids((size_t)o.nq * o.k);
        std::vector<song_query_trace> tr(o.nq);
        std::vector<double> lat(o.nq);
        for (int pq : pqs) {
            Row row;
            row.pq = pq;
            row.build_ms = build_ms;
            row.gt_ms = gt_ms;
            double wall = 0;
            song::check(song_cpu_search(g.data.data(), g.adj.data(), o.n, dim, g.block, query.data(), o.nq, o.k, pq,
                                        song::N_MULTIPROBE, o.threads, ids.data(), nullptr, tr.data(), lat.data(),
                                        &wall),
                        "song_cpu_search");
            row.telemetry_ms = wall;
            std::vector<double> times = {wall}, qps = {o.nq / (wall / 1000.0)};
            for (int r = 1; r < o.repeats; r++) {
                std::vector<int> scratch((size_t)o.nq * o.k);
                double w = 0;
                song::check(song_cpu_search(g.data.data(), g.adj.data(), o.n, dim, g.block, query.data(), o.nq, o.k,
                                            pq, song::N_MULTIPROBE, o.threads, scratch.data(), nullptr, nullptr,
                                            nullptr, &w),
                            "song_cpu_search");
                times.push_back(w);
                qps.push_back(o.nq / (w / 1000.0));
            }
            std::vector<double> rec(o.nq);
            for (int q = 0; q < o.nq; q++) rec[q] = recall_of(ids.data() + (size_t)q * o.k, gt[q], o.k);

            const fs::path dir(o.out_dir);
            row.ids_path = (dir / (pq_tag(pq) + ".ids")).string();
            row.qstats_path = (dir / (pq_tag(pq) + "_query_stats.csv")).string();
            write_ids(row.ids_path, ids, o.nq, o.k, true);
            if (!o.ids_out.empty() && pq == pqs.back()) write_ids(o.ids_out, ids, o.nq, o.k, true);
            std::vector<double> vis, exp, slo, cand, dev, fpu, rup, pkf, pkr, early;
            {
                std::ofstream f(row.qstats_path);
                f << "query_id,pq,latency_ms,recall,seed_count,visited_nodes,expanded_nodes,neighbor_slots_scanned,"
                     "candidate_nodes,distance_evals,frontier_pushes,result_updates,peak_frontier,peak_results,"
                     "result_count,early_stopped\n";
                for (int q = 0; q < o.nq; q++) {
                    const song_query_trace& t = tr[q];
                    f << q << "," << pq << "," << lat[q] << "," << rec[q] << "," << t.seed_count << ","
                      << t.visited_nodes << "," << t.expanded_nodes << "," << t.neighbor_slots_scanned << ","
                      << t.candidate_nodes << "," << t.distance_evals << "," << t.frontier_pushes << ","
                      << t.result_updates << "," << t.peak_frontier << "," << t.peak_results << "," << t.result_count
                      << "," << t.early_stopped << "\n";
                    vis.push_back(t.visited_nodes);
                    exp.push_back(t.expanded_nodes);
                    slo.push_back(t.neighbor_slots_scanned);
                    cand.push_back(t.candidate_nodes);
                    dev.push_back(t.distance_evals);
                    fpu.push_back(t.frontier_pushes);
                    rup.push_back(t.result_updates);
                    pkf.push_back(t.peak_frontier);
                    pkr.push_back(t.peak_results);
                    early.push_back(t.early_stopped);
                }
            }
            row.time = summarize(times);
            row.qps = summarize(qps);
            row.latency = summarize(lat);
            row.recall = summarize(rec);
            row.visited = summarize(vis);
            row.expanded = summarize(exp);
            row.slots = summarize(slo);
            row.candidates = summarize(cand);
            row.devals = summarize(dev);
            row.fpush = summarize(fpu);
            row.rupd = summarize(rup);
            row.peakf = summarize(pkf);
            row.peakr = summarize(pkr);
            row.early = summarize(early).mean;
            rows.push_back(row);
            std::cout << "[repro_cpu_raw] pq=" << pq << " build_ms=" << build_ms << " search_mean_ms=" << row.time.mean
                      << " qps_mean=" << row.qps.mean << " recall@" << o.k << "=" << row.recall.mean << "\n";
        }
        write_summary(fs::path(o.out_dir) / "sweep_summary.csv", rows);
        if (!o.json_out.empty()) {
            fs::path jp(o.json_out);
            if (jp.has_parent_path()) fs::create_directories(jp.parent_path());
            std::ofstream j(jp);
            const Row& p = rows.back();
            j << "{\n  \"method\": \"repro_cpu_raw\",\n  \"n\": " << o.n << ",\n  \"nq\": " << o.nq
              << ",\n  \"dim\": " << dim << ",\n  \"k\": " << o.k << ",\n  \"repeats\": " << o.repeats
              << ",\n  \"threads\": " << o.threads << ",\n  \"base_load_ms\": " << ms_between(t0, t1)
              << ",\n  \"query_load_ms\": " << ms_between(t1, t2) << ",\n  \"flatten_ms\": 0,\n  \"gt_ms\": " << gt_ms
              << ",\n  \"build_ms\": " << build_ms << ",\n  \"primary_pq\": " << p.pq << ",\n  \"search_ms\": "
              << p.time.mean << ",\n  \"qps\": " << p.qps.mean << ",\n  \"recall_at_k\": " << p.recall.mean
              << ",\n  \"degree_stats\": {\n    \"min_degree\": " << ds.min_degree << ",\n    \"max_degree\": "
              << ds.max_degree << ",\n    \"zero_degree_nodes\": " << ds.zero_degree_nodes
              << ",\n    \"mean_degree\": " << ds.mean_degree << ",\n    \"std_degree\": " << ds.std_degree
              << "\n  },\n  \"summary_csv\": \"" << (fs::path(o.out_dir) / "sweep_summary.csv").string()
              << "\",\n  \"pq_runs\": [\n";
            for (size_t i = 0; i < rows.size(); i++) {
                const Row& r = rows[i];
                j << "    {\n      \"pq\": " << r.pq << ",\n      \"search_ms_mean\": " << r.time.mean
                  << ",\n      \"search_ms_std\": " << r.time.stddev << ",\n      \"qps_mean\": " << r.qps.mean
                  << ",\n      \"recall_at_k\": " << r.recall.mean << ",\n      \"ids_path\": \"" << r.ids_path
                  << "\",\n      \"query_stats_path\": \"" << r.qstats_path << "\"\n    }"
                  << (i + 1 == rows.size() ? "\n" : ",\n");
            }
            j << "  ]\n}\n";
        }
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "bench_cpu: " << e.what() << "\n";
        return 1;
    }
}
