// bench_gpu.cpp -- command-line driver with the reference bench_gpu's flags
// and output files (bench_gpu.cu:27-580), running on libsong.so.
//
//   bench_gpu --base B.fvecs --query Q.fvecs [--n N] [--nq NQ] [--k K]
//             [--gt-npy gt.npy] [--pq-list 10,20,...|--pq-size P] [--repeats R]
//             [--batch-size S] [--out-dir DIR] [--raw-ids-out F] [--hash-ids-out F]
//             [--json-out F]
//   additions: [--threads T] [--batch-div D] [--bloom-bits B] [--no-hash 1]
//
// Outputs, as the reference writes them: <out>/<method>_pqNNNN.ids (best
// first), <method>_pqNNNN_query_stats.csv, <method>_pqNNNN_batch_stats.csv,
// <out>/sweep_summary.csv and the JSON summary; methods repro_gpu_raw and
// repro_gpu_hash (the names run_compare.py keys on).  QPS = nq / total_ms of
// the search call (query H2D + kernel + ids/trace D2H, allocation excluded).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <numeric>
#include <sstream>
#include <string>
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
    std::string out_dir = "results/repro_gpu", raw_ids_out, hash_ids_out, json_out = "results/repro_gpu.json";
    std::string pq_list = std::to_string(song::pq_size);
    int n = 10000, nq = 1000, k = song::TOPK, repeats = 3, batch_size = song::N_MULTIQUERY;
    int threads = 1, batch_div = 0, bloom_bits = song::BF_WORDS * 64, no_hash = 0;
};

Options parse(int argc, char** argv) {
    std::map<std::string, std::string> kv;  // flags come in pairs, unknown ones are ignored
    for (int i = 1; i + 1 < argc; i += 2) kv[argv[i]] = argv[i + 1];
    Options o;
    auto s = [&](const char* key, std::string& dst) { if (kv.count(key)) dst = kv[key]; };
    auto i = [&](const char* key, int& dst) { if (kv.count(key)) dst = std::stoi(kv[key]); };
    s("--base", o.base); s("--query", o.query); s("--gt-npy", o.gt_npy); s("--out-dir", o.out_dir);
    s("--raw-ids-out", o.raw_ids_out); s("--hash-ids-out", o.hash_ids_out); s("--json-out", o.json_out);
    s("--pq-list", o.pq_list); s("--pq-size", o.pq_list);  // --pq-size wins, as bench_gpu.cu:107
    i("--n", o.n); i("--nq", o.nq); i("--k", o.k); i("--repeats", o.repeats); i("--batch-size", o.batch_size);
    i("--threads", o.threads); i("--batch-div", o.batch_div); i("--bloom-bits", o.bloom_bits); i("--no-hash", o.no_hash);
    o.repeats = std::max(1, o.repeats);
    o.batch_size = std::max(1, o.batch_size);
    return o;
}

struct Row {
    std::string method;
    int pq = 0;
    double build_ms = 0, upload_ms = 0, projection_ms = 0, telemetry_ms = 0;
    Stats time, qps, latency, recall, visited, expanded, slots, devals, fpush, rupd, peakf, peakr;
    double recall_mean = 0, alloc_ms = 0, h2d_ms = 0, memset_ms = 0, kernel_ms = 0, d2h_ms = 0, early = 0;
    std::string ids_path, qstats_path, bstats_path;
};

template <class Fn>
Row run_variant(const std::string& method, const Options& o, int pq, Fn&& search,
                const std::vector<std::vector<int>>& gt) {
    song::GpuSearchTrace first;
    std::vector<int> ids = search(pq, &first);
    std::vector<double> times = {first.total_ms}, qps = {o.nq / (first.total_ms / 1000.0)};
    for (int r = 1; r < o.repeats; r++) {
        song::GpuSearchTrace t;
        (void)search(pq, &t);
        times.push_back(t.total_ms);
        qps.push_back(o.nq / (t.total_ms / 1000.0));
    }
    std::vector<double> rec(o.nq, 0.0);
    if (!gt.empty())
        for (int q = 0; q < o.nq; q++) rec[q] = recall_of(ids.data() + (size_t)q * o.k, gt[q], o.k);
    fs::path dir(o.out_dir);
    Row row;
    row.method = method;
    row.pq = pq;
    row.ids_path = (dir / (method + "_" + pq_tag(pq) + ".ids")).string();
    row.qstats_path = (dir / (method + "_" + pq_tag(pq) + "_query_stats.csv")).string();
    row.bstats_path = (dir / (method + "_" + pq_tag(pq) + "_batch_stats.csv")).string();
    write_ids(row.ids_path, ids, o.nq, o.k);
    {
        std::ofstream f(row.bstats_path);
        f << "pq,batch_index,batch_offset,batch_size,h2d_ms,memset_ms,kernel_ms,d2h_ms,total_ms,per_query_total_ms\n";
        for (auto& b : first.batches) {
            double tot = b.h2d_ms + b.memset_ms + b.kernel_ms + b.d2h_ms;
            f << pq << "," << b.batch_index << "," << b.batch_offset << "," << b.batch_size << "," << b.h2d_ms << ","
              << b.memset_ms << "," << b.kernel_ms << "," << b.d2h_ms << "," << tot << "," << tot / b.batch_size << "\n";
        }
    }
    std::vector<double> lat, vis, exp, slo, dev, fpu, rup, pkf, pkr, early;
    {
        std::ofstream f(row.qstats_path);
        f << "query_id,pq,batch_index,approx_h2d_ms,approx_memset_ms,approx_kernel_ms,approx_d2h_ms,approx_total_ms,"
             "recall,seed_count,visited_nodes,expanded_nodes,neighbor_slots_scanned,candidate_nodes,distance_evals,"
             "frontier_pushes,result_updates,peak_frontier,peak_results,result_count,early_stopped\n";
        for (int q = 0; q < o.nq; q++) {
            const auto& t = first.queries[q];
            const auto& b = first.batches[q / o.batch_size];
            double h = b.h2d_ms / b.batch_size, m = b.memset_ms / b.batch_size, kk = b.kernel_ms / b.batch_size,
                   d = b.d2h_ms / b.batch_size;
            f << q << "," << pq << "," << b.batch_index << "," << h << "," << m << "," << kk << "," << d << ","
              << (h + m + kk + d) << "," << rec[q] << "," << t.seed_count << "," << t.visited_nodes << ","
              << t.expanded_nodes << "," << t.neighbor_slots_scanned << "," << t.candidate_nodes << ","
              << t.distance_evals << "," << t.frontier_pushes << "," << t.result_updates << "," << t.peak_frontier
              << "," << t.peak_results << "," << t.result_count << "," << t.early_stopped << "\n";
            lat.push_back(h + m + kk + d);
            vis.push_back(t.visited_nodes);
            exp.push_back(t.expanded_nodes);
            slo.push_back(t.neighbor_slots_scanned);
            dev.push_back(t.distance_evals);
            fpu.push_back(t.frontier_pushes);
            rup.push_back(t.result_updates);
            pkf.push_back(t.peak_frontier);
            pkr.push_back(t.peak_results);
            early.push_back(t.early_stopped);
        }
    }
    row.telemetry_ms = first.total_ms;
    row.time = summarize(times);
    row.qps = summarize(qps);
    row.recall = summarize(rec);
    row.recall_mean = gt.empty() ? 0.0 : row.recall.mean;
    row.alloc_ms = first.alloc_ms;
    row.h2d_ms = first.h2d_ms;
    row.memset_ms = first.memset_ms;
    row.kernel_ms = first.kernel_ms;
    row.d2h_ms = first.d2h_ms;
    row.latency = summarize(lat);
    row.visited = summarize(vis);
    row.expanded = summarize(exp);
    row.slots = summarize(slo);
    row.devals = summarize(dev);
    row.fpush = summarize(fpu);
    row.rupd = summarize(rup);
    row.peakf = summarize(pkf);
    row.peakr = summarize(pkr);
    row.early = summarize(early).mean;
    return row;
}

void write_summary(const fs::path& path, const std::vector<Row>& rows) {
    std::ofstream f(path);
    f << "method,pq,build_ms,upload_ms,query_projection_ms,search_ms_telemetry,search_ms_mean,search_ms_std,"
         "search_ms_min,search_ms_max,qps_mean,qps_std,recall,alloc_ms,h2d_ms,memset_ms,kernel_ms,d2h_ms,"
         "approx_latency_mean_ms,approx_latency_p95_ms,recall_mean,recall_p95,visited_mean,visited_p95,"
         "expanded_mean,expanded_p95,neighbor_slots_mean,distance_evals_mean,distance_evals_p95,"
         "frontier_pushes_mean,result_updates_mean,peak_frontier_mean,peak_results_mean,early_stop_rate,"
         "ids_path,query_stats_path,batch_stats_path\n";
    for (auto& r : rows)
        f << r.method << "," << r.pq << "," << r.build_ms << "," << r.upload_ms << "," << r.projection_ms << ","
          << r.telemetry_ms << "," << r.time.mean << "," << r.time.stddev << "," << r.time.min << "," << r.time.max
          << "," << r.qps.mean << "," << r.qps.stddev << "," << r.recall_mean << "," << r.alloc_ms << "," << r.h2d_ms
          << "," << r.memset_ms << "," << r.kernel_ms << "," << r.d2h_ms << "," << r.latency.mean << ","
          << r.latency.p95 << "," << r.recall.mean << "," << r.recall.p95 << "," << r.visited.mean << ","
          << r.visited.p95 << "," << r.expanded.mean << "," << r.expanded.p95 << "," << r.slots.mean << ","
          << r.devals.mean << "," << r.devals.p95 << "," << r.fpush.mean << "," << r.rupd.mean << "," << r.peakf.mean
          << "," << r.peakr.mean << "," << r.early << "," << r.ids_path << "," << r.qstats_path << "," << r.bstats_path
          << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    try {
        Options o = parse(argc, argv);
        const std::vector<int> pqs = int_list(o.pq_list, song::pq_size);
        auto t0 = clk::now();
        int nb = 0, dim = 0, nqr = 0, qdim = 0;
        std::vector<float> base = read_fvecs(o.base, o.n, nb, dim);
        auto t1 = clk::now();
        std::vector<float> query = read_fvecs(o.query, o.nq, nqr, qdim);
        auto t2 = clk::now();
        if (nb == 0 || nqr == 0) {
            std::cerr << "empty base/query\n";
            return 1;
        }
        if (qdim != dim) throw song::Error("query dim differs from base dim");
        o.n = nb;
        o.nq = nqr;
        std::vector<std::vector<int>> gt;
        if (!o.gt_npy.empty()) gt = load_gt(o.gt_npy, o.nq, o.k);

        song::BuildConfig bc;
        bc.n_threads = o.threads;
        bc.batch_div = o.batch_div;
        bc.exact_order = o.batch_div == 0;
        song::SearchConfig sc;
        sc.bloom_bits = o.bloom_bits;

        auto tr0 = clk::now();
        song::KernelGraph kg = song::build_kernel_graph(base, o.n, dim, bc);
        auto tr1 = clk::now();
        kg.upload_to_gpu();
        auto tr2 = clk::now();
        double hash_build = 0, hash_proj = 0, hash_up = 0;
        std::vector<song::HashKernelGraph> hkg;
        std::vector<uint32_t> qhash;
        if (!o.no_hash) {
            auto th0 = clk::now();
            hkg.push_back(song::build_hashed_graph(base.data(), o.n, dim, bc));
            auto th1 = clk::now();
            // queries hashed on the GPU (the reference's BitHash::project_all codes, bit for bit)
            qhash = hkg[0].bh.project_all_gpu(query.data(), o.nq);
            auto th2 = clk::now();
            hkg[0].upload_to_gpu();
            auto th3 = clk::now();
            hash_build = ms_between(th0, th1);
            hash_proj = ms_between(th1, th2);
            hash_up = ms_between(th2, th3);
        }
        fs::create_directories(o.out_dir);
        std::vector<Row> rows;
        for (int pq : pqs) {
            Row raw = run_variant("repro_gpu_raw", o, pq, [&](int ef, song::GpuSearchTrace* t) {
                return song::gpu_search(kg, query.data(), o.nq, o.k, o.batch_size, ef, t, sc);
            }, gt);
            raw.build_ms = ms_between(tr0, tr1);
            raw.upload_ms = ms_between(tr1, tr2);
            rows.push_back(raw);
            if (!o.raw_ids_out.empty() && pq == pqs.back())
                fs::copy_file(raw.ids_path, o.raw_ids_out, fs::copy_options::overwrite_existing);
            std::cout << "[repro_gpu_raw] pq=" << pq << " build_ms=" << raw.build_ms << " upload_ms=" << raw.upload_ms
                      << " search_mean_ms=" << raw.time.mean << " qps_mean=" << raw.qps.mean << " recall@" << o.k
                      << "=" << raw.recall_mean << "\n";
            if (!o.no_hash) {
                Row h = run_variant("repro_gpu_hash", o, pq, [&](int ef, song::GpuSearchTrace* t) {
                    return song::gpu_search(hkg[0], qhash.data(), o.nq, o.k, o.batch_size, ef, t, sc);
                }, gt);
                h.build_ms = hash_build;
                h.upload_ms = hash_up;
                h.projection_ms = hash_proj;
                rows.push_back(h);
                if (!o.hash_ids_out.empty() && pq == pqs.back())
                    fs::copy_file(h.ids_path, o.hash_ids_out, fs::copy_options::overwrite_existing);
                std::cout << "[repro_gpu_hash] pq=" << pq << " build_ms=" << h.build_ms << " upload_ms=" << h.upload_ms
                          << " search_mean_ms=" << h.time.mean << " qps_mean=" << h.qps.mean << " recall@" << o.k
                          << "=" << h.recall_mean << "\n";
            }
        }
        write_summary(fs::path(o.out_dir) / "sweep_summary.csv", rows);
        if (!o.json_out.empty()) {
            fs::path jp(o.json_out);
            if (jp.has_parent_path()) fs::create_directories(jp.parent_path());
            std::ofstream j(jp);
            const Row* rp = nullptr;
            const Row* hp = nullptr;
            for (auto& r : rows) {
                if (r.method == "repro_gpu_raw") rp = &r;
                if (r.method == "repro_gpu_hash") hp = &r;
            }
            j << "{\n  \"n\": " << o.n << ",\n  \"nq\": " << o.nq << ",\n  \"dim\": " << dim << ",\n  \"k\": " << o.k
              << ",\n  \"repeats\": " << o.repeats << ",\n  \"batch_size\": " << o.batch_size
              << ",\n  \"base_load_ms\": " << ms_between(t0, t1) << ",\n  \"query_load_ms\": " << ms_between(t1, t2)
              << ",\n  \"flatten_ms\": 0,\n  \"raw_build_ms\": " << ms_between(tr0, tr1)
              << ",\n  \"raw_upload_ms\": " << ms_between(tr1, tr2) << ",\n  \"hash_build_ms\": " << hash_build
              << ",\n  \"hash_projection_ms\": " << hash_proj << ",\n  \"hash_upload_ms\": " << hash_up << ",\n";
            if (rp) j << "  \"raw_search_ms\": " << rp->time.mean << ",\n  \"raw_qps\": " << rp->qps.mean << ",\n";
            if (hp) j << "  \"hash_search_ms\": " << hp->time.mean << ",\n  \"hash_qps\": " << hp->qps.mean << ",\n";
            j << "  \"summary_csv\": \"" << (fs::path(o.out_dir) / "sweep_summary.csv").string() << "\"\n}\n";
        }
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "bench_gpu: " << e.what() << "\n";
        return 1;
    }
}
