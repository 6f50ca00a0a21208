// cli_common.hpp -- file formats and statistics shared by the bench_gpu and
// bench_cpu drivers: .fvecs input, .npy ground truth, the quantile rule of
// the reference's summaries, and the ids / pq-tag file conventions.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "song_graph.hpp"

namespace song {
namespace cli {

// .fvecs: per row an int32 dimension followed by that many float32 values
// (the format of parser.h:10-22); the first `limit` rows are read.
inline std::vector<float> read_fvecs(const std::string& path, int limit, int& rows, int& dim) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw Error("cannot open " + path);
    std::vector<float> out;
    rows = 0;
    dim = 0;
    int32_t d;
    while (rows < limit && f.read((char*)&d, 4)) {
        if (dim == 0) dim = d;
        if (d != dim) throw Error("inconsistent dim in " + path);
        size_t off = out.size();
        out.resize(off + d);
        if (!f.read((char*)(out.data() + off), (std::streamsize)d * 4)) throw Error("short row in " + path);
        rows++;
    }
    return out;
}

// The value text of `key` in a .npy header dictionary ({'descr': '<i4',
// 'fortran_order': False, 'shape': (10000, 100), }): from after the colon to
// the comma that closes it at bracket depth 0.
inline std::string npy_dict_value(const std::string& header, const std::string& key) {
    const size_t at = header.find("'" + key + "'");
    if (at == std::string::npos) return "";
    size_t i = header.find(':', at);
    if (i == std::string::npos) return "";
    int depth = 0;
    std::string v;
    for (++i; i < header.size(); i++) {
        const char c = header[i];
        if (c == '(' || c == '[') depth++;
        if (c == ')' || c == ']') depth--;
        if ((c == ',' && depth == 0) || c == '}') break;
        v += c;
    }
    const size_t b = v.find_first_not_of(" \t"), e = v.find_last_not_of(" \t");
    return b == std::string::npos ? "" : v.substr(b, e - b + 1);
}

// A 2-D integer array from a .npy file (NumPy's format: the magic
// "\x93NUMPY", a major and a minor version byte, the header length as a
// little-endian uint16 in version 1 or uint32 in versions 2 and 3, the header
// dictionary, then the data).  Accepts little-endian int32 and int64, C order.
inline std::vector<int64_t> read_npy_int2d(const std::string& path, int64_t& rows, int64_t& cols) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw Error("cannot open ground truth " + path);
    unsigned char pre[8];
    if (!in.read((char*)pre, 8) || std::memcmp(pre, "\x93NUMPY", 6) != 0)
        throw Error(path + " is not a .npy file");
    uint32_t hlen = 0;
    if (pre[6] == 1) {
        unsigned char b[2];
        if (!in.read((char*)b, 2)) throw Error(path + ": truncated .npy header");
        hlen = b[0] | (b[1] << 8);
    } else if (pre[6] == 2 || pre[6] == 3) {
        unsigned char b[4];
        if (!in.read((char*)b, 4)) throw Error(path + ": truncated .npy header");
        hlen = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    } else {
        throw Error(path + ": unknown .npy version " + std::to_string(pre[6]));
    }
    std::string header(hlen, '\0');
    if (!in.read(header.data(), hlen)) throw Error(path + ": truncated .npy header");
    const std::string descr = npy_dict_value(header, "descr");
    const int width = descr == "'<i4'" || descr == "'|i4'" ? 4 : descr == "'<i8'" ? 8 : 0;
    if (!width) throw Error(path + ": ground truth must be int32 or int64, found " + descr);
    if (npy_dict_value(header, "fortran_order") != "False") throw Error(path + ": Fortran-ordered arrays are not read");
    std::string shape = npy_dict_value(header, "shape");
    std::vector<int64_t> dims;
    for (size_t i = 0; i < shape.size();) {
        if (shape[i] >= '0' && shape[i] <= '9') {
            size_t j = i;
            while (j < shape.size() && shape[j] >= '0' && shape[j] <= '9') j++;
            dims.push_back(std::stoll(shape.substr(i, j - i)));
            i = j;
        } else {
            i++;
        }
    }
    if (dims.size() != 2) throw Error(path + ": ground truth must be 2-D, shape " + shape);
    rows = dims[0];
    cols = dims[1];
    std::vector<int64_t> out((size_t)(rows * cols));
    if (width == 8) {
        if (!in.read((char*)out.data(), (std::streamsize)out.size() * 8)) throw Error(path + ": short payload");
    } else {
        std::vector<int32_t> raw(out.size());
        if (!in.read((char*)raw.data(), (std::streamsize)raw.size() * 4)) throw Error(path + ": short payload");
        std::copy(raw.begin(), raw.end(), out.begin());
    }
    return out;
}

// The first k columns of the first nq rows of a ground-truth .npy.
inline std::vector<std::vector<int>> load_gt(const std::string& path, int nq, int k) {
    int64_t rows = 0, cols = 0;
    const std::vector<int64_t> a = read_npy_int2d(path, rows, cols);
    if (rows < nq || cols < k)
        throw Error(path + ": ground truth is " + std::to_string(rows) + " x " + std::to_string(cols) +
                    ", smaller than the " + std::to_string(nq) + " x " + std::to_string(k) + " asked for");
    std::vector<std::vector<int>> gt(nq, std::vector<int>(k));
    for (int r = 0; r < nq; r++)
        for (int c = 0; c < k; c++) gt[r][c] = (int)a[(size_t)r * cols + c];
    return gt;
}

struct Stats {
    double min = 0, max = 0, mean = 0, stddev = 0, p50 = 0, p90 = 0, p95 = 0, p99 = 0;
};

// The reference's summary rule (bench_gpu.cu:127-156, bench_cpu.cpp:128-151):
// quantiles by linear interpolation at q * (N - 1) of the sorted values, the
// population standard deviation.
inline Stats summarize(std::vector<double> v) {
    Stats s;
    if (v.empty()) return s;
    std::sort(v.begin(), v.end());
    auto q = [&](double p) {
        if (v.size() == 1) return v[0];
        const double pos = p * (v.size() - 1);
        const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
        return v[lo] * (1.0 - (pos - lo)) + v[hi] * (pos - lo);
    };
    s.min = v.front();
    s.max = v.back();
    s.mean = std::accumulate(v.begin(), v.end(), 0.0) / v.size();
    double var = 0;
    for (double x : v) var += (x - s.mean) * (x - s.mean);
    s.stddev = std::sqrt(var / v.size());
    s.p50 = q(0.5);
    s.p90 = q(0.9);
    s.p95 = q(0.95);
    s.p99 = q(0.99);
    return s;
}

// "pq0050" -- the queue-size tag of the result file names.
inline std::string pq_tag(int pq) {
    std::ostringstream o;
    o << "pq" << std::setw(4) << std::setfill('0') << pq;
    return o.str();
}

// A comma-separated list of queue sizes, sorted and without repeats.
inline std::vector<int> int_list(const std::string& text, int fallback) {
    std::vector<int> v;
    std::stringstream ss(text);
    std::string tok;
    while (std::getline(ss, tok, ','))
        if (!tok.empty()) v.push_back(std::stoi(tok));
    if (v.empty()) v.push_back(fallback);
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    return v;
}

// Fraction of the k ground-truth ids among the first k result ids.
inline double recall_of(const int* ids, const std::vector<int>& gt, int k) {
    int hit = 0;
    for (int j = 0; j < k; j++)
        for (int r = 0; r < k && r < (int)gt.size(); r++)
            if (gt[r] == ids[j]) {
                hit++;
                break;
            }
    return (double)hit / k;
}

// One line per query, ids separated by single spaces; skip_pad leaves out the
// -1 padding of queries with fewer than k results (a CPU result vector is only
// as long as what was found).
inline void write_ids(const std::string& path, const std::vector<int>& ids, int nq, int k, bool skip_pad = false) {
    std::filesystem::path p(path);
    if (p.has_parent_path()) std::filesystem::create_directories(p.parent_path());
    std::ofstream f(path);
    for (int q = 0; q < nq; q++) {
        int w = 0;
        for (int j = 0; j < k; j++) {
            const int id = ids[(size_t)q * k + j];
            if (skip_pad && id < 0) continue;
            f << (w++ ? " " : "") << id;
        }
        f << '\n';
    }
}

}  // namespace cli
}  // namespace song
