#pragma once
// song_search_impl.h -- the SONG graph-search kernels for MI355X (gfx950).
// Included by song_search_part.hip, compiled once per metric and set size
// (slots per lane), which instantiates the launch templates of that pair, and
// by song_search.hip (dispatch and the visit-table kernel).
//
// One 64-lane wavefront searches one query.  The algorithm is the reference's
// kg_search_kernel / hkg_search_kernel (kernalGraph.h:184-299,
// hashKernelGraph.h:238-352) with the same results bit for bit; only the
// execution is re-planned for a wave64 machine:
//
//  * Candidates / Results (gpuHeap.h:8-109) keep the reference's
//    unsorted-array semantics -- append while not full, otherwise replace the
//    first-index maximum if strictly better; pop = first-index minimum, hole
//    filled by the last element.  Entry i lives in VGPR slot i/64 of lane i%64,
//    so every scan is a wave reduction (DPP max/min + 4 readlanes) followed by
//    one ballot per slot to find the first index holding that value.
//  * The Bloom visited set (bloomFilter.h:14-49, same hash, same word/bit
//    layout) lives in LDS.  All slots of an adjacency row are hashed, tested
//    and set lane-parallel; the reference's in-row order (a bit set by slot j
//    is seen by slot j' > j) is honoured exactly: a 64-bit ds_or_rtn reveals
//    any position two fresh slots share, and only then the row is replayed
//    slot by slot.
//  * Neighbour rows are gathered with 16-byte loads, G lanes per row, and the
//    partial sums reduced by an xor butterfly built from DPP / swizzle.  The
//    summation order (per-lane chunk order, explicit fmaf, butterfly) is fixed
//    and is restated by the oracle.
//  * The reference's ordered per-neighbour pushes become one batched step:
//    appends are a lane rotation, and the remaining pushes on a full set are
//    resolved in order by ballots -- only pushes that are accepted cost a
//    reduction, rejected ones are skipped in bulk.
//
// Nothing here allocates; every buffer is passed in.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "song_kernels.h"

namespace song {

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int dppi(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7-i within 8
constexpr int DPP_MIRROR = 0x140;      // lane i <-> 15-i within 16

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, dppf<DPP_XOR1>(v));
    v = fmaxf(v, dppf<DPP_XOR2>(v));
    v = fmaxf(v, dppf<DPP_HALF_MIRROR>(v));
    v = fmaxf(v, dppf<DPP_MIRROR>(v));
    float a = fmaxf(readlane_f(v, 0), readlane_f(v, 16));
    float b = fmaxf(readlane_f(v, 32), readlane_f(v, 48));
    return fmaxf(a, b);
}

__device__ __forceinline__ float wave_min(float v) {
    v = fminf(v, dppf<DPP_XOR1>(v));
    v = fminf(v, dppf<DPP_XOR2>(v));
    v = fminf(v, dppf<DPP_HALF_MIRROR>(v));
    v = fminf(v, dppf<DPP_MIRROR>(v));
    float a = fminf(readlane_f(v, 0), readlane_f(v, 16));
    float b = fminf(readlane_f(v, 32), readlane_f(v, 48));
    return fminf(a, b);
}

__device__ __forceinline__ float wave_sum(float v) {
    v = v + dppf<DPP_XOR1>(v);
    v = v + dppf<DPP_XOR2>(v);
    v = v + dppf<DPP_HALF_MIRROR>(v);
    v = v + dppf<DPP_MIRROR>(v);
    const float a = readlane_f(v, 0) + readlane_f(v, 16);
    const float b = readlane_f(v, 32) + readlane_f(v, 48);
    return a + b;
}

// Sum over aligned groups of G lanes, xor-butterfly order: for s = 1, 2, ...
// < G every lane t takes acc[t] + acc[t ^ s].  Steps 4 and 8 use the mirror
// patterns, whose partner lies in the same already-uniform 4 / 8 block as the
// xor partner, so the value is identical.  Every lane of the group ends with
// the group's sum.
template <int G>
__device__ __forceinline__ float group_sum(float v) {
    if constexpr (G > 1) v = v + dppf<DPP_XOR1>(v);
    if constexpr (G > 2) v = v + dppf<DPP_XOR2>(v);
    if constexpr (G > 4) v = v + dppf<DPP_HALF_MIRROR>(v);
    if constexpr (G > 8) v = v + dppf<DPP_MIRROR>(v);
    if constexpr (G > 16)
        v = v + __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (16 << 10)));
    if constexpr (G > 32) v = v + __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int ffs64(uint64_t m) { return __ffsll((unsigned long long)m) - 1; }

// ---------------------------------------------------------------- Bloom filter

// BF_RAND, bloomFilter.h:14-19.  Returns the low 32 bits of hash h of id;
// the bit count is a power of two, so the position is these bits masked.
__device__ __forceinline__ uint32_t bloom_hash(int h, uint32_t id) {
    constexpr uint64_t R[8] = {
        0x4bcb391f924ed183ULL, 0xa0ab69ccd854fc0aULL, 0x91086b9cecf5e3b7ULL, 0xc68e01641bead407ULL,
        0x3a7b976128a30449ULL, 0x6d122efabfc4d99fULL, 0xe6700ef8715030e2ULL, 0x80dd0c3bffcfb45bULL,
    };
    // bf_hash, bloomFilter.h:23-30.  id < 2^31, so the first x ^= x >> 33 is
    // the identity.
    uint64_t x = (uint64_t)id;
    x *= R[2 * h];
    x ^= x >> 33;
    x *= R[2 * h + 1];
    x ^= x >> 33;
    return (uint32_t)x;
}
// "% nbits" of bf_hash (bloomFilter.h:30) for a power-of-two bit count.
__device__ __forceinline__ uint32_t bloom_pos(int h, uint32_t id, uint32_t nbits_mask) {
    return bloom_hash(h, id) & nbits_mask;
}

// Test-then-set of one row of ids in slot order (the reference's per-slot
// bf_test / bf_set, kernalGraph.h:262-272).  Returns the mask of lanes whose
// id was not (possibly) visited; their bits are now set.
__device__ __forceinline__ uint64_t bloom_row(unsigned long long* bloom, uint32_t wmask, int wshift,
                                              uint32_t nbits_mask, int id, bool valid) {
    uint32_t word[4], bitn[4];
    unsigned long long pre[4];
    bool miss[4];
    bool all = valid;
#pragma unroll
    for (int h = 0; h < 4; h++) {
        uint32_t pos = valid ? bloom_pos(h, (uint32_t)id, nbits_mask) : 0u;
        word[h] = pos & wmask;  // word = pos % BF_WORDS
        bitn[h] = pos >> wshift; // bit  = pos / BF_WORDS
        pre[h] = valid ? bloom[word[h]] : 0ull;
        miss[h] = !((pre[h] >> bitn[h]) & 1ull);
        all = all && !miss[h];
    }
    const bool fresh = valid && !all;
    bool dup = false;
    if (fresh) {
#pragma unroll
        for (int h = 0; h < 4; h++) {
            if (miss[h]) {
                unsigned long long bit = 1ull << bitn[h];
                unsigned long long old = atomicOr(&bloom[word[h]], bit);
                dup = dup || ((old & bit) != 0ull);
            }
        }
    }
    const uint64_t fresh_mask = __ballot(fresh);
    if (__ballot(dup) == 0ull) return fresh_mask;

    // Two fresh slots share a position (or one slot hashes twice to it): put
    // the words back as they were before this row and replay in slot order.
    if (fresh) {
#pragma unroll
        for (int h = 0; h < 4; h++)
            if (miss[h]) bloom[word[h]] = pre[h];
    }
    wave_lds_sync();
    uint64_t acc = 0ull;
    uint64_t fm = fresh_mask;
    const int lane = lane_id();
    while (fm) {
        const int j = ffs64(fm);
        fm &= fm - 1ull;
        bool take = false;
        if (lane == j) {
            bool hit = true;
#pragma unroll
            for (int h = 0; h < 4; h++) hit = hit && ((bloom[word[h]] >> bitn[h]) & 1ull);
            if (!hit) {
#pragma unroll
                for (int h = 0; h < 4; h++) bloom[word[h]] |= (1ull << bitn[h]);
            }
            take = !hit;
        }
        wave_lds_sync();
        if (__ballot(take)) acc |= 1ull << j;
    }
    return acc;
}

// The same test-then-set for rows of at most 32 ids, with the hashing split
// over both half-waves: lane j computes hashes 0 and 1 of slot j, lane j + 32
// hashes 2 and 3, so each lane runs two of the four 64-bit hash chains.  The
// caller gives every lane its slot's id (lane j and lane j + 32 both hold slot
// j: the row is loaded twice from the same cache lines, so no lane has to
// pass it to the other half).  A slot is "all hit" when both its lanes see
// their two bits set.  Duplicates (two fresh slots, or two hashes of one slot,
// on one bit) are replayed slot by slot exactly as in bloom_row.  Returns the
// mask of fresh slots (bits 0..31).  PRE: the lane's two positions come
// precomputed in vp (hash 2 * half + t in bits 16 t .. 16 t + 15, see
// SearchArgs::vpos) instead of being hashed here.
template <bool PRE>
__device__ __forceinline__ uint64_t bloom_row_pairs(unsigned long long* bloom, uint32_t wmask, int wshift,
                                                    uint32_t nbits_mask, int sid, bool sv, int lane,
                                                    uint32_t vp) {
    const int half = lane >> 5;
    uint32_t word[2], bitn[2];
    unsigned long long pre[2];
    bool miss[2];
    bool mine = sv;
#pragma unroll
    for (int t = 0; t < 2; t++) {
        uint32_t pos = 0u;
        if constexpr (PRE) {
            pos = (vp >> (16 * t)) & nbits_mask;
        } else {
            if (sv) pos = half ? bloom_pos(2 + t, (uint32_t)sid, nbits_mask) : bloom_pos(t, (uint32_t)sid, nbits_mask);
        }
        word[t] = pos & wmask;
        bitn[t] = pos >> wshift;
        if constexpr (PRE) {
            // the word index is in range for any lane: read without a branch
            const unsigned long long w = bloom[word[t]];
            pre[t] = sv ? w : 0ull;
        } else {
            pre[t] = sv ? bloom[word[t]] : 0ull;
        }
        miss[t] = !((pre[t] >> bitn[t]) & 1ull);
        mine = mine && !miss[t];
    }
    const uint64_t hitb = __ballot(mine);
    const uint32_t all_hit = (uint32_t)hitb & (uint32_t)(hitb >> 32);  // per slot
    const bool fresh = sv && !((all_hit >> (lane & 31)) & 1u);
    bool dup = false;
    if (fresh) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (miss[t]) {
                unsigned long long bit = 1ull << bitn[t];
                unsigned long long old = atomicOr(&bloom[word[t]], bit);
                dup = dup || ((old & bit) != 0ull);
            }
        }
    }
    const uint64_t fresh_mask = __ballot(fresh) & 0xffffffffull;
    if (__ballot(dup) == 0ull) return fresh_mask;

    if (fresh) {
#pragma unroll
        for (int t = 0; t < 2; t++)
            if (miss[t]) bloom[word[t]] = pre[t];
    }
    wave_lds_sync();
    uint64_t acc = 0ull;
    uint64_t fm = fresh_mask;
    while (fm) {
        const int j = ffs64(fm);
        fm &= fm - 1ull;
        const bool me = (lane & 31) == j;
        bool hit = true;
        if (me) {
#pragma unroll
            for (int t = 0; t < 2; t++) hit = hit && ((bloom[word[t]] >> bitn[t]) & 1ull);
        }
        const uint64_t hb = __ballot(me && hit);
        const bool slot_hit = ((hb >> j) & 1ull) && ((hb >> (j + 32)) & 1ull);
        if (me && !slot_hit) {
            // both lanes of the slot may hit one word: atomic, not read-modify-write
#pragma unroll
            for (int t = 0; t < 2; t++) atomicOr(&bloom[word[t]], 1ull << bitn[t]);
        }
        wave_lds_sync();
        if (!slot_hit) acc |= 1ull << j;
    }
    return acc;
}

// ---------------------------------------------------------------- the two sets

// Set entries are compared as 32-bit integer keys, not floats: integer
// min / max fuse with DPP and need no NaN canonicalisation, which halves the
// cost of a wave reduction.  The key of a distance d is its bit pattern,
// mapped so that signed integer order is float order:
//  * L2 and Hamming distances are >= +0 (never -0, never negative), whose bit
//    patterns already order as integers -- the key is the bit pattern;
//  * inner-product distances (-dot) can be negative: negative patterns get
//    their magnitude bits flipped (an involution, so the same map decodes).
// The empty slot holds INT_MIN: below every key for a signed max, and (as
// 2^31) above every non-negative key for an unsigned min.
constexpr int KEY_EMPTY = (int)0x80000000u;

template <bool SIGNED>
__device__ __forceinline__ int dist_key(float d) {
    const int b = __float_as_int(d);
    if constexpr (SIGNED) return b ^ ((b >> 31) & 0x7fffffff);
    else return b;
}
template <bool SIGNED>
__device__ __forceinline__ float key_dist(int k) {
    if constexpr (SIGNED) return __int_as_float(k ^ ((k >> 31) & 0x7fffffff));
    else return __int_as_float(k);
}

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
// In-row steps as in wave_max, then row_bcast:15 (rows 1 and 3 take lane 15
// of the row below) and row_bcast:31 (rows 2 and 3 take lane 31): lane 63
// ends with the maximum of all four rows.  Rows a broadcast skips keep the
// identity, so the max leaves them as they were.
constexpr int DPP_WAVE_SHR1 = 0x138;   // lane i takes lane i - 1 across the whole wave
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;
__device__ __forceinline__ int wave_max_i(int v) {
    v = max(v, dpp_i<DPP_XOR1>(v));
    v = max(v, dpp_i<DPP_XOR2>(v));
    v = max(v, dpp_i<DPP_HALF_MIRROR>(v));
    v = max(v, dpp_i<DPP_MIRROR>(v));
    v = max(v, __builtin_amdgcn_update_dpp((int)0x80000000u, v, DPP_ROW_BCAST15, 0xA, 0xF, false));
    v = max(v, __builtin_amdgcn_update_dpp((int)0x80000000u, v, DPP_ROW_BCAST31, 0xC, 0xF, false));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ uint32_t wave_min_u(uint32_t v) {
    v = min(v, (uint32_t)dpp_i<DPP_XOR1>((int)v));
    v = min(v, (uint32_t)dpp_i<DPP_XOR2>((int)v));
    v = min(v, (uint32_t)dpp_i<DPP_HALF_MIRROR>((int)v));
    v = min(v, (uint32_t)dpp_i<DPP_MIRROR>((int)v));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, DPP_ROW_BCAST15, 0xA, 0xF, false));
    v = min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, DPP_ROW_BCAST31, 0xC, 0xF, false));
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Unsorted bounded set of (dist, id): the key of entry i in VGPR slot i/64 of
// lane i%64, its id in LDS at ids[i] (in a VGPR when S = 1).  Every scan (min, max, first index of a
// key) runs on the keys; ids are only read or written at one known index, so
// they cost one LDS access instead of a select per slot.  Slots at or beyond
// sz hold KEY_EMPTY, which no scan picks and == never matches, so scans need
// no validity test.  The maximum is cached (wave uniform) and recomputed only
// after it can have changed.  SIGNED: keys of inner-product distances.
template <int S, bool SIGNED>
struct WSet {
    // ext_vector_type keeps the slots in VGPRs: a plain array gets its
    // per-slot selects folded back into an indexed scratch access.
    typedef int kvec __attribute__((ext_vector_type(S == 1 ? 2 : S)));
    kvec d;
    // Ids: with one slot per lane (S = 1) entry i's id sits in lane i of idv,
    // read with one readlane and written with one select (no LDS round trip);
    // with more slots in this wave's LDS array of 64 * S ids.
    static constexpr bool VID = S == 1;
    int idv;
    int* ids;
    int sz;    // wave-uniform
    int cap;   // wave-uniform
    int mx;    // cached maximum key, valid when mx_ok
    bool mx_ok;

    __device__ __forceinline__ void init(int c, int* lds_ids) {
        ids = lds_ids;
        idv = -1;
        sz = 0;
        cap = c < 1 ? 1 : c;
        mx_ok = false;
        mx = KEY_EMPTY;
#pragma unroll
        for (int s = 0; s < S; s++) d[s] = KEY_EMPTY;
    }
    __device__ __forceinline__ int vmax(int lane) const {
        int m = d[0];
#pragma unroll
        for (int s = 1; s < S; s++) m = max(m, d[s]);
        return wave_max_i(m);
    }
    __device__ __forceinline__ int cmax(int lane) {
        if (!mx_ok) {
            mx = vmax(lane);
            mx_ok = true;
        }
        return mx;
    }
    // Minimum key of the slots k (the set's own or a copy of them).
    __device__ __forceinline__ static int kmin(const kvec& k) {
        uint32_t m = 0xffffffffu;
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint32_t u = (uint32_t)k[s];
            // signed keys: move to unsigned order, the empty slot to the top
            if constexpr (SIGNED) u = k[s] == KEY_EMPTY ? 0xffffffffu : u ^ 0x80000000u;
            m = min(m, u);
        }
        m = wave_min_u(m);
        return SIGNED ? (int)(m ^ 0x80000000u) : (int)m;
    }
    __device__ __forceinline__ int vmin(int lane) const { return kmin(d); }
    // First index whose key equals v (v is a key present in the set).
    __device__ __forceinline__ static int first_eq_of(const kvec& k, int v) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint64_t b = __ballot(k[s] == v);
            if (b) return s * 64 + ffs64(b);
        }
        return 0;
    }
    __device__ __forceinline__ int first_eq(int v, int lane) const { return first_eq_of(d, v); }
    // Slot access at a wave-uniform runtime index: a compare-and-branch per
    // slot up to the one addressed, which then costs one VALU operation --
    // a select chain would cost an operation per slot plus scalar mask work.
    // (The empty asm keeps the compiler from folding the branches back into
    // such a chain.)
    __device__ __forceinline__ int get_d(int idx) const {
        const int slot = idx >> 6, ln = idx & 63;
        int r = KEY_EMPTY;
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (slot == s) {
                asm volatile("");
                r = __builtin_amdgcn_readlane(d[s], ln);
            }
        }
        return r;
    }
    // Id of entry idx (a wave-uniform index).
    __device__ __forceinline__ int get_id(int idx) const {
        if constexpr (VID) return __builtin_amdgcn_readlane(idv, idx);
        // every lane reads the same word: readfirstlane makes the value scalar, so the
        // branches and row addresses that use it stay scalar too
        else return __builtin_amdgcn_readfirstlane(ids[idx]);
    }
    __device__ __forceinline__ void set_id(int idx, int iv, int lane) {
        if constexpr (VID) idv = lane == idx ? iv : idv;
        else if (lane == 0) ids[idx] = iv;
    }
    __device__ __forceinline__ void get(int idx, int& kv, int& iv) const {
        kv = get_d(idx);
        iv = get_id(idx);
    }
    // Only lane idx%64 of slot idx/64 changes; one lane writes the id.
    __device__ __forceinline__ void put_d(int idx, int kv, int lane) {
        const int slot = idx >> 6;
        const bool mine = lane == (idx & 63);
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (slot == s) {
                asm volatile("");
                d[s] = mine ? kv : d[s];
            }
        }
    }
    __device__ __forceinline__ void put(int idx, int kv, int iv, int lane) {
        put_d(idx, kv, lane);
        set_id(idx, iv, lane);
    }
    // Remove entry p, filling the hole with the last entry (gpuHeap.h:44-48);
    // the vacated last slot becomes empty.  The caller keeps mx_ok as fits.
    __device__ __forceinline__ void remove_at(int p, int lane) { remove_at(p, get_id(sz - 1), lane); }
    // The same with the last entry's id already read (li = ids[sz - 1]).
    __device__ __forceinline__ void remove_at(int p, int li, int lane) {
        const int last = sz - 1;
        const int lk = get_d(last);
        put(p, lk, li, lane);
        put_d(last, KEY_EMPTY, lane);
        sz--;
    }
    // Replace the first entry holding key M (present in the set) by (kv, iv):
    // the first-index search and the write in one pass, so the write happens
    // in the branch of the slot found, at a slot fixed at compile time.
    __device__ __forceinline__ void replace_first(int M, int kv, int iv, int lane) {
        bool done = false;
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (!done) {
                const uint64_t b = __ballot(d[s] == M);
                if (b) {
                    const int l = ffs64(b);
                    d[s] = lane == l ? kv : d[s];
                    set_id(s * 64 + l, iv, lane);
                    done = true;
                }
            }
        }
    }
    // Append the first min(F, cap - sz) items (lanes 0..) while the set is not
    // full, as one lane rotation; returns how many.
    __device__ __forceinline__ int append_items(int it_k, int it_id, int F, int lane) {
        const int a = min(F, cap - sz);
        if (a > 0) {
            const int src = (lane - sz) & 63;
            const int mk = __shfl(it_k, src, 64);
            const int slot = src < a ? (sz + src) >> 6 : -1;
#pragma unroll
            for (int s = 0; s < S; s++) d[s] = slot == s ? mk : d[s];
            if constexpr (VID) {
                const int mi = __shfl(it_id, src, 64);
                idv = src < a ? mi : idv;
            } else {
                if (lane < a) ids[sz + lane] = it_id;
            }
            sz += a;
            mx_ok = false;
        }
        return a > 0 ? a : 0;
    }
    // Push items it_k/it_id held by lanes 0..F-1, in lane order, with the
    // Candidates::push / Results::try_push rule (gpuHeap.h:23-36, 78-87).
    // Returns how many were accepted.
    __device__ __forceinline__ int push_items(int it_k, int it_id, int F, int lane) {
        const int a = append_items(it_k, it_id, F, lane);
        int accepted = a;
        if (F > a) {
            int M = cmax(lane);
            int start = a;
            for (;;) {
                const uint64_t m = __ballot(lane >= start && lane < F && it_k < M);
                if (m == 0ull) break;
                const int r = ffs64(m);
                const int kv = __builtin_amdgcn_readlane(it_k, r);
                const int iv = __builtin_amdgcn_readlane(it_id, r);
                replace_first(M, kv, iv, lane);
                accepted++;
                mx_ok = false;
                start = r + 1;
                if (start >= F) break;
                M = cmax(lane);
            }
        }
        return accepted;
    }
};

// ---------------------------------------------------------------- distances

// Distances of the rows stage_id[0..F) to the query; lane 0 of each group of
// G lanes writes stage_d[r].  Lane t of a group accumulates the 16-byte
// chunks c = t, t+G, ... (c < dstride/4) of the row in order, with explicit
// fmaf; the group sum is the xor butterfly of group_sum.  distance_passes
// handles passes p0 .. p0+NP-1 (64/G rows each) with all their loads issued
// before the first sum; `shadow` runs between the loads and the sums (work
// that does not depend on the rows, hidden behind their memory latency).
// Groups past the last row read row 0 and drop its sum, so no load waits
// behind a branch.  FULL: the group's lanes cover the row exactly
// (G * QCH == dstride / 4), so no chunk needs a range test.  OFF32: row
// offsets fit 31 bits and node ids 23 bits (the host checks), so a row's
// byte offset is one 24-bit multiply and the load takes a 32-bit offset from
// the base address instead of a 64-bit address per lane.
template <int METRIC, int G, int QCH, int NP, bool FULL, bool OFF32, class Shadow>
__device__ __forceinline__ void distance_passes(const float4* __restrict__ data, int d4, const float4 (&q)[QCH],
                                                int F, int p0, const int* stage_id, float* stage_d, int lane,
                                                Shadow&& shadow) {
    constexpr int NG = 64 / G;
    const int grp = lane / G, t = lane % G;
    float4 x[NP][QCH];
    int rr[NP];
#pragma unroll
    for (int pp = 0; pp < NP; pp++) {
        const int r = (p0 + pp) * NG + grp;  // < 64: passes cover at most the 64 staged rows
        rr[pp] = r;
        const int sid = stage_id[r];
        const int node = r < F ? sid : 0;
        if constexpr (OFF32) {
            const uint32_t off = (uint32_t)__mul24(node, d4 * 16) + (uint32_t)t * 16u;
            const char* row = (const char*)data + off;
#pragma unroll
            for (int i = 0; i < QCH; i++) {
                if (FULL || t + i * G < d4) x[pp][i] = *(const float4*)(row + i * G * 16);
                else x[pp][i] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        } else {
            const float4* row = data + (size_t)node * d4;
#pragma unroll
            for (int i = 0; i < QCH; i++) {
                const int c = t + i * G;
                if (FULL || c < d4) x[pp][i] = row[c];
                else x[pp][i] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
        }
    }
    // every row load issues before any sum: the scheduler would otherwise
    // start the first pass's sums, waiting on its loads, before issuing the
    // next pass's loads
    __builtin_amdgcn_sched_barrier(0);
    shadow();
#pragma unroll
    for (int pp = 0; pp < NP; pp++) {
        float acc = 0.f;
#pragma unroll
        for (int i = 0; i < QCH; i++) {
            const int c = t + i * G;
            if (FULL || c < d4) {
                if constexpr (METRIC == METRIC_IP) {
                    acc = __builtin_fmaf(q[i].x, x[pp][i].x, acc);
                    acc = __builtin_fmaf(q[i].y, x[pp][i].y, acc);
                    acc = __builtin_fmaf(q[i].z, x[pp][i].z, acc);
                    acc = __builtin_fmaf(q[i].w, x[pp][i].w, acc);
                } else {
                    float e;
                    e = q[i].x - x[pp][i].x;
                    acc = __builtin_fmaf(e, e, acc);
                    e = q[i].y - x[pp][i].y;
                    acc = __builtin_fmaf(e, e, acc);
                    e = q[i].z - x[pp][i].z;
                    acc = __builtin_fmaf(e, e, acc);
                    e = q[i].w - x[pp][i].w;
                    acc = __builtin_fmaf(e, e, acc);
                }
            }
        }
        acc = group_sum<G>(acc);
        if (t == 0 && rr[pp] < F) stage_d[rr[pp]] = METRIC == METRIC_IP ? -acc : acc;
    }
}

// Passes in flight: two for rows shared by several lane groups; with one row
// per pass (64 lanes per row: long rows, where few waves run) up to 8, held to
// about 128 VGPRs of row data.
__host__ __device__ constexpr int rows_in_flight(int g, int qch) { return g == 64 && qch <= 4 ? 8 : 2; }

// NP passes at a time while that many remain, then one at a time (a row's sum
// does not depend on which passes share a step).  `shadow` runs once, behind
// the first step's loads.
template <int METRIC, int G, int QCH, int NP, bool FULL, bool OFF32, class Shadow>
__device__ __forceinline__ void float_distances_v(const float4* __restrict__ data, int d4,
                                                  const float4 (&q)[QCH], int F,
                                                  const int* stage_id, float* stage_d, int lane, Shadow&& shadow) {
    constexpr int NG = 64 / G;
    const int passes = (F + NG - 1) / NG;
    auto none = []() {};
    int p0 = 0;
    if constexpr (NP == 8) {
        // one row per pass (long rows): the largest step of 8, 4, 2 or 1 passes that fits,
        // then steps of 8, 4, 2, 1 for the rest, so a hop's rows take at most four round trips
        const int first = passes >= 8 ? 8 : passes >= 4 ? 4 : passes >= 2 ? 2 : 1;
        if (first == 8) distance_passes<METRIC, G, QCH, 8, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        else if (first == 4) distance_passes<METRIC, G, QCH, 4, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        else if (first == 2) distance_passes<METRIC, G, QCH, 2, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        else distance_passes<METRIC, G, QCH, 1, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        p0 = first;
        for (; passes - p0 >= 8; p0 += 8)
            distance_passes<METRIC, G, QCH, 8, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
        if (passes - p0 >= 4) {
            distance_passes<METRIC, G, QCH, 4, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
            p0 += 4;
        }
        if (passes - p0 >= 2) {
            distance_passes<METRIC, G, QCH, 2, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
            p0 += 2;
        }
        if (passes - p0 >= 1)
            distance_passes<METRIC, G, QCH, 1, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
        return;
    }
    if (passes >= NP) {
        distance_passes<METRIC, G, QCH, NP, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        for (p0 = NP; p0 + NP <= passes; p0 += NP)
            distance_passes<METRIC, G, QCH, NP, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
    } else {
        distance_passes<METRIC, G, QCH, 1, FULL, OFF32>(data, d4, q, F, 0, stage_id, stage_d, lane, shadow);
        p0 = 1;
    }
    if constexpr (NP > 1)
        for (; p0 < passes; p0++)
            distance_passes<METRIC, G, QCH, 1, FULL, OFF32>(data, d4, q, F, p0, stage_id, stage_d, lane, none);
}

// off32: the index's rows are addressed with 32-bit offsets (SearchArgs::off32).
template <int METRIC, int G, int QCH, int NP, class Shadow>
__device__ __forceinline__ void float_distances(const float4* __restrict__ data, int d4, bool off32,
                                                const float4 (&q)[QCH], int F,
                                                const int* stage_id, float* stage_d, int lane, Shadow&& shadow) {
    if (off32 && d4 == G * QCH)
        float_distances_v<METRIC, G, QCH, NP, true, true>(data, d4, q, F, stage_id, stage_d, lane, shadow);
    else if (off32)
        float_distances_v<METRIC, G, QCH, NP, false, true>(data, d4, q, F, stage_id, stage_d, lane, shadow);
    else
        float_distances_v<METRIC, G, QCH, NP, false, false>(data, d4, q, F, stage_id, stage_d, lane, shadow);
}

// Hamming on 256-bit codes (4 x u64): two lanes per code, 16 bytes each
// (warp_hamming, hashKernelGraph.h:215-224).
template <class Shadow>
__device__ __forceinline__ void hamming_distances(const ulonglong2* __restrict__ codes, ulonglong2 q,
                                                  int F, const int* stage_id, float* stage_d, int lane,
                                                  Shadow&& shadow) {
    const int grp = lane >> 1, t = lane & 1;
    for (int p = 0; p * 32 < F; p++) {
        const int r = p * 32 + grp;
        const int node = r < F ? stage_id[r] : -1;
        ulonglong2 x = make_ulonglong2(0ull, 0ull);
        if (node >= 0) x = codes[(size_t)node * 2 + t];
        if (p == 0) shadow();
        int c = __popcll(q.x ^ x.x) + __popcll(q.y ^ x.y);
        c = c + dppi<DPP_XOR1>(c);
        if (t == 0 && r < F) stage_d[r] = (float)c;
    }
}

// ---------------------------------------------------------------- learned entry / termination

// The stop network's weights in the kernel's layout (the host converts them
// once, stop_weights_to_kernel_layout): w1 [64][8], b1 [64], w2 interleaved so
// lane j reads 16 conflict-free, coalesced float4 (element (j, i) at
// (i/4)*256 + j*4 + i%4), b2 [64], w3 [64].  The search reads them from
// global memory: 19.2 KB, shared by every wave of a compute unit through its
// L1, so LDS holds only the waves' own state (staged per workgroup, the
// weights would leave LDS room for two 8-wave workgroups per CU, 16 waves).
// stage_stop_weights serves the stand-alone inference kernel.
constexpr int SW_B1 = LM_H * LM_F;         // b1 [64]
constexpr int SW_W2 = SW_B1 + LM_H;        // w2 [64][64], interleaved
constexpr int SW_B2 = SW_W2 + LM_H * LM_H; // b2 [64]
constexpr int SW_W3 = SW_B2 + LM_H;        // w3 [64]

__device__ __forceinline__ void stage_stop_weights(const float* __restrict__ src, float* lds) {
    for (int i = threadIdx.x; i < LM_STOP_FLOATS; i += blockDim.x) lds[i] = src[i];
}

// logit of the stop network for wave-uniform features f[0..7].  W: the
// weights in the kernel layout (global memory or LDS).  hbuf: 64 floats of
// this wave's LDS; the hidden layer goes through it so that every lane reads
// it back as broadcast float4s (no per-unit lane reads).  The weights' base
// address passes through an empty asm so that the compiler forms each load's
// address here, from the wave-uniform base plus the lane's offset, instead of
// hoisting 16 per-lane 64-bit addresses out of the search loop; the second
// layer's loop stays rolled (one row block in flight), which keeps the
// search within its register budget.
__device__ __forceinline__ float stop_logit(const float* __restrict__ W, const float (&f)[LM_F], float b3,
                                            int lane, float* hbuf) {
    asm volatile("" : "+s"(W));
    const float4* w1 = (const float4*)W;
    const float4 wa = w1[2 * lane];
    const float4 wb = w1[2 * lane + 1];
    float h = W[SW_B1 + lane];
    h = __builtin_fmaf(wa.x, f[0], h);
    h = __builtin_fmaf(wa.y, f[1], h);
    h = __builtin_fmaf(wa.z, f[2], h);
    h = __builtin_fmaf(wa.w, f[3], h);
    h = __builtin_fmaf(wb.x, f[4], h);
    h = __builtin_fmaf(wb.y, f[5], h);
    h = __builtin_fmaf(wb.z, f[6], h);
    h = __builtin_fmaf(wb.w, f[7], h);
    hbuf[lane] = fmaxf(h, 0.f);
    wave_lds_sync();
    float acc = W[SW_B2 + lane];
    const float4* w2 = (const float4*)(W + SW_W2);
#pragma unroll 1
    for (int i4 = 0; i4 < 16; i4++) {
        const float4 w = w2[i4 * 64 + lane];
        const float4 x = *(const float4*)(hbuf + 4 * i4);
        acc = __builtin_fmaf(w.x, x.x, acc);
        acc = __builtin_fmaf(w.y, x.y, acc);
        acc = __builtin_fmaf(w.z, x.z, acc);
        acc = __builtin_fmaf(w.w, x.w, acc);
    }
    wave_lds_sync();
    acc = fmaxf(acc, 0.f);
    return wave_sum(acc * W[SW_W3 + lane]) + b3;
}

// The k smallest distances seen so far, ascending in lanes 0..k-1 (k <= 64):
// the k best of the result set (no value among them can have been evicted).
struct TopK {
    float v;     // this lane's entry (+inf beyond what has been seen)
    float dk;    // wave-uniform: the k-th smallest (+inf until k are seen)
    float d1;    // wave-uniform: the smallest
    int k;
    int inserts; // insertions since the last read

    __device__ __forceinline__ void init(int k_) {
        v = INFINITY;
        dk = INFINITY;
        d1 = INFINITY;
        k = k_;
        inserts = 0;
    }
    // Insert the distances of lanes 0..F-1 in lane order.
    __device__ __forceinline__ void insert(float it_d, int F, int lane) {
        uint64_t m = __ballot(lane < F && it_d < dk);
        while (m) {
            const int r = ffs64(m);
            const float d = readlane_f(it_d, r);
            const int pos = __popcll(__ballot(lane < k && v <= d));
            // v of lane - 1 (wave_shr:1; lane 0 keeps its own, which it never uses)
            const float up = __int_as_float(
                __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), DPP_WAVE_SHR1, 0xF, 0xF, false));
            if (lane < k) v = lane < pos ? v : (lane == pos ? d : up);
            dk = readlane_f(v, k - 1);
            d1 = readlane_f(v, 0);
            inserts++;
            m = __ballot(lane > r && lane < F && it_d < dk);
        }
    }
};

// ---------------------------------------------------------------- one query

// The search of query qid by the calling wave (kg_search_kernel body,
// kernalGraph.h:192-298).  lds: this wave's bloom[W] | stage_id[64] | stage_d[64] | set ids.
template <int METRIC, int S, int G, int QCH, int LEARN>
__device__ __forceinline__ void search_query(const SearchArgs& P, int qid, int lane, unsigned long long* lds,
                                             const float* stop_lds) {
    unsigned long long* bloom = lds;
    int* stage_id = (int*)(bloom + P.bloom_words);
    float* stage_d = (float*)(stage_id + 64);
    int* set_ids = (int*)(stage_d + 64);  // candidates' ids [64 * S] | results' ids [64 * S]
    const uint32_t wmask = (uint32_t)P.bloom_words - 1u;
    const int wshift = P.bloom_shift;
    const uint32_t nbits_mask = (uint32_t)P.bloom_words * 64u - 1u;
    // per-query budgets (serving / persistent queues mix them); ef >= k.  The
    // host checks every budget before the launch (1 <= k <= k_stride,
    // max(ef, k) <= P.ef); the clamps only keep a budget that bypassed those
    // checks inside the output row and the register sets.
    const int k = max(1, min(P.ks ? P.ks[qid] : P.k, P.k_stride));
    const int ef = min(P.efs ? max(P.efs[qid], k) : P.ef, min(P.ef, 64 * S));

    for (int w = lane; w < P.bloom_words; w += 64) bloom[w] = 0ull;

    // query in registers (each lane holds the chunks it multiplies)
    float4 qf[QCH];
    ulonglong2 qc = make_ulonglong2(0ull, 0ull);
    const int d4 = P.dstride >> 2;
    if constexpr (METRIC == METRIC_HAMMING) {
        qc = ((const ulonglong2*)P.queries)[(size_t)qid * 2 + (lane & 1)];
    } else {
        const float4* qrow = (const float4*)P.queries + (size_t)qid * d4;
        const int t = lane % G;
#pragma unroll
        for (int i = 0; i < QCH; i++) {
            const int c = t + i * G;
            qf[i] = c < d4 ? qrow[c] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    wave_lds_sync();

    constexpr bool SK = METRIC == METRIC_IP;
    WSet<S, SK> cand, res;
    // Learned stop with a result cap below ef: the result set keeps only
    // result_cap (>= k) entries, and the reference's early stop -- which
    // compares with the worst of ef results -- is left to the stop network
    // and the hop cap (the host requires one).  The traversal (candidates,
    // Bloom) does not depend on the result set, and its k best are the same.
    // LEARN 1: the learned stop; 3: the same with a late entry (its own
    // instantiation, so the plain learned search keeps its registers);
    // 2: record the stop's features instead
    constexpr bool STOPL = LEARN == 1 || LEARN == 3;
    const bool small_res = STOPL && P.L.result_cap > 0 && P.L.result_cap < ef;
    cand.init(ef, set_ids);
    res.init(small_res ? max(P.L.result_cap, k) : ef, set_ids + 64 * S);
    // QueryTrace, telemetry.h:5-18 (wave-uniform counters)
    // (distance_evals = visited, candidate_nodes = visited - seeds,
    // neighbor_slots = block per expansion that read its row, and the result
    // set never shrinks before the trim, so its peak is its final size: those
    // four are derived at the end instead of carried through the loop.)
    int tr_seed = 0, tr_visited = 0, tr_expanded = 0;
    int tr_fpush = 0, tr_rupd = 0, tr_peakf = 0, tr_early = 0;
    // learned termination state (k <= 64 on this path)
    // LEARN 1: the learned stop; LEARN 2: record its features instead
    const bool tracing = LEARN == 2 && P.L.trace != nullptr;
    const bool stopping = STOPL && P.L.use_stop;
    TopK tk;
    tk.init(k);
    float dk_prev = INFINITY;
    int n_evals = 0;
    // consulted hops: min_hops, min_hops + interval, ... (hop counts start at
    // 1, so with min_hops 0 the first is interval)
    int next_eval = LEARN ? (P.L.min_hops > 0 ? P.L.min_hops : P.L.interval) : 0;
    int stop_votes = 0;
    int last_insert_hop = 0;
    // feature scales (learned search only): 1 / distance scale, 1 / k, 1 / ef
    float inv_sc = 0.f, inv_k = 0.f, inv_ef = 0.f;
    if constexpr (LEARN) {
        inv_sc = 1.0f / P.L.dist_scale;
        inv_k = 1.0f / (float)k;
        inv_ef = 1.0f / (float)ef;
    }

    // Each expansion is gather (Bloom test-then-set of a row of ids in lane
    // order, distances of the fresh ones: lane r < F holds the r-th fresh id
    // and its key), then the pushes of those F items to both sets in order.
    int it_k = 0, it_id = -1;
    float it_d = INFINITY;
    // id / valid: slot j of the row in lane j, and for rows of at most 32
    // slots also in lane j + 32 (see bloom_row_pairs).
    // shadow: work that does not depend on the row, run behind the row loads
    // vp: the slot's precomputed filter positions when `pre` (adjacency rows
    // of at most 32 slots with a position table); seeds are hashed here.
    auto gather_row = [&](int id, bool valid, int width, bool pre, uint32_t vp,
                          auto&& shadow) __attribute__((always_inline)) -> int {
        const uint64_t acc =
            pre ? bloom_row_pairs<true>(bloom, wmask, wshift, nbits_mask, id, valid, lane, vp)
                : width <= 32 ? bloom_row_pairs<false>(bloom, wmask, wshift, nbits_mask, id, valid, lane, 0u)
                              : bloom_row(bloom, wmask, wshift, nbits_mask, id, valid);
        const int F = __popcll(acc);
        if (F == 0) {
            shadow();
            return 0;
        }
        const int rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(acc >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)acc, 0u));
        if ((acc >> lane) & 1ull) stage_id[rank] = id;
        wave_lds_sync();
        if constexpr (METRIC == METRIC_HAMMING) {
            hamming_distances((const ulonglong2*)P.data, qc, F, stage_id, stage_d, lane, shadow);
        } else {
            float_distances<METRIC, G, QCH, rows_in_flight(G, QCH)>((const float4*)P.data, d4, P.off32 != 0, qf, F,
                                                                   stage_id, stage_d, lane,
                                               shadow);
        }
        wave_lds_sync();
        // read unconditionally (64 staged entries), then keep the first F
        const float sd = stage_d[lane];
        const int si = stage_id[lane];
        it_d = lane < F ? sd : INFINITY;
        it_id = lane < F ? si : -1;
        it_k = dist_key<SK>(it_d);
        return F;
    };
    auto push_row = [&](int F) __attribute__((always_inline)) {
        if (F > 0) {
            tr_fpush += cand.push_items(it_k, it_id, F, lane);
            tr_rupd += res.push_items(it_k, it_id, F, lane);
            tr_peakf = max(tr_peakf, cand.sz);
            if constexpr (LEARN) {
                const int before = tk.inserts;
                tk.insert(it_d, F, lane);
                if (tk.inserts != before) last_insert_hop = tr_expanded;
            }
        }
        wave_lds_sync();
    };

    // The next pop, known before the pushes: the smallest key of the
    // candidates after them is the smaller of (the set's minimum before them,
    // the best fresh key) -- a push only evicts a maximum, so an old minimum
    // stays unless every entry equals it, and the best fresh item is accepted
    // unless every entry already equals it.  The pop then only looks up the
    // first index of that key.  The adjacency row of the node that holds it
    // is loaded before the pushes, so its latency hides behind them; with
    // ties the pop can land on another node, which then loads its own row.
    int nx_key = 0;
    int pf_node = -1, nb_pref = -1;
    uint32_t vp_pref = 0u;
    // the adjacency slot this lane loads: rows of at most 32 slots are loaded
    // into both half-waves (the layout bloom_row_pairs takes), and with the
    // position table lane j of half h also loads plane h of slot j
    const int slot = P.block <= 32 ? (lane & 31) : lane;
    const bool use_vpos = P.vpos != nullptr && P.block <= 32;
    const int vslot = (lane >> 5) * P.block + slot;  // within a node's 2 x block words
    auto load_row = [&](int node, int& nbv, uint32_t& vpv) __attribute__((always_inline)) {
        nbv = slot < P.block ? P.adj[(size_t)node * P.block + slot] : -1;
        if (use_vpos) vpv = slot < P.block ? P.vpos[(size_t)node * 2 * P.block + vslot] : 0u;
    };
    auto predict = [&](bool old_ok, int old_min, int old_id, int F) __attribute__((always_inline)) {
        bool take_new = false;
        int new_min = 0, new_id = -1;
        if (F > 0) {
            const bool in = lane < F;
            uint32_t u = (uint32_t)it_k;
            if constexpr (SK) u ^= 0x80000000u;
            new_min = (int)wave_min_u(in ? u : 0xffffffffu);
            if constexpr (SK) new_min = (int)((uint32_t)new_min ^ 0x80000000u);
            new_id = __builtin_amdgcn_readlane(it_id, ffs64(__ballot(in && it_k == new_min)));
            take_new = !old_ok || new_min < old_min;
        }
        if (!take_new && !old_ok) return;  // the candidates end empty
        nx_key = take_new ? new_min : old_min;
        pf_node = take_new ? new_id : old_id;
        load_row(pf_node, nb_pref, vp_pref);
    };

    // Seeds: kernalGraph.h:215-231 -- ids (size_t)p * n / n_seeds, or the
    // caller's (or the planning pass's) entry list; with a late entry the
    // reference seeds first and the entry list at hop entry_late_hops.
    bool late_entry = false;
    if constexpr (LEARN == 3) late_entry = P.L.entry_late_hops > 0 && P.entries != nullptr;
    {
        const int e = P.n_entry <= 32 ? (lane & 31) : lane;
        int sid = 0;
        bool sv = e < P.n_entry;
        if (sv) {
            if (P.entries && !late_entry) {
                sid = P.entries[(size_t)qid * P.n_entry + e];
                sv = sid >= 0 && sid < P.n;
            } else {
                sid = (int)(((uint64_t)e * (uint64_t)P.n) / (uint64_t)P.n_entry);
            }
        }
        const int F = gather_row(sid, sv, P.n_entry, false, 0u, []() {});
        tr_seed += F;
        tr_visited += F;
        predict(false, 0, -1, F);
        push_row(F);
    }

    // Main loop: kernalGraph.h:233-289.
    for (;;) {
        if (cand.sz == 0) break;
        const int c_min = nx_key;
        const int p = cand.first_eq(c_min, lane);
        // both ids in one LDS round trip: the popped node and the last entry, which fills its hole
        const int cn = cand.get_id(p);
        const int li = cand.get_id(cand.sz - 1);
        int nb = nb_pref;
        uint32_t vp = vp_pref;
        if (cn != pf_node) load_row(cn, nb, vp);
        cand.remove_at(p, li, lane);  // removing the minimum leaves the maximum unless the set empties
        if (cand.sz == 0) cand.mx_ok = false;
        tr_expanded++;
        if (!small_res && res.sz == res.cap) {
            const int worst = res.cmax(lane);
            if (c_min > worst) {
                tr_early = 1;
                break;
            }
        }
        if constexpr (LEARN) {
            if (stopping && P.L.max_hops > 0 && tr_expanded > P.L.max_hops) {
                tr_early = 3;  // hop cap of the learned termination
                break;
            }
            // Learned termination: every `interval` hops from min_hops on, the
            // stop network sees where the search stands; trace mode records
            // the features instead (labels come from the k-th best at the end).
            // next_eval: the next consulted hop (min_hops, then every interval;
            // from late_hops on every late_interval)
            bool due = false;
            if (tr_expanded == next_eval) {
                due = true;
                next_eval += P.L.late_hops > 0 && tr_expanded >= P.L.late_hops ? P.L.late_interval : P.L.interval;
            }
            if ((tracing || stopping) && tk.dk < INFINITY && due && tr_expanded >= P.L.min_hops) {
                const int dkk = dist_key<SK>(tk.dk);
                int closer = 0;  // frontier entries still closer than the k-th best
#pragma unroll
                for (int s = 0; s < S; s++) closer += __popcll(__ballot(cand.d[s] != KEY_EMPTY && cand.d[s] < dkk));
                float f[LM_F];
                // products with reciprocals rounded once per query (no division per consult)
                f[0] = (float)tr_expanded * (1.0f / 64.0f);
                f[1] = (key_dist<SK>(c_min) - tk.dk) * inv_sc;
                f[2] = (tk.dk - tk.d1) * inv_sc;
                f[3] = (float)(tr_expanded - last_insert_hop) * (1.0f / 16.0f);
                f[4] = (float)closer * inv_k;
                f[5] = (float)tk.inserts * inv_k;
                f[6] = (float)cand.sz * inv_ef;
                f[7] = dk_prev < INFINITY ? (dk_prev - tk.dk) * inv_sc : 0.f;
#pragma unroll
                for (int i = 0; i < LM_F; i++) f[i] = fminf(fmaxf(f[i], -8.f), 8.f);
                dk_prev = tk.dk;
                tk.inserts = 0;
                if constexpr (LEARN == 2) {
                    if (n_evals < P.L.max_evals && lane <= LM_F) {
                        float v = tk.dk;
#pragma unroll
                        for (int i = 0; i < LM_F; i++)
                            if (lane == i) v = f[i];
                        P.L.trace[((size_t)qid * P.L.max_evals + n_evals) * (LM_F + 1) + lane] = v;
                    }
                    n_evals++;
                } else {
                    const float logit = stop_logit(stop_lds, f, P.L.s_b3, lane, stage_d);
                    // a NaN logit never stops; a stop needs `confirm` consecutive votes
                    stop_votes = logit > P.L.stop_logit ? stop_votes + 1 : 0;
                    if (stop_votes >= P.L.confirm) {
                        tr_early = 2;
                        break;
                    }
                }
            }
        }
        // the entry list joins (late entry): tested and pushed like a row,
        // before the popped node's row, whose next-pop prediction then sees it
        if constexpr (LEARN == 3) {
            if (late_entry && tr_expanded == P.L.entry_late_hops) {
                const int e = P.n_entry <= 32 ? (lane & 31) : lane;
                bool sv = e < P.n_entry;
                const int sid = sv ? P.entries[(size_t)qid * P.n_entry + e] : 0;
                sv = sv && sid >= 0 && sid < P.n;
                const int Fe = gather_row(sid, sv, P.n_entry, false, 0u, []() {});
                tr_visited += Fe;
                push_row(Fe);
            }
        }
        // the candidates' minimum without cn, for the next prediction, found
        // while the row's vectors are in flight
        const bool old_ok = cand.sz > 0;
        int old_min = 0, old_id = -1;
        const int F = gather_row(nb, nb >= 0, P.block, use_vpos, vp, [&]() __attribute__((always_inline)) {
            if (old_ok) {
                old_min = cand.vmin(lane);
                old_id = cand.get_id(cand.first_eq(old_min, lane));
            }
        });
        tr_visited += F;
        predict(old_ok, old_min, old_id, F);
        push_row(F);
    }
    const int tr_peakr = res.sz;

    // Trim to k, then drain best-first (kernalGraph.h:292-298, gpuHeap.h:90-108).
    // When each of the cnt smallest distances occurs once in the set, the
    // trim removes exactly the entries above the cnt-th smallest and the drain
    // returns the rest in ascending order, wherever they sit: extract them
    // directly, one per lane.  Any repeated value among them falls back to the
    // step-by-step replay, whose first-index rules decide ties.
    const int cnt = res.sz < k ? res.sz : k;
    int32_t* out = P.ids_out + (size_t)qid * P.k_stride;
    float* dout = P.dists_out ? P.dists_out + (size_t)qid * P.k_stride : nullptr;
    bool fast = cnt <= 64;
    if (fast) {
        typename WSet<S, SK>::kvec t = res.d;
        int od = 0;
        int oi = -1;
        for (int i = 0; i < cnt; i++) {
            const int m = WSet<S, SK>::kmin(t);
            int c = 0, p = -1;
#pragma unroll
            for (int s = 0; s < S; s++) {
                const uint64_t b = __ballot(t[s] == m);
                c += __popcll(b);
                if (p < 0 && b) p = s * 64 + ffs64(b);
            }
            if (c != 1) {
                fast = false;
                break;
            }
            const int iv = res.get_id(p);
            if (lane == i) {
                od = m;
                oi = iv;
            }
            const int slot = p >> 6;
            const bool mine = lane == (p & 63);
#pragma unroll
            for (int s = 0; s < S; s++) t[s] = (mine && slot == s) ? KEY_EMPTY : t[s];
        }
        if (fast && lane < cnt) {
            out[lane] = oi;
            if (dout) dout[lane] = key_dist<SK>(od);
        }
    }
    if (!fast) {
        while (res.sz > k) {
            const int M = res.vmax(lane);
            res.remove_at(res.first_eq(M, lane), lane);
        }
        res.mx_ok = false;
        for (int i = 0; i < cnt; i++) {
            const int m = res.vmin(lane);
            const int p = res.first_eq(m, lane);
            int kv, iv;
            res.get(p, kv, iv);
            res.remove_at(p, lane);
            if (lane == 0) {
                out[i] = iv;
                if (dout) dout[i] = key_dist<SK>(kv);
            }
        }
    }
    for (int i = cnt + lane; i < P.k_stride; i += 64) {
        out[i] = -1;
        if (dout) dout[i] = INFINITY;
    }
    if (tracing && lane == 0) {
        P.L.trace_count[qid] = min(n_evals, P.L.max_evals);
        P.L.final_dk[qid] = tk.dk;
    }
    if (P.trace_out && lane == 0) {
        int32_t* t = P.trace_out + (size_t)qid * 12;
        t[0] = tr_seed;
        t[1] = tr_visited;
        t[2] = tr_expanded;
        t[3] = (tr_expanded - (tr_early ? 1 : 0)) * P.block;
        t[4] = tr_visited - tr_seed;
        t[5] = tr_visited;
        t[6] = tr_fpush;
        t[7] = tr_rupd;
        t[8] = tr_peakf;
        t[9] = tr_peakr;
        t[10] = cnt;
        t[11] = tr_early;
    }
    wave_lds_sync();
}

// ---------------------------------------------------------------- kernels

// LDS of a workgroup: [stop weights when learned] then per wave
// bloom[W] | stage_id[64] | stage_d[64] | set ids [2 x 64 * S] (8-byte words).
template <int S>
__host__ __device__ constexpr size_t wave_lds_u64(int bloom_words) { return (size_t)bloom_words + 64 + 64 * S; }

// Register budget: with two 16-byte chunks per lane and the wave's LDS base
// addresses in scalar registers, the plain search fits in 72 VGPRs without a
// spill, so 7 waves share a SIMD (3 % faster than 6 at 80 VGPRs; before the
// LDS bases went scalar it needed 80 with spills).  The search with the
// learned stop is held to 80 with more set slots (6 waves), and with one set
// slot per lane (queue <= 64) to 72 (7 waves): 2 % faster than 6; 8 waves
// (64 VGPRs) gained nothing more.  A few of its per-query values spill, none
// inside the hop loop.  Four or more 16-byte chunks per lane (rows longer than
// 128 floats at 16 lanes, or 8 lanes per 128-float row) and the
// feature-recording search keep the compiler's own choice: with 8 lanes per
// 128-float row the plain search takes 94 VGPRs (5 waves per SIMD) and runs
// 2-3 % faster than at 16 lanes; held to 80 or 72 VGPRs it spilled and ran 11 %
// and 39 % slower.
__host__ __device__ constexpr int waves_per_simd(int qch, int learn, int s) {
    return qch > 2 ? 1 : learn == 0 ? 7 : learn == 1 || learn == 3 ? (s == 1 ? 7 : 6) : 1;
}

// One wave per query (the reference's one block per query, kernalGraph.h:354).
template <int METRIC, int S, int G, int QCH, int LEARN>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(waves_per_simd(QCH, LEARN, S))))
search_kernel(SearchArgs P) {
    extern __shared__ unsigned long long smem[];
    unsigned long long* waves_lds = smem;
    const float* stop_lds = LEARN == 1 || LEARN == 3 ? P.L.s_weights : nullptr;
    // the wave's index is the same in every lane: read it once so the wave's LDS base
    // addresses are scalar (derived from threadIdx.x they would count as varying and take
    // vector registers)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pos = blockIdx.x * (blockDim.x >> 6) + wave;
    if (pos >= P.nq) return;
    const int qid = P.order ? P.order[pos] : pos;
    search_query<METRIC, S, G, QCH, LEARN>(P, qid, lane_id(), waves_lds + (size_t)wave * wave_lds_u64<S>(P.bloom_words),
                                           stop_lds);
}

// The resident serving grid (a request ring in host memory, claimed wave by
// wave) is compiled out: in its device run some claimed requests read stale
// ring rows and some waves did not leave at shutdown.  The stale rows have a
// shown cause (no acquire between the claim and the request's reads, fixed
// below); the waves that stayed have none, so the host refuses the mode while
// this is false.
constexpr bool kResidentServing = false;

// Resident serving: the next published request, or -1 to leave (stop set, or
// nothing to claim for idle_ticks).  Lane 0 only.  Claims are exactly-once:
// head only moves by a successful compare-and-swap below tail.  The inner loop
// ends because every failed swap means another wave advanced head.  Polls are
// plain system-scope loads (no cache invalidation: the ring lives in coherent
// host memory, which the device does not cache), spaced by a growing sleep
// while idle, so idle waves cost neither L2 contents nor much bus traffic.
__device__ __forceinline__ int64_t resident_claim(const ResidentQueue& Q) {
    const uint64_t start = __builtin_amdgcn_s_memrealtime();
    int naps = 1;
    for (;;) {
        if (__hip_atomic_load(Q.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) return -1;
        const int64_t t = __hip_atomic_load(Q.tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        int64_t h = __hip_atomic_load(Q.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (h < t) {
            if (__hip_atomic_compare_exchange_strong(Q.head, &h, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return h;
        }
        if (__builtin_amdgcn_s_memrealtime() - start > (uint64_t)Q.idle_ticks) return -1;
        // about 3.4 us per nap; 1, 2, 4, ... 32 naps between polls
        for (int i = 0; i < naps; i++) __builtin_amdgcn_s_sleep(127);
        naps = naps < 32 ? 2 * naps : 32;
    }
}

// Persistent: a fixed grid of waves drains the query queue; each wave takes
// the next query index from a device counter until all nq are taken.  Every
// wave leaves once the counter passes nq, so the grid always drains.  With a
// resident queue the waves serve the request ring instead (resident_claim),
// publishing each request's results before its done flag.
template <int METRIC, int S, int G, int QCH, int LEARN>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(waves_per_simd(QCH, LEARN, S))))
persistent_kernel(SearchArgs P) {
    extern __shared__ unsigned long long smem[];
    unsigned long long* waves_lds = smem;
    const float* stop_lds = LEARN == 1 || LEARN == 3 ? P.L.s_weights : nullptr;
    // the wave's index is the same in every lane: read it once so the wave's LDS base
    // addresses are scalar (derived from threadIdx.x they would count as varying and take
    // vector registers)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = lane_id();
    unsigned long long* lds = waves_lds + (size_t)wave * wave_lds_u64<S>(P.bloom_words);
    if constexpr (kResidentServing) {
      if (P.resident) {
        const ResidentQueue Q = *P.resident;
        for (;;) {
            int64_t item = -1;
            if (lane == 0) item = resident_claim(Q);
            const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)item);
            const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)item >> 32));
            item = (int64_t)(((uint64_t)hi << 32) | lo);
            if (item < 0) break;
            // Every lane reads the request (row, k, queue size) after the host
            // wrote it: its acquire read of a tail that covers the request
            // synchronizes with the host's release store of that tail.  The
            // claim's relaxed read orders nothing; without this a wave read
            // the slot's previous request (tests/resident_ring_model.cpp shows
            // the race).  Tail only grows and lane 0 has read such a value, so
            // the first read ends the loop.
            while (__hip_atomic_load(Q.tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM) <= item)
                __builtin_amdgcn_s_sleep(1);
            const int slot = (int)(item % Q.ring);
            search_query<METRIC, S, G, QCH, LEARN>(P, slot, lane, lds, stop_lds);
            // every lane's ids / distances of the slot become visible to the host before its flag
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            if (lane == 0) __hip_atomic_store(Q.done + slot, item + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        return;
      }
    }
    for (;;) {
        int item = 0;
        if (lane == 0) item = atomicAdd(P.queue_head, 1);
        item = __builtin_amdgcn_readfirstlane(item);
        if (item >= P.nq) break;
        search_query<METRIC, S, G, QCH, LEARN>(P, P.order ? P.order[item] : item, lane, lds, stop_lds);
    }
}

static __global__ void zero_counter_kernel(int* c) {
    if (threadIdx.x == 0) *c = 0;
}

// ---------------------------------------------------------------- dispatch

template <int METRIC, int S, int G, int QCH, int LEARN>
static hipError_t launch_impl(const SearchArgs& a, const LaunchShape& ls, hipStream_t st) {
    const int threads = 64 * ls.waves_per_block;
    size_t lds = (size_t)ls.waves_per_block * wave_lds_u64<S>(a.bloom_words) * 8;
    if constexpr (LEARN == 2) {  // recording features: one launch shape (same results in every mode)
        const int blocks = (a.nq + ls.waves_per_block - 1) / ls.waves_per_block;
        hipLaunchKernelGGL((search_kernel<METRIC, S, G, QCH, 2>), dim3(blocks), dim3(threads), lds, st, a);
    } else if (kResidentServing && a.resident) {  // the ring's head persists across launches
        hipLaunchKernelGGL((persistent_kernel<METRIC, S, G, QCH, LEARN>), dim3(ls.persistent_blocks), dim3(threads),
                           lds, st, a);
    } else if (a.queue_head) {
        // no more workgroups than the queries need: with fewer queries than the grid holds, the
        // idle waves would only share SIMDs with the working ones
        const int need = (a.nq + ls.waves_per_block - 1) / ls.waves_per_block;
        const int blocks = std::max(1, std::min(ls.persistent_blocks, need));
        hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, st, a.queue_head);
        hipLaunchKernelGGL((persistent_kernel<METRIC, S, G, QCH, LEARN>), dim3(blocks), dim3(threads), lds, st, a);
    } else {
        const int blocks = (a.nq + ls.waves_per_block - 1) / ls.waves_per_block;
        hipLaunchKernelGGL((search_kernel<METRIC, S, G, QCH, LEARN>), dim3(blocks), dim3(threads), lds, st, a);
    }
    return hipGetLastError();
}

// The learned variants are built for rows of at most 4 chunks per lane
// (dim <= 1024); the host checks before launching.
template <int METRIC, int S, int G, int QCH>
static hipError_t launch_one(const SearchArgs& a, const LaunchShape& ls, hipStream_t st) {
    if (!a.learned) return launch_impl<METRIC, S, G, QCH, 0>(a, ls, st);
    if constexpr (QCH <= 4) {
        if (a.L.trace) return launch_impl<METRIC, S, G, QCH, 2>(a, ls, st);
        // the late entry is built for the short queues of the learned search only
        if constexpr (S == 1)
            if (a.L.entry_late_hops > 0 && a.entries) return launch_impl<METRIC, S, G, QCH, 3>(a, ls, st);
        return launch_impl<METRIC, S, G, QCH, 1>(a, ls, st);
    }
    return hipErrorInvalidValue;
}

template <int METRIC, int S>
static hipError_t launch_g(const SearchArgs& a, int G, int qch, const LaunchShape& wpb, hipStream_t st) {
    if constexpr (METRIC == METRIC_HAMMING) {
        return launch_one<METRIC, S, 2, 1>(a, wpb, st);
    } else {
        if (qch == 2) {
            switch (G) {
            case 2: return launch_one<METRIC, S, 2, 2>(a, wpb, st);
            case 4: return launch_one<METRIC, S, 4, 2>(a, wpb, st);
            case 8: return launch_one<METRIC, S, 8, 2>(a, wpb, st);
            case 16: return launch_one<METRIC, S, 16, 2>(a, wpb, st);
            case 32: return launch_one<METRIC, S, 32, 2>(a, wpb, st);
            case 64: return launch_one<METRIC, S, 64, 2>(a, wpb, st);
            }
        } else if (qch == 4) {
            switch (G) {
            case 2: return launch_one<METRIC, S, 2, 4>(a, wpb, st);
            case 4: return launch_one<METRIC, S, 4, 4>(a, wpb, st);
            case 8: return launch_one<METRIC, S, 8, 4>(a, wpb, st);
            case 16: return launch_one<METRIC, S, 16, 4>(a, wpb, st);
            case 32: return launch_one<METRIC, S, 32, 4>(a, wpb, st);
            case 64: return launch_one<METRIC, S, 64, 4>(a, wpb, st);
            }
        } else if (G == 64 && qch == 8) {
            return launch_one<METRIC, S, 64, 8>(a, wpb, st);
        } else if (G == 64 && qch == 16) {
            return launch_one<METRIC, S, 64, 16>(a, wpb, st);
        }
        return hipErrorInvalidValue;
    }
}

}  // namespace song
