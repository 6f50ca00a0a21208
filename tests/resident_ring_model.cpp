// CPU model of the resident serving ring: the claim / publish / retire protocol
// of persistent_kernel's resident branch (resident_claim, the done flag) and of
// the host's dispatcher_loop_resident, with host threads standing for the
// waves.  Each side keeps the memory orders of the code it models; the order of
// the waves' tail read is a compile-time choice (TAIL_ORDER), so that
// ThreadSanitizer can judge it:
//   * relaxed (the kernel as it was): nothing orders a wave's reads of the
//     request row, k and queue size after the host's writes of them, so a
//     wave can read the slot's previous request -- the stale rows seen on the
//     device;
//   * acquire: the read of a tail that covers the claimed request
//     synchronizes with the host's release store of that tail, which follows
//     the host's writes of the row.
// The device's release fence before its relaxed store of the done flag is
// modelled as a release store (ThreadSanitizer does not follow stand-alone
// fences; the two are equivalent for the flag's readers).
//
// Built and run by tests/test_resident_model.py with g++ -fsanitize=thread.
// Exit status: 0 when every request got its own result, 3 otherwise.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <thread>
#include <vector>

#ifndef TAIL_ORDER
#define TAIL_ORDER std::memory_order_acquire
#endif

namespace {

constexpr int kRing = 8;
constexpr int kRow = 16;
constexpr int kWaves = 4;
constexpr int64_t kRequests = 4000;

struct Ring {
    // host-written request rows (coherent host memory on the device)
    float q[kRing][kRow];
    int32_t ks[kRing];
    int32_t efs[kRing];
    // device-written results
    float out[kRing];
    std::atomic<int64_t> tail{0};  // requests published
    std::atomic<int64_t> head{0};  // next request to claim (device memory)
    std::atomic<int64_t> done[kRing];
    std::atomic<int32_t> stop{0};
};

float expected(int64_t req) {
    float s = 0.f;
    for (int j = 0; j < kRow; j++) s += (float)(req % 1000) + (float)j;
    return s + (float)(req % 7) + 2.f * (float)(req % 5);
}

// resident_claim: the next published request, or -1 once stop is set
int64_t claim(Ring& r) {
    for (;;) {
        if (r.stop.load(std::memory_order_relaxed)) return -1;
        const int64_t t = r.tail.load(TAIL_ORDER);
        int64_t h = r.head.load(std::memory_order_relaxed);
        while (h < t) {
            if (r.head.compare_exchange_strong(h, h + 1, std::memory_order_relaxed, std::memory_order_relaxed))
                return h;
        }
        std::this_thread::yield();
    }
}

// a wave: claim, read the slot's request, write its result, publish done
void wave(Ring& r) {
    for (;;) {
        const int64_t item = claim(r);
        if (item < 0) return;
        const int slot = (int)(item % kRing);
        float s = 0.f;
        for (int j = 0; j < kRow; j++) s += r.q[slot][j];
        s += (float)r.ks[slot] + 2.f * (float)r.efs[slot];
        r.out[slot] = s;
        r.done[slot].store(item + 1, std::memory_order_release);
    }
}

}  // namespace

int main() {
    static Ring r;
    for (auto& d : r.done) d.store(0, std::memory_order_relaxed);
    std::vector<std::thread> waves;
    for (int w = 0; w < kWaves; w++) waves.emplace_back(wave, std::ref(r));
    int64_t wrong = 0, retired = 0;
    auto retire = [&](int64_t req) {
        const int slot = (int)(req % kRing);
        while (r.done[slot].load(std::memory_order_acquire) != req + 1) std::this_thread::yield();
        if (r.out[slot] != expected(req)) wrong++;
        retired++;
    };
    // dispatcher_loop_resident: a slot is rewritten once its previous request retired
    for (int64_t req = 0; req < kRequests; req++) {
        if (req >= kRing) retire(req - kRing);
        const int slot = (int)(req % kRing);
        for (int j = 0; j < kRow; j++) r.q[slot][j] = (float)(req % 1000) + (float)j;
        r.ks[slot] = (int32_t)(req % 7);
        r.efs[slot] = (int32_t)(req % 5);
        r.tail.store(req + 1, std::memory_order_release);
    }
    for (int64_t req = kRequests - kRing; req < kRequests; req++) retire(req);
    r.stop.store(1, std::memory_order_release);
    for (auto& t : waves) t.join();
    std::printf("requests %lld retired %lld wrong %lld\n", (long long)kRequests, (long long)retired, (long long)wrong);
    return wrong ? 3 : 0;
}
