// Continuous-batching serving loop (midint_server_*).
//
// Requests (one add or mul each, operands 1..4096 limbs) are queued on the
// host at submit time. poll() drives the scheduler without blocking:
//
//  1. retire: micro-batches whose completion event has fired hand their
//     results to their requests (latency = completion - submission) and give
//     their device region back to the pool;
//  2. admit: while a stream slot is free and requests are queued, a new
//     micro-batch takes queued requests in arrival order as long as its
//     device region fits the free part of the HBM pool (the admission
//     budget) and the staging limits; a request that does not fit waits
//     (deferred, never dropped);
//  3. launch: the batch's operands go to HBM with one transfer kernel from
//     its pinned staging buffer (the operands of requests above 1024 limbs
//     packed there and spread into their zero-padded class slots on the
//     way), its kernels run on the slot's stream, results come back with
//     one transfer kernel (the big requests' results gathered packed from
//     their padded outputs), and an event marks completion.
//
// The transfers are kernels reading and writing the mapped staging memory,
// not copy-engine transfers: hipMemcpyAsync hands copies above a few tens of
// KB to a DMA engine, and the first use of an engine queue blocks the
// calling thread for milliseconds (measured: 6-7 ms stalls of poll() at
// random points of a run), while a kernel on the slot's own stream costs a
// launch and keeps the whole batch on one hardware queue.
//
// Several micro-batches are in flight on different streams, so work of
// requests admitted later overlaps on the device with work admitted earlier.
// Kernels per micro-batch: for ops whose operands are <= 1024 limbs, the
// ragged plans' kernels (ragged.wave = 1, the default: the additions and
// the products whose shorter operand has at most 128 limbs one op per wave,
// the other products on matrix-core kernels with tiles and waves per size
// class; ragged.wave = 0: the matrix-core kernel for multiplications whose
// shorter operand has at least ragged.mfma_min limbs and the ragged
// persistent kernel for the rest); the add kernel for larger additions (padded
// to a 2048- or 4096-limb class, carry-out kept); the matrix-core kernel for
// larger products (one workgroup each up to 2048 limbs, or four full
// 2048-limb quarter products plus the combine kernel above 2048 limbs; the
// NTT kernel when no launch configuration fits). Every path is exact, so a
// request's result does not depend on what it was batched with.
#include <hip/hip_runtime.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <deque>
#include <map>
#include <new>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "midint.h"
#include "midint_internal.h"
#include "mul_core.hpp"

using namespace midint;

namespace {

constexpr uint32_t kServeMaxLimbs = 4096;
constexpr uint32_t kNttM = 2048;

double now_s() {
    timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Host copies of a micro-batch (operands in at submit, staging fills, result
// copies out) run on several threads once a call moves enough bytes to pay
// for starting them (measured: one thread copies about 13 GB/s, and the
// host side of a 3000-request burst took 12.6 ms against 1.6 ms of device
// work per 1000 requests).
constexpr size_t kParallelCopyBytes = 1u << 20;
constexpr unsigned kMaxCopyThreads = 8;

// Runs body(i) for i in [0, n): items split into contiguous ranges of about
// equal weight(i) (bytes), one range per thread; on the calling thread
// alone below kParallelCopyBytes in total, or where a thread cannot start.
template <class W, class F>
void parallel_items(size_t n, W weight, F body) {
    size_t total = 0;
    for (size_t i = 0; i < n; i++) total += weight(i);
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned nt = (unsigned)std::min<size_t>({(size_t)std::min(kMaxCopyThreads, hw), n,
                                                    total / (kParallelCopyBytes / 4) + 1});
    if (nt <= 1 || total < kParallelCopyBytes) {
        for (size_t i = 0; i < n; i++) body(i);
        return;
    }
    std::vector<size_t> cut(nt + 1, n);
    cut[0] = 0;
    size_t acc = 0;
    unsigned c = 1;
    for (size_t i = 0; i < n && c < nt; i++) {
        acc += weight(i);
        while (c < nt && acc * nt >= total * c) cut[c++] = i + 1;
    }
    auto run = [&](unsigned k) {
        for (size_t i = cut[k]; i < cut[k + 1]; i++) body(i);
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    unsigned k = 1;
    try {
        for (; k < nt; k++) th.emplace_back(run, k);
    } catch (...) {
        for (; k < nt; k++) run(k);  // no more threads: the rest here
    }
    run(0);
    for (std::thread& t : th) t.join();
}

struct Request {
    uint32_t kind, na, nb;
    std::vector<uint64_t> a, b;   // host copies until the request is staged
    std::vector<uint64_t> result; // filled at completion
    double t_submit = 0, t_done = 0;
    bool done = false;
};

// Where a request's result sits in the batch's output staging (the packed
// results that come back).
struct ResultRef {
    uint64_t id;
    size_t out_off;  // limbs from the start of the output area
    uint32_t len;    // limbs of the result
};

enum class Route { Ragged, AddClass2048, AddClass4096, Ntt2048, Ntt4096 };

Route route_of(uint32_t kind, uint32_t na, uint32_t nb) {
    const uint32_t mx = std::max(na, nb);
    if (mx <= kRgMaxLimbs) return Route::Ragged;
    if (kind == MIDINT_OP_ADD) return mx <= 2048 ? Route::AddClass2048 : Route::AddClass4096;
    return mx <= kNttM ? Route::Ntt2048 : Route::Ntt4096;
}

// Device bytes a request needs inside a batch region (upper bound including
// its share of the metadata): ragged ops their operands and result; the
// others their zero-padded class slots, the class-size outputs and their
// slot / result entries (their packed operands and results only pass
// through the staging buffer, which is never larger than the region).
size_t request_bytes(uint32_t kind, uint32_t na, uint32_t nb) {
    constexpr size_t kBigMeta = 8 * sizeof(ServeSlot) + sizeof(ServeGather) + sizeof(MfRgItem);
    switch (route_of(kind, na, nb)) {
        case Route::Ragged:
            return 8ull * (na + nb + (kind == MIDINT_OP_MUL ? na + nb : std::max(na, nb) + 1)) + 2 * sizeof(RgOp);
        case Route::AddClass2048: return 8ull * 3 * 2048 + 1 + kBigMeta;
        case Route::AddClass4096: return 8ull * 3 * 4096 + 1 + kBigMeta;
        case Route::Ntt2048: return 8ull * (2 * kNttM + 2 * kNttM) + kBigMeta;
        case Route::Ntt4096: return 4 * 8ull * (2 * kNttM + 2 * kNttM) + 4 * kBigMeta;
    }
    return 0;
}

struct Pool {  // first-fit allocator over one device allocation
    uint8_t* base = nullptr;
    size_t total = 0, used = 0;
    std::map<size_t, size_t> free_;  // offset -> size
    void init(uint8_t* b, size_t n) {
        base = b;
        total = n;
        free_.clear();
        free_[0] = n;
        used = 0;
    }
    bool alloc(size_t n, size_t* off) {
        n = align_up(n, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n;
                const size_t noff = it->first + n;
                free_.erase(it);
                if (rest) free_[noff] = rest;
                used += n;
                return true;
            }
        }
        return false;
    }
    size_t largest() const {
        size_t m = 0;
        for (auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }
    void release(size_t off, size_t n) {
        n = align_up(n, 256);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    uint8_t* pinned = nullptr;  // staging: inputs then outputs
    uint8_t* pinned_dev = nullptr;  // the staging buffer's device address
    size_t pinned_bytes = 0;
    bool busy = false;
    // in-flight batch
    size_t region_off = 0, region_bytes = 0, out_off_bytes = 0, out_bytes = 0;
    std::vector<ResultRef> refs;
    std::vector<uint64_t> ids;
};

}  // namespace

struct midint_server {
    midint_ctx* ctx = nullptr;
    int device = 0;
    uint64_t max_batch_bytes = 0;
    uint32_t max_batch_ops = 0;
    Pool pool;
    uint8_t* d_pool = nullptr;
    std::vector<Slot> slots;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Request> reqs;
    uint64_t next_id = 1;
    midint_server_counters st = {};
    uint32_t cus = 0;
    int ragged_per_cu = 1;
    size_t ragged_lds = 0;
};

namespace {

void destroy_server(midint_server* s) {
    if (!s) return;
    for (Slot& sl : s->slots) {
        if (sl.stream) {
            (void)hipStreamSynchronize(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
    }
    if (s->d_pool) (void)hipFree(s->d_pool);
    delete s;
}

// (Re)allocates a slot's staging buffer: pinned, mapped into the device
// address space and coherent, so the copy kernels read and write it directly
// and results written by the device are visible to the host once the batch's
// event has completed. The host touches every page once here: the first
// host write to a fresh pinned page is a fault (measured: a 33 MB memset
// into an untouched buffer 12 ms, into a touched one about 1 ms), which
// would otherwise land on the first large micro-batch of each slot.
hipError_t alloc_staging(Slot& sl, size_t bytes) {
    if (sl.pinned) (void)hipHostFree(sl.pinned);
    sl.pinned = nullptr;
    sl.pinned_dev = nullptr;
    sl.pinned_bytes = 0;
    hipError_t e = hipHostMalloc((void**)&sl.pinned, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&sl.pinned_dev, sl.pinned, 0);
    if (e == hipSuccess) memset(sl.pinned, 0, bytes);
    if (e != hipSuccess) {
        if (sl.pinned) (void)hipHostFree(sl.pinned);
        sl.pinned = nullptr;
        sl.pinned_dev = nullptr;
        return e;
    }
    sl.pinned_bytes = bytes;
    return hipSuccess;
}

// Builds and launches one micro-batch on slot `si` from the queue head.
// Returns MIDINT_EBUSY when the head request does not fit the free pool.
midint_status launch_batch(midint_server* s, size_t si) {
    Slot& sl = s->slots[si];
    // ---- choose requests in arrival order
    std::vector<uint64_t> take;
    size_t est = 64 * 1024;  // metadata slack
    const size_t free_max = s->pool.largest();
    while (!s->queue.empty() && take.size() < s->max_batch_ops) {
        const Request& r = s->reqs[s->queue.front()];
        const size_t need = request_bytes(r.kind, r.na, r.nb);
        if (est + need > s->max_batch_bytes || est + need > free_max) break;
        est += need;
        take.push_back(s->queue.front());
        s->queue.pop_front();
    }
    if (take.empty()) {
        s->st.deferred_admissions++;
        return MIDINT_EBUSY;
    }

    // ---- lay out the batch: [inputs | outputs] in limbs
    std::vector<midint_op> rops;        // ragged ops (offsets filled below)
    std::vector<uint64_t> rids;
    std::vector<uint64_t> add2k, add4k, ntt2k, ntt4k;
    for (uint64_t id : take) {
        Request& r = s->reqs[id];
        switch (route_of(r.kind, r.na, r.nb)) {
            case Route::Ragged: rids.push_back(id); break;
            case Route::AddClass2048: add2k.push_back(id); break;
            case Route::AddClass4096: add4k.push_back(id); break;
            case Route::Ntt2048: ntt2k.push_back(id); break;
            case Route::Ntt4096: ntt4k.push_back(id); break;
        }
    }
    // inputs
    size_t in = 0;
    const size_t rg_a = in;
    for (uint64_t id : rids) {
        Request& r = s->reqs[id];
        midint_op o = {r.kind, r.na, r.nb, 0, in - rg_a, 0, 0};
        rops.push_back(o);
        in += r.na;
    }
    const size_t rg_b = in;
    for (size_t k = 0; k < rids.size(); k++) {
        rops[k].off_b = in - rg_b;
        in += s->reqs[rids[k]].nb;
    }
    // the requests above the ragged limit (in the order add2k, add4k,
    // ntt2k, ntt4k) cross the bus packed; they stay in the staging buffer
    // after the inputs that go to the device (offsets below)
    std::vector<uint64_t> big;
    big.reserve(add2k.size() + add4k.size() + ntt2k.size() + ntt4k.size());
    big.insert(big.end(), add2k.begin(), add2k.end());
    big.insert(big.end(), add4k.begin(), add4k.end());
    big.insert(big.end(), ntt2k.begin(), ntt2k.end());
    big.insert(big.end(), ntt4k.begin(), ntt4k.end());
    std::vector<size_t> big_a(big.size()), big_b(big.size()), big_r(big.size());
    // their zero-padded class slots, on the device only (limbs from the
    // start of the padded area), filled from the packed operands by the
    // inbound transfer
    size_t pad = 0;
    const size_t a2k = pad;
    pad += (size_t)add2k.size() * 2048 * 2;
    const size_t a4k = pad;
    pad += (size_t)add4k.size() * 4096 * 2;
    const size_t nttx = pad;
    const size_t n_ntt = ntt2k.size() + 4 * ntt4k.size();
    pad += n_ntt * kNttM * 2;
    // ragged metadata (as limbs)
    std::vector<RgItem> items;
    std::vector<RgOp> rgops;
    uint32_t conv_cap = 0;
    // outputs
    size_t out = 0;
    const size_t rg_r = out;
    for (size_t k = 0; k < rids.size(); k++) {
        Request& r = s->reqs[rids[k]];
        rops[k].off_r = out - rg_r;
        out += r.kind == MIDINT_OP_MUL ? r.na + r.nb : std::max(r.na, r.nb) + 1;
    }
    const size_t rg_out = out;
    const size_t a2k_r = out;
    out += add2k.size() * 2048;
    const size_t a4k_r = out;
    out += add4k.size() * 4096;
    const size_t ntt_r = out;
    out += n_ntt * 2 * kNttM;
    const size_t carry_bytes = align_up(add2k.size() + add4k.size(), 8);
    const size_t items_off = in;  // limbs
    // ragged ops, as the ragged plans run them: with ragged.wave = 1 (the
    // default) the additions and the products whose shorter operand has at
    // most rgw_max_short() limbs one op per wave, the other products on
    // matrix-core kernels (tiles and waves per size class: ragged.m16@C,
    // ragged.m16_waves@C, ragged.mfma_waves@C; one launch per tile shape);
    // with ragged.wave = 0 the products whose shorter operand has at least
    // ragged.mfma_min limbs on one matrix-core launch and the rest on the
    // persistent kernel
    const bool by_wave = tune_get("ragged.wave", 1) != 0;
    const int64_t mf_min = by_wave ? (int64_t)rgw_max_short() + 1 : tune_get("ragged.mfma_min", 32);
    std::vector<midint_op> pack_ops;
    std::vector<size_t> wv_add_idx, wv_mul_idx;
    struct MfGroup {
        std::vector<size_t> idx;
        uint32_t na = 0, nb = 0, waves = kMfRgWaves;
        MfRgConfig cfg = {};
        size_t meta_off = 0;
    };
    MfGroup mfg[2];  // by_wave: [0] 32 x 32 x 32 tiles, [1] 16 x 16 x 64; else [0] only
    for (size_t k = 0; k < rops.size(); k++) {
        const midint_op& o = rops[k];
        const uint32_t nl = (uint32_t)std::max(o.n_a, o.n_b), ns = (uint32_t)std::min(o.n_a, o.n_b);
        if (by_wave && o.kind == MIDINT_OP_ADD) {
            wv_add_idx.push_back(k);
        } else if (by_wave && o.kind == MIDINT_OP_MUL && ns <= rgw_max_short()) {
            wv_mul_idx.push_back(k);
        } else if (o.kind == MIDINT_OP_MUL && mf_min > 0 && ns >= (uint64_t)mf_min) {
            int g = 0;
            if (by_wave) {
                const int64_t mode = tune_get_class("ragged.m16", nl, 0);
                const int64_t ratio = std::max<int64_t>(1, tune_get_class("ragged.m16_ratio", nl, 2));
                g = (mode == 1 || (mode == 2 && (uint64_t)ns * ratio >= nl)) ? 1 : 0;
                // waves: the largest class's setting
                if (nl > mfg[g].na)
                    mfg[g].waves = (uint32_t)tune_get_class(g ? "ragged.m16_waves" : "ragged.mfma_waves", nl,
                                                            kMfRgWaves);
            }
            mfg[g].idx.push_back(k);
            mfg[g].na = std::max(mfg[g].na, nl);
            mfg[g].nb = std::max(mfg[g].nb, ns);
        } else {
            pack_ops.push_back(o);
        }
    }
    for (int g = 0; g < 2; g++) {
        MfGroup& G = mfg[g];
        if (G.idx.empty()) continue;
        bool ok;
        if (!by_wave) {
            ok = mfma_rg_config(G.na, G.nb, G.na + G.nb, &G.cfg);
        } else {
            auto cfg = [&](uint32_t w) {
                return g ? mfma16_rg_config(G.na, G.nb, G.na + G.nb, &G.cfg, w)
                         : mfma32_rg_config(G.na, G.nb, G.na + G.nb, &G.cfg, w);
            };
            ok = (G.waves == 1 || G.waves == 2 || G.waves == 4 || G.waves == 8) && cfg(G.waves);
            if (!ok) ok = cfg(kMfRgWaves);
        }
        if (!ok) {  // no fitting configuration: the persistent kernel takes them
            for (size_t k : G.idx) pack_ops.push_back(rops[k]);
            G.idx.clear();
            continue;
        }
        std::stable_sort(G.idx.begin(), G.idx.end(), [&](size_t x, size_t y) {
            return (uint64_t)rops[x].n_a * rops[x].n_b > (uint64_t)rops[y].n_a * rops[y].n_b;
        });
    }
    // one-op-per-wave items, the costliest first (as in the ragged plans)
    auto wv_cost = [&](size_t k) {
        const uint64_t nl = std::max(rops[k].n_a, rops[k].n_b), ns = std::min(rops[k].n_a, rops[k].n_b);
        return ((nl + ns + 127) / 128) * ((8 * ns + 15) / 64 + 1);
    };
    std::stable_sort(wv_mul_idx.begin(), wv_mul_idx.end(), [&](size_t x, size_t y) { return wv_cost(x) > wv_cost(y); });
    std::stable_sort(wv_add_idx.begin(), wv_add_idx.end(), [&](size_t x, size_t y) {
        return std::max(rops[x].n_a, rops[x].n_b) > std::max(rops[y].n_a, rops[y].n_b);
    });
    midint_status ps = ragged_pack(pack_ops.data(), pack_ops.size(), 0, 0, 0, items, rgops, &conv_cap);
    if (ps != MIDINT_OK) return ps;
    // products above the ragged limit: matrix-core items over the padded
    // 2048-limb operand slots (one per 2048-class product, four full-size
    // quarter products per split product) on the tiles and waves the
    // uniform 2048-limb batches use (mul.m16@2048, default 16 x 16 x 64;
    // mul.m16_waves@2048, default 8 waves), with the NTT kernel as the
    // fallback when no configuration fits
    MfRgConfig bigmf = {};
    bool big_mf = false;
    if (n_ntt > 0) {
        if (tune_get_class("mul.m16", kNttM, 1) != 0) {
            const int64_t w = tune_get_class("mul.m16_waves", kNttM, 0);
            big_mf = mfma16_rg_config(kNttM, kNttM, 2 * kNttM, &bigmf, w > 0 && w <= 16 ? (uint32_t)w : 8u);
        }
        if (!big_mf) big_mf = mfma_rg_config(kNttM, kNttM, 2 * kNttM, &bigmf);
    }
    // metadata: [RgItem | RgOp | MfRgItem groups (16-byte aligned) | big MfRgItem | wave additions | wave
    // products | counter limb]
    size_t meta_at = align_up(items.size() * sizeof(RgItem) + rgops.size() * sizeof(RgOp), 16);
    for (MfGroup& G : mfg) {
        G.meta_off = meta_at;
        meta_at += G.idx.size() * sizeof(MfRgItem);
    }
    const size_t big_meta_off = meta_at;
    const size_t wv_add_off = big_meta_off + (big_mf ? n_ntt : 0) * sizeof(MfRgItem);
    const size_t wv_mul_off = wv_add_off + wv_add_idx.size() * sizeof(MfRgItem);
    // the big requests' slot and result tables
    const size_t n_slots = 2 * (add2k.size() + add4k.size() + n_ntt);
    const size_t slot_meta_off = align_up(wv_mul_off + wv_mul_idx.size() * sizeof(MfRgItem), 16);
    const size_t gather_meta_off = align_up(slot_meta_off + n_slots * sizeof(ServeSlot), 16);
    const size_t meta_limbs = align_up(gather_meta_off + big.size() * sizeof(ServeGather), 8) / 8 + 1 /* counter */;
    in += meta_limbs;
    in = align_up(in, 32);
    const size_t in_copy = in;  // limbs that go to the device
    // staging only: the big requests' packed operands (their a's, then their
    // b's), then the results that come back: the ragged results as they lie
    // on the device, then the big requests' results, packed
    for (size_t i = 0; i < big.size(); i++) {
        big_a[i] = in;
        in += s->reqs[big[i]].na;
    }
    for (size_t i = 0; i < big.size(); i++) {
        big_b[i] = in;
        in += s->reqs[big[i]].nb;
    }
    const size_t host_in = align_up(in, 32);
    size_t out_copy = rg_out;
    for (size_t i = 0; i < big.size(); i++) {
        const Request& r = s->reqs[big[i]];
        big_r[i] = out_copy;
        out_copy += r.kind == MIDINT_OP_MUL ? r.na + r.nb : std::max(r.na, r.nb) + 1;
    }
    // device region: [inputs | padded slots | outputs | carry bytes]
    const size_t in_bytes = in_copy * 8;
    const size_t pad_bytes = align_up(pad, 32) * 8;
    const size_t out_bytes = out * 8 + carry_bytes;
    const size_t region = in_bytes + pad_bytes + out_bytes;
    const size_t staging = host_in * 8 + out_copy * 8;
    if (staging > sl.pinned_bytes) {
        // rare: staging too small for this batch; grow it
        if (alloc_staging(sl, staging) != hipSuccess) {
            for (auto it = take.rbegin(); it != take.rend(); ++it) s->queue.push_front(*it);
            return MIDINT_ENOMEM;
        }
    }
    size_t off = 0;
    if (!s->pool.alloc(region, &off)) {
        for (auto it = take.rbegin(); it != take.rend(); ++it) s->queue.push_front(*it);
        s->st.deferred_admissions++;
        return MIDINT_EBUSY;
    }

    // ---- fill the staging inputs: the packed operands, in parallel; the
    // metadata area after them starts zeroed
    uint64_t* h = reinterpret_cast<uint64_t*>(sl.pinned);
    struct Fill {
        uint64_t* dst;
        const uint64_t* src;
        size_t n;  // limbs
    };
    std::vector<Fill> fills;
    fills.reserve(2 * (rids.size() + big.size()));
    for (size_t k = 0; k < rids.size(); k++) {
        const Request& r = s->reqs[rids[k]];
        fills.push_back({h + rg_a + rops[k].off_a, r.a.data(), r.na});
        fills.push_back({h + rg_b + rops[k].off_b, r.b.data(), r.nb});
    }
    for (size_t i = 0; i < big.size(); i++) {
        const Request& r = s->reqs[big[i]];
        fills.push_back({h + big_a[i], r.a.data(), r.na});
        fills.push_back({h + big_b[i], r.b.data(), r.nb});
    }
    parallel_items(
        fills.size(), [&](size_t i) { return 8 * fills[i].n; },
        [&](size_t i) { memcpy(fills[i].dst, fills[i].src, 8 * fills[i].n); });
    memset(h + items_off, 0, (in_copy - items_off) * 8);
    // ---- device pointers
    uint8_t* dreg = s->pool.base + off;
    uint64_t* d_in = reinterpret_cast<uint64_t*>(dreg);
    uint64_t* d_pad = reinterpret_cast<uint64_t*>(dreg + in_bytes);
    uint64_t* d_out = reinterpret_cast<uint64_t*>(dreg + in_bytes + pad_bytes);
    uint8_t* d_carry = dreg + in_bytes + pad_bytes + out * 8;

    uint8_t* meta = reinterpret_cast<uint8_t*>(h + items_off);
    memcpy(meta, items.data(), items.size() * sizeof(RgItem));
    memcpy(meta + items.size() * sizeof(RgItem), rgops.data(), rgops.size() * sizeof(RgOp));
    {
        // matrix-core and one-per-wave product items: device addresses inside
        // this batch's region, the longer operand first
        auto product_item = [&](size_t k) {
            const midint_op& o = rops[k];
            const bool sw = o.n_a < o.n_b;
            const uint64_t* x = sw ? d_in + rg_b + o.off_b : d_in + rg_a + o.off_a;
            const uint64_t* y = sw ? d_in + rg_a + o.off_a : d_in + rg_b + o.off_b;
            return mfma_rg_item(x, y, d_out + rg_r + o.off_r, (uint32_t)(sw ? o.n_b : o.n_a),
                                (uint32_t)(sw ? o.n_a : o.n_b));
        };
        for (const MfGroup& G : mfg) {
            MfRgItem* mfi = reinterpret_cast<MfRgItem*>(meta + G.meta_off);
            for (size_t q = 0; q < G.idx.size(); q++) mfi[q] = product_item(G.idx[q]);
        }
        MfRgItem* wvm = reinterpret_cast<MfRgItem*>(meta + wv_mul_off);
        for (size_t q = 0; q < wv_mul_idx.size(); q++) wvm[q] = product_item(wv_mul_idx[q]);
        // additions in the op's order of operands
        MfRgItem* wva = reinterpret_cast<MfRgItem*>(meta + wv_add_off);
        for (size_t q = 0; q < wv_add_idx.size(); q++) {
            const midint_op& o = rops[wv_add_idx[q]];
            MfRgItem it = {};
            it.a = d_in + rg_a + o.off_a;
            it.b = d_in + rg_b + o.off_b;
            it.r = d_out + rg_r + o.off_r;
            it.n_a = (uint32_t)o.n_a;
            it.n_b = (uint32_t)o.n_b;
            wva[q] = it;
        }
        if (big_mf) {
            // instance inst: operands at X / Y + inst kNttM (zero-padded),
            // product at ntt_r + inst 2 kNttM
            MfRgItem* bi = reinterpret_cast<MfRgItem*>(meta + big_meta_off);
            const uint64_t* X = d_pad + nttx;
            const uint64_t* Y = d_pad + nttx + n_ntt * kNttM;
            uint64_t* Z = d_out + ntt_r;
            size_t inst = 0;
            for (uint64_t id : ntt2k) {
                const Request& r = s->reqs[id];
                const bool sw = r.na < r.nb;
                const uint64_t* x = (sw ? Y : X) + inst * kNttM;
                const uint64_t* y = (sw ? X : Y) + inst * kNttM;
                bi[inst] = mfma_rg_item(x, y, Z + inst * 2 * kNttM, sw ? r.nb : r.na, sw ? r.na : r.nb);
                inst++;
            }
            for (; inst < n_ntt; inst++)  // the split's quarters: all 2 kNttM limbs feed the combine
                bi[inst] = mfma_rg_item(X + inst * kNttM, Y + inst * kNttM, Z + inst * 2 * kNttM, kNttM, kNttM);
        }
        // the big requests' padded slots, filled from their packed operands
        // in the staging buffer (its device-mapped address)
        const uint64_t* hdev = reinterpret_cast<const uint64_t*>(sl.pinned_dev);
        ServeSlot* sv = reinterpret_cast<ServeSlot*>(meta + slot_meta_off);
        size_t q = 0;
        auto slot = [&](size_t i, bool b_op, uint32_t skip, uint32_t n, uint64_t* dst, uint32_t size) {
            sv[q++] = ServeSlot{hdev + (b_op ? big_b[i] : big_a[i]) + skip, dst, n, size};
        };
        size_t i = 0;
        for (size_t k = 0; k < add2k.size(); k++, i++) {
            const Request& r = s->reqs[big[i]];
            slot(i, false, 0, r.na, d_pad + a2k + k * 2048, 2048);
            slot(i, true, 0, r.nb, d_pad + a2k + (add2k.size() + k) * 2048, 2048);
        }
        for (size_t k = 0; k < add4k.size(); k++, i++) {
            const Request& r = s->reqs[big[i]];
            slot(i, false, 0, r.na, d_pad + a4k + k * 4096, 4096);
            slot(i, true, 0, r.nb, d_pad + a4k + (add4k.size() + k) * 4096, 4096);
        }
        uint64_t* PX = d_pad + nttx;
        uint64_t* PY = d_pad + nttx + n_ntt * kNttM;
        size_t inst = 0;
        for (size_t k = 0; k < ntt2k.size(); k++, i++, inst++) {
            const Request& r = s->reqs[big[i]];
            slot(i, false, 0, r.na, PX + inst * kNttM, kNttM);
            slot(i, true, 0, r.nb, PY + inst * kNttM, kNttM);
        }
        for (size_t k = 0; k < ntt4k.size(); k++, i++) {
            const Request& r = s->reqs[big[i]];
            // halves: a = a1 X + a0, b = b1 X + b0; instances a0b0, a1b1, a0b1, a1b0
            const uint32_t a0 = std::min(r.na, kNttM), a1 = r.na > kNttM ? r.na - kNttM : 0;
            const uint32_t b0 = std::min(r.nb, kNttM), b1 = r.nb > kNttM ? r.nb - kNttM : 0;
            const uint32_t xo[4] = {0, a0, 0, a0}, xl[4] = {a0, a1, a0, a1};
            const uint32_t yo[4] = {0, b0, b0, 0}, yl[4] = {b0, b1, b1, b0};
            for (int u = 0; u < 4; u++, inst++) {
                slot(i, false, xo[u], xl[u], PX + inst * kNttM, kNttM);
                slot(i, true, yo[u], yl[u], PY + inst * kNttM, kNttM);
            }
        }
        // and their results, into their packed places in the staging buffer
        uint64_t* hout = reinterpret_cast<uint64_t*>(sl.pinned_dev) + host_in;
        ServeGather* gv = reinterpret_cast<ServeGather*>(meta + gather_meta_off);
        i = 0;
        for (size_t k = 0; k < add2k.size() + add4k.size(); k++, i++) {
            const Request& r = s->reqs[big[i]];
            const bool c4 = k >= add2k.size();
            const uint32_t C = c4 ? 4096 : 2048;
            const size_t kk = c4 ? k - add2k.size() : k;
            const uint32_t len = std::max(r.na, r.nb) + 1;
            // the full-width sum at the class size: its top limb is the kernel's carry-out
            gv[i] = ServeGather{d_out + (c4 ? a4k_r : a2k_r) + kk * C, hout + big_r[i], len - 1 == C ? d_carry + k : nullptr,
                                len, 0};
        }
        inst = 0;
        for (size_t k = 0; k < ntt2k.size() + ntt4k.size(); k++, i++) {
            const Request& r = s->reqs[big[i]];
            gv[i] = ServeGather{d_out + ntt_r + inst * 2 * kNttM, hout + big_r[i], nullptr, r.na + r.nb, 0};
            inst += k < ntt2k.size() ? 1 : 4;
        }
    }
    RgItem* d_items = reinterpret_cast<RgItem*>(d_in + items_off);
    RgOp* d_rgops = reinterpret_cast<RgOp*>(reinterpret_cast<uint8_t*>(d_items) + items.size() * sizeof(RgItem));
    uint32_t* d_counter = reinterpret_cast<uint32_t*>(d_in + items_off + meta_limbs - 1);

    // the ragged kernel's work counter is the last metadata limb, zeroed by
    // the memset of the staging above, so the input copy clears it
    hipStream_t st = sl.stream;
    // (the slot and result tables are read from the staging buffer)
    const uint8_t* h_meta = sl.pinned_dev + items_off * 8;
    hipError_t e = launch_stage_in(sl.pinned_dev, d_in, in_bytes,
                                   reinterpret_cast<const ServeSlot*>(h_meta + slot_meta_off), (uint32_t)n_slots, st);
    uint8_t* d_meta = reinterpret_cast<uint8_t*>(d_in + items_off);
    if (e == hipSuccess && !items.empty()) {
        const uint32_t n_items = (uint32_t)items.size();
        const uint64_t* dA = d_in + rg_a;
        const uint64_t* dB = d_in + rg_b;
        uint64_t* dR = d_out + rg_r;
        uint32_t cap = conv_cap;
        void* args[] = {(void*)&d_items, (void*)&d_rgops, (void*)&n_items, (void*)&d_counter,
                        (void*)&dA,      (void*)&dB,      (void*)&dR,      (void*)&cap};
        const size_t lds = ragged_lds_size(256, cap);
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(n_items, s->cus * s->ragged_per_cu));
        if (e == hipSuccess)
            e = hipLaunchKernel(ragged_kernel_ptr(256), dim3(grid), dim3(256), args, lds, st);
    }
    for (const MfGroup& G : mfg) {
        if (e != hipSuccess || G.idx.empty()) continue;
        MfRgItem* d_mf = reinterpret_cast<MfRgItem*>(d_meta + G.meta_off);
        uint32_t pow8 = G.cfg.pow8;
        void* args[] = {(void*)&d_mf, (void*)&pow8};
        if (G.cfg.lds > 64 * 1024)
            e = hipFuncSetAttribute(G.cfg.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)G.cfg.lds);
        if (e == hipSuccess)
            e = hipLaunchKernel(G.cfg.func, dim3((uint32_t)G.idx.size()), dim3(G.cfg.threads), args, G.cfg.lds, st);
    }
    for (int k = 0; k < 2 && e == hipSuccess; k++) {  // one op per wave: products, additions
        const bool mul = k == 0;
        uint32_t n_wv = (uint32_t)(mul ? wv_mul_idx.size() : wv_add_idx.size());
        if (n_wv == 0) continue;
        MfRgItem* d_wv = reinterpret_cast<MfRgItem*>(d_meta + (mul ? wv_mul_off : wv_add_off));
        uint32_t pow8 = 256;
        void* args[] = {(void*)&d_wv, (void*)&n_wv, (void*)&pow8};
        e = hipLaunchKernel(mul ? rgw_mul_kernel_ptr(0) : rgw_add_kernel_ptr(), dim3((n_wv + rgw_waves() - 1) / rgw_waves()),
                            dim3(64 * rgw_waves()), args, mul ? rgw_lds_bytes(0) : 0, st);
    }
    if (e == hipSuccess && !add2k.empty())
        e = launch_add((uint32_t)add2k.size(), 2048, d_pad + a2k, d_pad + a2k + add2k.size() * 2048, d_out + a2k_r,
                       d_carry, false, st);
    if (e == hipSuccess && !add4k.empty())
        e = launch_add((uint32_t)add4k.size(), 4096, d_pad + a4k, d_pad + a4k + add4k.size() * 4096, d_out + a4k_r,
                       d_carry + add2k.size(), false, st);
    if (e == hipSuccess && n_ntt > 0 && big_mf) {
        MfRgItem* d_big = reinterpret_cast<MfRgItem*>(d_meta + big_meta_off);
        uint32_t pow8 = bigmf.pow8;
        void* args[] = {(void*)&d_big, (void*)&pow8};
        if (bigmf.lds > 64 * 1024)
            e = hipFuncSetAttribute(bigmf.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bigmf.lds);
        if (e == hipSuccess)
            e = hipLaunchKernel(bigmf.func, dim3((uint32_t)n_ntt), dim3(bigmf.threads), args, bigmf.lds, st);
    } else if (e == hipSuccess && n_ntt > 0) {
        NttTables t;
        midint_status ts = ctx_ntt_tables(s->ctx, ntt_lg_for_limbs(kNttM), &t);
        if (ts != MIDINT_OK) e = hipErrorUnknown;
        if (e == hipSuccess)
            e = launch_mul_ntt((uint32_t)n_ntt, kNttM, d_pad + nttx, d_pad + nttx + n_ntt * kNttM, d_out + ntt_r, t,
                               st);
    }
    if (e == hipSuccess && !ntt4k.empty())
        e = launch_bigmul_combine((uint32_t)ntt4k.size(), d_out + ntt_r + ntt2k.size() * 2 * kNttM, st);
    if (e == hipSuccess)
        e = launch_stage_out(d_out, sl.pinned_dev + host_in * 8, rg_out * 8,
                             reinterpret_cast<const ServeGather*>(h_meta + gather_meta_off), (uint32_t)big.size(), st);
    if (e == hipSuccess) e = hipEventRecord(sl.ev, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        s->pool.release(off, region);
        for (auto it = take.rbegin(); it != take.rend(); ++it) s->queue.push_front(*it);
        return MIDINT_EHIP;
    }

    // ---- bookkeeping for retirement
    sl.busy = true;
    sl.region_off = off;
    sl.region_bytes = region;
    sl.out_off_bytes = host_in * 8;
    sl.out_bytes = out_copy * 8;
    sl.refs.clear();
    sl.ids = take;
    for (size_t k = 0; k < rids.size(); k++) {
        Request& r = s->reqs[rids[k]];
        const uint32_t len = r.kind == MIDINT_OP_MUL ? r.na + r.nb : std::max(r.na, r.nb) + 1;
        sl.refs.push_back({rids[k], rg_r + rops[k].off_r, len});
    }
    for (size_t i = 0; i < big.size(); i++) {
        Request& r = s->reqs[big[i]];
        const uint32_t len = r.kind == MIDINT_OP_MUL ? r.na + r.nb : std::max(r.na, r.nb) + 1;
        sl.refs.push_back({big[i], big_r[i], len});
    }
    for (uint64_t id : take) {
        Request& r = s->reqs[id];
        r.a.clear();
        r.a.shrink_to_fit();
        r.b.clear();
        r.b.shrink_to_fit();
    }
    s->st.batches_launched++;
    s->st.inflight_batches++;
    s->st.ops_in_batches += take.size();
    return MIDINT_OK;
}

void retire(midint_server* s, Slot& sl, double t, std::vector<uint64_t>& done) {
    const uint64_t* outp = reinterpret_cast<const uint64_t*>(sl.pinned + sl.out_off_bytes);
    std::vector<Request*> req(sl.refs.size());
    for (size_t i = 0; i < sl.refs.size(); i++) req[i] = &s->reqs[sl.refs[i].id];
    // the result copies in parallel (each request's own vector)
    parallel_items(
        sl.refs.size(), [&](size_t i) { return 8 * (size_t)sl.refs[i].len; },
        [&](size_t i) {
            const ResultRef& rr = sl.refs[i];
            req[i]->result.assign(outp + rr.out_off, outp + rr.out_off + rr.len);
        });
    for (size_t i = 0; i < sl.refs.size(); i++) {
        req[i]->done = true;
        req[i]->t_done = t;
        done.push_back(sl.refs[i].id);
    }
    s->pool.release(sl.region_off, sl.region_bytes);
    sl.busy = false;
    s->st.inflight_batches--;
    s->st.completed += sl.refs.size();
}

// Runs one small request of every route on every slot and waits for it.
// The first dispatch on a stream attaches it to a hardware queue and the
// first launch of a kernel loads its code object; both block the host for
// milliseconds, which would otherwise land on the first live requests of a
// slot (measured: ~6.5 ms stalls of the submit loop). Counters are reset
// afterwards, so a fresh server reports only live traffic.
midint_status warm_up(midint_server* s) {
    const uint32_t shapes[][3] = {{MIDINT_OP_MUL, 8, 8},       {MIDINT_OP_MUL, 64, 64},  {MIDINT_OP_MUL, 1024, 1024},
                                  {MIDINT_OP_MUL, 200, 150},   {MIDINT_OP_MUL, 300, 200}, {MIDINT_OP_ADD, 8, 8},
                                  {MIDINT_OP_ADD, 2048, 8},    {MIDINT_OP_ADD, 4096, 8},  {MIDINT_OP_MUL, 2048, 8},
                                  {MIDINT_OP_MUL, 4096, 8}};
    std::vector<uint64_t> ones(kServeMaxLimbs, 1);
    std::vector<uint64_t> done;
    for (size_t si = 0; si < s->slots.size(); si++) {
        for (const auto& q : shapes) {
            uint64_t id = 0;
            if (midint_server_submit(s, q[0], ones.data(), q[1], ones.data(), q[2], &id) != MIDINT_OK)
                continue;  // too large for this server's limits
            const midint_status ls = launch_batch(s, si);
            if (ls == MIDINT_EBUSY) {  // does not fit the pool: skip this shape
                s->queue.clear();
                s->reqs.erase(id);
                continue;
            }
            if (ls != MIDINT_OK) return ls;
            Slot& sl = s->slots[si];
            if (hipEventSynchronize(sl.ev) != hipSuccess) return MIDINT_EHIP;
            done.clear();
            retire(s, sl, now_s(), done);
            for (uint64_t d : done) s->reqs.erase(d);
        }
    }
    s->queue.clear();
    s->reqs.clear();
    s->next_id = 1;
    s->st = {};
    return MIDINT_OK;
}

}  // namespace

extern "C" {

midint_status midint_server_create(midint_ctx* ctx, const midint_server_config* cfg, midint_server** out) {
    if (!ctx || !out) return MIDINT_EINVAL;
    *out = nullptr;
    midint_server* s = new (std::nothrow) midint_server;
    if (!s) return MIDINT_ENOMEM;
    s->ctx = ctx;
    const uint64_t budget = (cfg && cfg->hbm_budget_bytes) ? cfg->hbm_budget_bytes : (1ull << 30);
    s->max_batch_bytes = (cfg && cfg->max_batch_bytes) ? cfg->max_batch_bytes : (64ull << 20);
    s->max_batch_ops = (cfg && cfg->max_batch_ops) ? cfg->max_batch_ops : 8192;
    uint32_t nstreams = (cfg && cfg->streams) ? cfg->streams : 4;
    if (nstreams > 8) nstreams = 8;
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc((void**)&s->d_pool, budget) != hipSuccess) {
        (void)hipGetLastError();
        destroy_server(s);
        return MIDINT_ENOMEM;
    }
    s->pool.init(s->d_pool, budget);
    s->slots.resize(nstreams);
    for (Slot& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess ||
            alloc_staging(sl, s->max_batch_bytes + (4u << 20)) != hipSuccess) {
            destroy_server(s);
            return MIDINT_EHIP;
        }
    }
    int cus = 0;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device);
    s->cus = (uint32_t)std::max(1, cus);
    s->ragged_lds = ragged_lds_size(256, 4 * kRgWindow);
    int per = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, ragged_kernel_ptr(256), 256, s->ragged_lds) ==
        hipSuccess)
        s->ragged_per_cu = std::max(1, per);
    NttTables t;
    if (ctx_ntt_tables(ctx, ntt_lg_for_limbs(kNttM), &t) != MIDINT_OK) {
        destroy_server(s);
        return MIDINT_EHIP;
    }
    const midint_status ws = warm_up(s);
    if (ws != MIDINT_OK) {
        destroy_server(s);
        return ws;
    }
    s->st.pool_bytes_total = budget;
    *out = s;
    return MIDINT_OK;
}

midint_status midint_server_submit(midint_server* s, uint32_t kind, const uint64_t* a, uint32_t n_a,
                                   const uint64_t* b, uint32_t n_b, uint64_t* id) {
    if (!s || !a || !b || !id || n_a == 0 || n_b == 0) return MIDINT_EINVAL;
    if (kind != MIDINT_OP_ADD && kind != MIDINT_OP_MUL) return MIDINT_EINVAL;
    if (n_a > kServeMaxLimbs || n_b > kServeMaxLimbs) return MIDINT_EUNSUPPORTED;
    if (request_bytes(kind, n_a, n_b) + 64 * 1024 > std::min<uint64_t>(s->max_batch_bytes, s->pool.total))
        return MIDINT_EUNSUPPORTED;
    const uint64_t k = s->next_id++;
    Request& r = s->reqs[k];
    r.kind = kind;
    r.na = n_a;
    r.nb = n_b;
    r.a.assign(a, a + n_a);
    r.b.assign(b, b + n_b);
    r.t_submit = now_s();
    s->queue.push_back(k);
    s->st.submitted++;
    *id = k;
    return MIDINT_OK;
}

midint_status midint_server_submit_many(midint_server* s, size_t n, const uint32_t* kinds,
                                        const uint64_t* const* a, const uint32_t* n_a,
                                        const uint64_t* const* b, const uint32_t* n_b, uint64_t* ids) {
    if (!s || (n > 0 && (!kinds || !a || !n_a || !b || !n_b || !ids))) return MIDINT_EINVAL;
    for (size_t i = 0; i < n; i++) {
        if (!a[i] || !b[i] || n_a[i] == 0 || n_b[i] == 0) return MIDINT_EINVAL;
        if (kinds[i] != MIDINT_OP_ADD && kinds[i] != MIDINT_OP_MUL) return MIDINT_EINVAL;
        if (n_a[i] > kServeMaxLimbs || n_b[i] > kServeMaxLimbs) return MIDINT_EUNSUPPORTED;
        if (request_bytes(kinds[i], n_a[i], n_b[i]) + 64 * 1024 > std::min<uint64_t>(s->max_batch_bytes, s->pool.total))
            return MIDINT_EUNSUPPORTED;
    }
    // the requests in order (ids, queue), then their operand copies in
    // parallel (each request's own vectors)
    std::vector<Request*> req(n);
    for (size_t i = 0; i < n; i++) {
        const uint64_t k = s->next_id++;
        Request& r = s->reqs[k];
        r.kind = kinds[i];
        r.na = n_a[i];
        r.nb = n_b[i];
        req[i] = &r;
        ids[i] = k;
    }
    parallel_items(
        n, [&](size_t i) { return 8 * ((size_t)n_a[i] + n_b[i]); },
        [&](size_t i) {
            req[i]->a.assign(a[i], a[i] + n_a[i]);
            req[i]->b.assign(b[i], b[i] + n_b[i]);
        });
    const double t = now_s();
    for (size_t i = 0; i < n; i++) {
        req[i]->t_submit = t;
        s->queue.push_back(ids[i]);
    }
    s->st.submitted += n;
    return MIDINT_OK;
}

midint_status midint_server_poll(midint_server* s, uint64_t* done_ids, size_t cap, size_t* n_done) {
    if (!s) return MIDINT_EINVAL;
    std::vector<uint64_t> done;
    const double t = now_s();
    for (Slot& sl : s->slots) {
        if (!sl.busy) continue;
        hipError_t q = hipEventQuery(sl.ev);
        if (q == hipSuccess) {
            retire(s, sl, t, done);
        } else if (q != hipErrorNotReady) {
            return MIDINT_EHIP;
        }
    }
    for (size_t i = 0; i < s->slots.size() && !s->queue.empty(); i++) {
        if (s->slots[i].busy) continue;
        midint_status ls = launch_batch(s, i);
        if (ls == MIDINT_EBUSY) break;
        if (ls != MIDINT_OK) return ls;
    }
    s->st.queued = s->queue.size();
    s->st.pool_bytes_in_use = s->pool.used;
    const size_t k = std::min(cap, done.size());
    if (done_ids)
        for (size_t i = 0; i < k; i++) done_ids[i] = done[i];
    if (n_done) *n_done = done_ids ? k : done.size();
    // ids beyond `cap` are still reported done by midint_server_result
    return MIDINT_OK;
}

midint_status midint_server_result(midint_server* s, uint64_t id, uint64_t* r, size_t cap_limbs, size_t* n_limbs,
                                   double* latency_s) {
    if (!s) return MIDINT_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return MIDINT_ENOTFOUND;
    Request& q = it->second;
    if (!q.done) return MIDINT_EBUSY;
    if (n_limbs) *n_limbs = q.result.size();
    if (latency_s) *latency_s = q.t_done - q.t_submit;
    if (r) {
        if (cap_limbs < q.result.size()) return MIDINT_EINVAL;
        memcpy(r, q.result.data(), q.result.size() * 8);
    }
    s->reqs.erase(it);
    return MIDINT_OK;
}

midint_status midint_server_result_many(midint_server* s, size_t n, const uint64_t* ids, uint64_t* r,
                                        const uint64_t* offsets, size_t cap_limbs, size_t* n_limbs,
                                        double* latency_s) {
    if (!s || (n > 0 && (!ids || !r || !offsets))) return MIDINT_EINVAL;
    std::unordered_set<uint64_t> seen;
    for (size_t i = 0; i < n; i++) {
        if (!seen.insert(ids[i]).second) return MIDINT_EINVAL;  // an id given twice
        auto it = s->reqs.find(ids[i]);
        if (it == s->reqs.end()) return MIDINT_ENOTFOUND;
        if (!it->second.done) return MIDINT_EBUSY;
        if (offsets[i] > cap_limbs || cap_limbs - offsets[i] < it->second.result.size()) return MIDINT_EINVAL;
    }
    std::vector<const Request*> req(n);
    for (size_t i = 0; i < n; i++) {
        const Request& q = s->reqs.find(ids[i])->second;
        req[i] = &q;
        if (n_limbs) n_limbs[i] = q.result.size();
        if (latency_s) latency_s[i] = q.t_done - q.t_submit;
    }
    parallel_items(
        n, [&](size_t i) { return 8 * req[i]->result.size(); },
        [&](size_t i) { memcpy(r + offsets[i], req[i]->result.data(), req[i]->result.size() * 8); });
    for (size_t i = 0; i < n; i++) s->reqs.erase(ids[i]);
    return MIDINT_OK;
}

midint_status midint_server_stats(const midint_server* s, midint_server_counters* out) {
    if (!s || !out) return MIDINT_EINVAL;
    *out = s->st;
    return MIDINT_OK;
}

midint_status midint_server_destroy(midint_server* s) {
    if (!s) return MIDINT_EINVAL;
    destroy_server(s);
    return MIDINT_OK;
}

}  // extern "C"
