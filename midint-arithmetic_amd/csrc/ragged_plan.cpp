// Host side of ragged batches: packing ops into work items, upload, and the
// hipGraph replayed by midint_ragged_launch. With ragged.wave = 1 (the
// default) the graph is
//  * the additions, one per wave (add_rgw_kernel), and the multiplications
//    whose shorter operand has at most 128 limbs, one per wave on 16 x 16 x 64
//    tiles (mul_rgw_kernel; ragged.wave_split = 1: one node up to 70 limbs,
//    one above), costliest first;
//  * the other multiplications on the matrix-core nodes below;
// all as parallel branches. With ragged.wave = 0 it is a counter reset node
// followed by
//  * the multiplications whose shorter operand has at least ragged.mfma_min
//    limbs on the matrix cores, one product per workgroup: 32 x 32 x 32
//    tiles with ragged.mfma_waves@C waves (1, 2, 4 or 8) for size class C (a
//    power of two; the class of an op is that of its longer operand, (C/2, C]
//    limbs), or 16 x 16 x 64 tiles with ragged.m16_waves@C waves for the
//    products ragged.m16 sends there (0, the default: none; 1: all; 2: those
//    whose shorter operand has at least 1 / ragged.m16_ratio of the longer
//    one's limbs, where those tiles issue less work); one node per (tiles,
//    wave count), largest products first, LDS sized for the node's largest
//    product;
//  * one persistent kernel per (threads, Q) pair that the size classes of
//    the other ops select (ragged.threads@C: 256 or 512; ragged.q@C: 4 or 8
//    columns per lane group), each with its own work counter;
// all as parallel branches after the reset (ragged.overlap = 1), or the
// persistent kernels after every matrix-core node (0).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "midint.h"
#include "midint_internal.h"
#include "mul_core.hpp"

using namespace midint;

// One persistent launch: items of the ops whose size class selects these
// threads and Q.
struct RgLaunch {
    uint32_t threads = 256, q = 4;
    RgItem* d_items = nullptr;
    RgOp* d_ops = nullptr;
    uint32_t* d_counter = nullptr;  // element of the plan's counter array
    uint32_t n_items = 0, conv_cap = 0, grid = 0;
    size_t lds = 0;
    const void* fn = nullptr;
};

// One matrix-core node: the products of the size classes with one wave count.
struct MfLaunch {
    uint32_t waves = 0;
    MfRgItem* d_items = nullptr;
    uint32_t n_items = 0;
    MfRgConfig cfg = {};
};

// One node of one-op-per-wave items (products with a short operand, or
// additions).
struct WvLaunch {
    bool mul = false;
    int part = 0;  // products: rgw_mul_kernel_ptr part
    MfRgItem* d_items = nullptr;
    uint32_t n_items = 0;
};

struct midint_ragged_plan {
    int device = 0;
    uint32_t* d_counter = nullptr;  // one counter per persistent launch
    std::vector<RgLaunch> rg;
    std::vector<MfLaunch> mf;
    std::vector<WvLaunch> wv;
    uint32_t pow8 = 256;
    uint32_t n_counters = 0;
    const uint64_t* A = nullptr;
    const uint64_t* B = nullptr;
    uint64_t* R = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr uint32_t kConvLimit = 4 * kRgWindow;  // words: the L/H window area is reused for staging

uint32_t words_pad(uint32_t limbs, uint32_t Q) { return (2 * limbs + Q - 1) / Q * Q; }

void free_plan(midint_ragged_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    for (RgLaunch& l : p->rg) {
        if (l.d_items) (void)hipFree(l.d_items);
        if (l.d_ops) (void)hipFree(l.d_ops);
    }
    for (MfLaunch& l : p->mf)
        if (l.d_items) (void)hipFree(l.d_items);
    for (WvLaunch& l : p->wv)
        if (l.d_items) (void)hipFree(l.d_items);
    if (p->d_counter) (void)hipFree(p->d_counter);
    delete p;
}

}  // namespace

namespace midint {

// Packs ops into work items (see ragged.hip). Returns EINVAL / EUNSUPPORTED
// for malformed or oversized ops. Offsets in the returned RgOps are the ops'
// own limb offsets plus the given bases.
midint_status ragged_pack(const midint_op* ops, size_t n_ops, uint64_t base_a, uint64_t base_b, uint64_t base_r,
                          std::vector<RgItem>& h_items, std::vector<RgOp>& h_ops, uint32_t* conv_cap_out, uint32_t Q) {
    if (Q != 4 && Q != 8) return MIDINT_EINVAL;
    std::vector<RgOp> muls, adds;
    for (size_t i = 0; i < n_ops; i++) {
        const midint_op& o = ops[i];
        if (o.reserved != 0 || (o.kind != MIDINT_OP_ADD && o.kind != MIDINT_OP_MUL)) return MIDINT_EINVAL;
        if (o.n_a == 0 || o.n_b == 0) return MIDINT_EINVAL;
        if (o.n_a > kRgMaxLimbs || o.n_b > kRgMaxLimbs) return MIDINT_EUNSUPPORTED;
        RgOp r = {};
        r.off_r = base_r + o.off_r;
        if (o.kind == MIDINT_OP_MUL) {
            r.swap = o.n_a > o.n_b;
            r.off_s = r.swap ? base_b + o.off_b : base_a + o.off_a;
            r.off_l = r.swap ? base_a + o.off_a : base_b + o.off_b;
            r.ns = r.swap ? o.n_b : o.n_a;
            r.nl = r.swap ? o.n_a : o.n_b;
            r.ns_w = words_pad(r.ns, Q);
            r.nl_w = words_pad(r.nl, Q);
            r.lanes = r.nl_w / Q;
            r.seg_len = (r.ns_w + r.nl_w) / 2;
            r.out_len = o.n_a + o.n_b;
            muls.push_back(r);
        } else {
            r.off_s = base_a + o.off_a;
            r.off_l = base_b + o.off_b;
            r.ns = o.n_a;
            r.nl = o.n_b;
            r.seg_len = std::max(o.n_a, o.n_b) + 1;
            r.out_len = r.seg_len;
            adds.push_back(r);
        }
    }
    // ---- pack multiplications: sorted by the loop length (ns), then nl
    std::sort(muls.begin(), muls.end(), [](const RgOp& x, const RgOp& y) {
        return x.ns_w != y.ns_w ? x.ns_w < y.ns_w : x.nl_w < y.nl_w;
    });
    std::vector<RgItem> items;
    std::vector<uint64_t> cost;
    std::vector<std::vector<RgOp>> item_ops;
    uint32_t conv_cap = 0;
    for (size_t i = 0; i < muls.size();) {
        RgItem it = {kRgMul, 0, 0, 0, 0, 0};
        std::vector<RgOp> v;
        const uint32_t ns0 = muls[i].ns_w;
        uint64_t c = 0;
        while (i < muls.size() && v.size() < 255) {
            RgOp o = muls[i];
            const uint32_t conv = (o.ns_w + Q) + (o.nl_w + 2 * Q);
            if (!v.empty() && (it.lanes + o.lanes > kRgMaxLanes || it.window + o.seg_len > kRgWindow ||
                               it.conv_words + conv > kConvLimit || o.ns_w > ns0 + ns0 / 4 + 4))
                break;
            o.lane_base = it.lanes;
            o.seg_base = it.window;
            o.lds_base = it.conv_words;
            it.lanes += o.lanes;
            it.window += o.seg_len;
            it.conv_words += conv;
            c += (uint64_t)o.lanes * (o.ns_w / Q + 1) * Q * Q + 8ull * o.seg_len;
            v.push_back(o);
            i++;
        }
        it.count = (uint32_t)v.size();
        conv_cap = std::max(conv_cap, it.conv_words);
        items.push_back(it);
        cost.push_back(c);
        item_ops.push_back(std::move(v));
    }
    // ---- pack additions by output length
    std::sort(adds.begin(), adds.end(), [](const RgOp& x, const RgOp& y) { return x.seg_len < y.seg_len; });
    for (size_t i = 0; i < adds.size();) {
        RgItem it = {kRgAdd, 0, 0, 0, 0, 0};
        std::vector<RgOp> v;
        while (i < adds.size() && v.size() < 255 && it.window + adds[i].seg_len <= kRgWindow) {
            RgOp o = adds[i];
            o.seg_base = it.window;
            it.window += o.seg_len;
            v.push_back(o);
            i++;
        }
        it.count = (uint32_t)v.size();
        items.push_back(it);
        cost.push_back(4ull * it.window);
        item_ops.push_back(std::move(v));
    }
    // ---- longest items first
    std::vector<size_t> order(items.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    h_items.clear();
    h_ops.clear();
    for (size_t k : order) {
        RgItem it = items[k];
        it.first = (uint32_t)h_ops.size();
        h_items.push_back(it);
        h_ops.insert(h_ops.end(), item_ops[k].begin(), item_ops[k].end());
    }

    *conv_cap_out = (conv_cap + 3) / 4 * 4;
    return MIDINT_OK;
}

}  // namespace midint

extern "C" {

uint32_t midint_max_limbs_ragged(void) { return kRgMaxLimbs; }

midint_status midint_ragged_plan_create(midint_ctx* ctx, const midint_op* ops, size_t n_ops, const uint64_t* d_a,
                                        const uint64_t* d_b, uint64_t* d_r, const midint_ragged_tuning* tuning,
                                        midint_ragged_plan** out) {
    if (!ctx || !out || (n_ops > 0 && (!ops || !d_a || !d_b || !d_r))) return MIDINT_EINVAL;
    *out = nullptr;
    // multiplications for the matrix-core nodes: shorter operand >= mfma_min
    // limbs (0: none), grouped by the size class of the longer operand; the
    // other ops go to the persistent launch their class selects
    const int64_t mf_min = tune_get("ragged.mfma_min", 32);
    // ragged.wave = 1 (default): additions and the products whose shorter
    // operand has at most rgw_max_short() limbs run one op per wave
    const bool by_wave = tune_get("ragged.wave", 1) != 0;
    // ragged.wave_split = 1: the products up to rgw_small_max_short() limbs on
    // a kernel of their own (a second node)
    const bool wv_split = tune_get("ragged.wave_split", 0) != 0;
    std::vector<std::vector<MfRgItem>> mf_by_class(32);  // index: log2 of the size class
    std::vector<midint_op> rest_by_launch[4];  // index: (threads == 512) * 2 + (q == 8)
    std::vector<MfRgItem> wv_mul, wv_mul_small, wv_add;
    for (size_t i = 0; i < n_ops; i++) {
        const midint_op& o = ops[i];
        if (o.reserved != 0 || (o.kind != MIDINT_OP_ADD && o.kind != MIDINT_OP_MUL)) return MIDINT_EINVAL;
        if (o.n_a == 0 || o.n_b == 0) return MIDINT_EINVAL;
        if (o.n_a > kRgMaxLimbs || o.n_b > kRgMaxLimbs) return MIDINT_EUNSUPPORTED;
        const uint32_t nmax = std::max(o.n_a, o.n_b);
        if (by_wave && o.kind == MIDINT_OP_ADD) {
            MfRgItem it = {};
            it.a = d_a + o.off_a;
            it.b = d_b + o.off_b;
            it.r = d_r + o.off_r;
            it.n_a = (uint32_t)o.n_a;
            it.n_b = (uint32_t)o.n_b;
            wv_add.push_back(it);
        } else if (by_wave && o.kind == MIDINT_OP_MUL && std::min(o.n_a, o.n_b) <= rgw_max_short()) {
            const bool sw = o.n_a < o.n_b;  // the longer operand is the kernel's a
            MfRgItem it = {};
            it.a = sw ? d_b + o.off_b : d_a + o.off_a;
            it.b = sw ? d_a + o.off_a : d_b + o.off_b;
            it.r = d_r + o.off_r;
            it.n_a = (uint32_t)(sw ? o.n_b : o.n_a);
            it.n_b = (uint32_t)(sw ? o.n_a : o.n_b);
            (wv_split && it.n_b <= rgw_small_max_short() ? wv_mul_small : wv_mul).push_back(it);
        } else if (o.kind == MIDINT_OP_MUL && (by_wave || (mf_min > 0 && std::min(o.n_a, o.n_b) >= (uint64_t)mf_min))) {
            const bool sw = o.n_a < o.n_b;  // the longer operand is the kernel's a
            const uint64_t* x = sw ? d_b + o.off_b : d_a + o.off_a;
            const uint64_t* y = sw ? d_a + o.off_a : d_b + o.off_b;
            const uint32_t nx = (uint32_t)(sw ? o.n_b : o.n_a), ny = (uint32_t)(sw ? o.n_a : o.n_b);
            uint32_t c = 0;
            while ((1u << c) < nmax) c++;
            mf_by_class[c].push_back(mfma_rg_item(x, y, d_r + o.off_r, nx, ny));
        } else {
            int64_t th = (tuning && tuning->threads) ? tuning->threads : tune_get_class("ragged.threads", nmax, 256);
            const int64_t q = tune_get_class("ragged.q", nmax, 4);
            rest_by_launch[(th == 512 ? 2 : 0) + (q == 8 ? 1 : 0)].push_back(o);
        }
    }

    midint_ragged_plan* p = new (std::nothrow) midint_ragged_plan;
    if (!p) return MIDINT_ENOMEM;
    p->device = 0;
    if (hipGetDevice(&p->device) != hipSuccess) {
        free_plan(p);
        return MIDINT_EHIP;
    }
    p->A = d_a;
    p->B = d_b;
    p->R = d_r;
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device);

    // ---- matrix-core nodes, one per (tiles, wave count). A class whose
    // products do not fit its wave count's rows (one wave holds at most 1024
    // output limbs) runs with the default count.
    auto config = [](int k16, uint32_t na, uint32_t nb, uint32_t pl, MfRgConfig* out, uint32_t w) {
        return k16 ? mfma16_rg_config(na, nb, pl, out, w) : mfma32_rg_config(na, nb, pl, out, w);
    };
    std::vector<MfRgItem> mf_by_waves[2][4];  // index: 16 x 16 x 64 tiles, log2(waves)
    for (int c = 0; c < 32; c++) {
        if (mf_by_class[c].empty()) continue;
        const int64_t m16_mode = tune_get_class("ragged.m16", 1u << c, 0);
        const int64_t m16_ratio = std::max<int64_t>(1, tune_get_class("ragged.m16_ratio", 1u << c, 2));
        std::vector<MfRgItem> part[2];
        for (const MfRgItem& it : mf_by_class[c]) {
            const bool k16 = m16_mode == 1 || (m16_mode == 2 && (uint64_t)it.n_b * m16_ratio >= it.n_a);
            part[k16 ? 1 : 0].push_back(it);
        }
        for (int k16 = 0; k16 < 2; k16++) {
            if (part[k16].empty()) continue;
            uint32_t na = 0, nb = 0, pl = 0;
            for (const MfRgItem& it : part[k16]) {
                na = std::max(na, it.n_a);
                nb = std::max(nb, it.n_b);
                pl = std::max(pl, it.n_a + it.n_b);
            }
            auto pick_waves = [&](int kk) {
                const int64_t w = kk ? tune_get_class("ragged.m16_waves", 1u << c, kMfRgWaves)
                                     : tune_get_class("ragged.mfma_waves", 1u << c, kMfRgWaves);
                MfRgConfig probe;
                return (w == 1 || w == 2 || w == 4 || w == 8) && config(kk, na, nb, pl, &probe, (uint32_t)w)
                           ? (uint32_t)w
                           : kMfRgWaves;
            };
            int kk = k16;
            uint32_t waves = pick_waves(kk);
            MfRgConfig probe;
            if (kk && !config(kk, na, nb, pl, &probe, waves)) {  // 16 x 16 x 64 tiles do not fit
                kk = 0;
                waves = pick_waves(kk);
            }
            const int lw = waves == 1 ? 0 : waves == 2 ? 1 : waves == 8 ? 3 : 2;
            mf_by_waves[kk][lw].insert(mf_by_waves[kk][lw].end(), part[k16].begin(), part[k16].end());
        }
    }
    for (int g = 7; e == hipSuccess && g >= 0; g--) {
        const int k16 = g & 1, lw = g >> 1;
        std::vector<MfRgItem>& v = mf_by_waves[k16][lw];
        if (v.empty()) continue;
        std::stable_sort(v.begin(), v.end(), [](const MfRgItem& x, const MfRgItem& y) {
            return (uint64_t)x.n_a * x.n_b > (uint64_t)y.n_a * y.n_b;
        });
        uint32_t na = 0, nb = 0, pl = 0;
        for (const MfRgItem& it : v) {
            na = std::max(na, it.n_a);
            nb = std::max(nb, it.n_b);
            pl = std::max(pl, it.n_a + it.n_b);
        }
        MfLaunch l;
        l.waves = 1u << lw;
        if (!config(k16, na, nb, pl, &l.cfg, l.waves)) {
            free_plan(p);
            return MIDINT_EUNSUPPORTED;
        }
        l.n_items = (uint32_t)v.size();
        e = hipMalloc((void**)&l.d_items, v.size() * sizeof(MfRgItem));
        if (e == hipSuccess) e = hipMemcpy(l.d_items, v.data(), v.size() * sizeof(MfRgItem), hipMemcpyHostToDevice);
        if (e == hipSuccess && l.cfg.lds > 64 * 1024)
            e = hipFuncSetAttribute(l.cfg.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.cfg.lds);
        p->mf.push_back(l);
    }

    // ---- one-op-per-wave nodes: the costliest ops first (the workgroups are
    // dispatched in order as the CUs free up)
    auto by_cost = [](const MfRgItem& x, const MfRgItem& y) {
        const uint64_t cx = (uint64_t)((x.n_a + x.n_b + 127) / 128) * ((8 * x.n_b + 15) / 64 + 1);
        const uint64_t cy = (uint64_t)((y.n_a + y.n_b + 127) / 128) * ((8 * y.n_b + 15) / 64 + 1);
        return cx > cy;
    };
    std::stable_sort(wv_mul.begin(), wv_mul.end(), by_cost);
    std::stable_sort(wv_mul_small.begin(), wv_mul_small.end(), by_cost);
    std::stable_sort(wv_add.begin(), wv_add.end(), [](const MfRgItem& x, const MfRgItem& y) {
        return std::max(x.n_a, x.n_b) > std::max(y.n_a, y.n_b);
    });
    for (int k = 0; e == hipSuccess && k < 3; k++) {
        std::vector<MfRgItem>& v = k == 0 ? wv_mul : k == 1 ? wv_mul_small : wv_add;
        if (v.empty()) continue;
        WvLaunch l;
        l.mul = k < 2;
        l.part = !wv_split ? 0 : k == 1 ? 1 : 2;
        l.n_items = (uint32_t)v.size();
        e = hipMalloc((void**)&l.d_items, v.size() * sizeof(MfRgItem));
        if (e == hipSuccess) e = hipMemcpy(l.d_items, v.data(), v.size() * sizeof(MfRgItem), hipMemcpyHostToDevice);
        p->wv.push_back(l);
    }

    // ---- persistent launches
    for (int k = 0; e == hipSuccess && k < 4; k++) {
        if (rest_by_launch[k].empty()) continue;
        RgLaunch l;
        l.threads = (k & 2) ? 512 : 256;
        l.q = (k & 1) ? 8 : 4;
        std::vector<RgItem> h_items;
        std::vector<RgOp> h_ops;
        midint_status st = ragged_pack(rest_by_launch[k].data(), rest_by_launch[k].size(), 0, 0, 0, h_items, h_ops,
                                       &l.conv_cap, l.q);
        if (st != MIDINT_OK) {
            free_plan(p);
            return st;
        }
        l.n_items = (uint32_t)h_items.size();
        l.lds = ragged_lds_size(l.threads, l.conv_cap);
        l.fn = ragged_kernel_ptr(l.threads, l.q);
        if (!h_items.empty()) e = hipMalloc((void**)&l.d_items, h_items.size() * sizeof(RgItem));
        if (e == hipSuccess && !h_ops.empty()) e = hipMalloc((void**)&l.d_ops, h_ops.size() * sizeof(RgOp));
        if (e == hipSuccess && !h_items.empty())
            e = hipMemcpy(l.d_items, h_items.data(), h_items.size() * sizeof(RgItem), hipMemcpyHostToDevice);
        if (e == hipSuccess && !h_ops.empty())
            e = hipMemcpy(l.d_ops, h_ops.data(), h_ops.size() * sizeof(RgOp), hipMemcpyHostToDevice);
        if (e == hipSuccess && l.lds > 64 * 1024)
            e = hipFuncSetAttribute(l.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.lds);
        int per_cu = 0;
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, l.fn, (int)l.threads, l.lds);
        if (per_cu < 1) per_cu = 1;
        if (tuning && tuning->blocks_per_cu > 0 && (int)tuning->blocks_per_cu < per_cu)
            per_cu = (int)tuning->blocks_per_cu;
        l.grid = (uint32_t)std::max(1, std::min<int>(per_cu * cus, (int)std::max<uint32_t>(1, l.n_items)));
        p->rg.push_back(l);
    }
    p->n_counters = (uint32_t)p->rg.size();
    if (e == hipSuccess) e = hipMalloc((void**)&p->d_counter, 4 * sizeof(uint32_t));
    for (size_t k = 0; k < p->rg.size(); k++) p->rg[k].d_counter = p->d_counter + k;
    if (e != hipSuccess) {
        free_plan(p);
        return MIDINT_EHIP;
    }

    // ---- graph: reset the work counters, then the matrix-core nodes and the
    // persistent launches (the vectors are final: kernel arguments point into
    // them)
    e = hipGraphCreate(&p->graph, 0);
    hipGraphNode_t n_reset = nullptr;
    if (e == hipSuccess && !p->rg.empty()) {
        void* rargs[] = {(void*)&p->d_counter, (void*)&p->n_counters};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(ragged_reset_ptr());
        kp.gridDim = dim3(1);
        kp.blockDim = dim3(64);
        kp.sharedMemBytes = 0;
        kp.kernelParams = rargs;
        kp.extra = nullptr;
        e = hipGraphAddKernelNode(&n_reset, p->graph, nullptr, 0, &kp);
    }
    std::vector<hipGraphNode_t> mf_nodes;
    for (MfLaunch& l : p->mf) {
        if (e != hipSuccess) break;
        void* margs[] = {(void*)&l.d_items, (void*)&l.cfg.pow8};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(l.cfg.func);
        kp.gridDim = dim3(l.n_items);
        kp.blockDim = dim3(l.cfg.threads);
        kp.sharedMemBytes = (unsigned)l.cfg.lds;
        kp.kernelParams = margs;
        kp.extra = nullptr;
        hipGraphNode_t n = nullptr;
        e = hipGraphAddKernelNode(&n, p->graph, n_reset ? &n_reset : nullptr, n_reset ? 1 : 0, &kp);
        mf_nodes.push_back(n);
    }
    for (WvLaunch& l : p->wv) {
        if (e != hipSuccess) break;
        void* wargs[] = {(void*)&l.d_items, (void*)&l.n_items, (void*)&p->pow8};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(l.mul ? rgw_mul_kernel_ptr(l.part) : rgw_add_kernel_ptr());
        kp.gridDim = dim3((l.n_items + rgw_waves() - 1) / rgw_waves());
        kp.blockDim = dim3(64 * rgw_waves());
        kp.sharedMemBytes = l.mul ? (unsigned)rgw_lds_bytes(l.part) : 0u;
        kp.kernelParams = wargs;
        kp.extra = nullptr;
        hipGraphNode_t n = nullptr;
        e = hipGraphAddKernelNode(&n, p->graph, n_reset ? &n_reset : nullptr, n_reset ? 1 : 0, &kp);
        mf_nodes.push_back(n);
    }
    // ragged.overlap = 1: the persistent kernels do not wait for the
    // matrix-core nodes (the branches write disjoint result ranges)
    std::vector<hipGraphNode_t> deps = {n_reset};
    if (tune_get("ragged.overlap", 1) == 0 && !mf_nodes.empty()) deps = mf_nodes;
    for (RgLaunch& l : p->rg) {
        if (e != hipSuccess || l.n_items == 0) continue;
        void* kargs[] = {(void*)&l.d_items, (void*)&l.d_ops, (void*)&l.n_items, (void*)&l.d_counter,
                         (void*)&p->A,      (void*)&p->B,    (void*)&p->R,      (void*)&l.conv_cap};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(l.fn);
        kp.gridDim = dim3(l.grid);
        kp.blockDim = dim3(l.threads);
        kp.sharedMemBytes = (unsigned)l.lds;
        kp.kernelParams = kargs;
        kp.extra = nullptr;
        hipGraphNode_t n = nullptr;
        e = hipGraphAddKernelNode(&n, p->graph, deps.data(), deps.size(), &kp);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        free_plan(p);
        return MIDINT_EHIP;
    }
    *out = p;
    return MIDINT_OK;
}

midint_status midint_ragged_launch(midint_ragged_plan* p, void* stream) {
    if (!p) return MIDINT_EINVAL;
    if (p->rg.empty() && p->mf.empty() && p->wv.empty()) return MIDINT_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return MIDINT_EHIP;
    if (cur != p->device && hipSetDevice(p->device) != hipSuccess) return MIDINT_EHIP;
    return hipGraphLaunch(p->exec, (hipStream_t)stream) == hipSuccess ? MIDINT_OK : MIDINT_EHIP;
}

midint_status midint_ragged_plan_stats(const midint_ragged_plan* p, uint64_t* n_items, uint64_t* grid,
                                       uint64_t* lds_bytes) {
    if (!p) return MIDINT_EINVAL;
    uint64_t items = 0, g = 0, lds = 0;
    for (const RgLaunch& l : p->rg) {
        items += l.n_items;
        g += l.grid;
        lds = std::max<uint64_t>(lds, l.lds);
    }
    for (const MfLaunch& l : p->mf) items += l.n_items;
    for (const WvLaunch& l : p->wv) items += l.n_items;
    if (n_items) *n_items = items;
    if (grid) *grid = g;
    if (lds_bytes) *lds_bytes = lds;
    return MIDINT_OK;
}

midint_status midint_ragged_plan_destroy(midint_ragged_plan* p) {
    if (!p) return MIDINT_EINVAL;
    (void)hipSetDevice(p->device);
    (void)hipDeviceSynchronize();
    free_plan(p);
    return MIDINT_OK;
}

}  // extern "C"
