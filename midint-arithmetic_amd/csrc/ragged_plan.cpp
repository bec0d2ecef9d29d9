// Host side of ragged batches: packing ops into work items, upload, and the
// hipGraph replayed by midint_ragged_launch: counter reset node -> matrix-core
// node (multiplications whose shorter operand has at least ragged.mfma_min
// limbs, one product per workgroup, largest first) -> persistent kernel node
// (every other op; with ragged.overlap = 1 it depends on the reset only).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "midint.h"
#include "midint_internal.h"
#include "mul_core.hpp"

using namespace midint;

struct midint_ragged_plan {
    int device = 0;
    RgItem* d_items = nullptr;
    RgOp* d_ops = nullptr;
    uint32_t* d_counter = nullptr;
    MfRgItem* d_mf_items = nullptr;
    uint32_t n_items = 0;
    uint32_t n_mf_items = 0;
    MfRgConfig mf = {};
    uint32_t threads = 256;
    uint32_t conv_cap = 0;
    uint32_t grid = 0;
    size_t lds = 0;
    const uint64_t* A = nullptr;
    const uint64_t* B = nullptr;
    uint64_t* R = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr uint32_t kConvLimit = 4 * kRgWindow;  // words: the L/H window area is reused for staging

uint32_t words_pad(uint32_t limbs) { return (2 * limbs + kMulQ - 1) / kMulQ * kMulQ; }

void free_plan(midint_ragged_plan* p) {
    if (!p) return;
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->d_items) (void)hipFree(p->d_items);
    if (p->d_ops) (void)hipFree(p->d_ops);
    if (p->d_counter) (void)hipFree(p->d_counter);
    if (p->d_mf_items) (void)hipFree(p->d_mf_items);
    delete p;
}

}  // namespace

namespace midint {

// Packs ops into work items (see ragged.hip). Returns EINVAL / EUNSUPPORTED
// for malformed or oversized ops. Offsets in the returned RgOps are the ops'
// own limb offsets plus the given bases.
midint_status ragged_pack(const midint_op* ops, size_t n_ops, uint64_t base_a, uint64_t base_b, uint64_t base_r,
                          std::vector<RgItem>& h_items, std::vector<RgOp>& h_ops, uint32_t* conv_cap_out) {
    constexpr uint32_t Q = kMulQ;
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
            r.ns_w = words_pad(r.ns);
            r.nl_w = words_pad(r.nl);
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
    // multiplications for the matrix-core node: shorter operand >= mfma_min
    // limbs (0: none); the rest go to the persistent kernel
    const int64_t mf_min = tune_get("ragged.mfma_min", 32);
    std::vector<midint_op> rest;
    std::vector<MfRgItem> h_mf;
    uint32_t mf_na = 0, mf_nb = 0, mf_pl = 0;
    rest.reserve(n_ops);
    for (size_t i = 0; i < n_ops; i++) {
        const midint_op& o = ops[i];
        const bool ok = o.reserved == 0 && o.kind == MIDINT_OP_MUL && o.n_a <= kRgMaxLimbs && o.n_b <= kRgMaxLimbs;
        if (ok && mf_min > 0 && std::min(o.n_a, o.n_b) >= (uint64_t)mf_min) {
            const bool sw = o.n_a < o.n_b;  // the longer operand is the kernel's a
            const uint64_t* x = sw ? d_b + o.off_b : d_a + o.off_a;
            const uint64_t* y = sw ? d_a + o.off_a : d_b + o.off_b;
            const uint32_t nx = (uint32_t)(sw ? o.n_b : o.n_a), ny = (uint32_t)(sw ? o.n_a : o.n_b);
            h_mf.push_back(mfma_rg_item(x, y, d_r + o.off_r, nx, ny));
            mf_na = std::max(mf_na, nx);
            mf_nb = std::max(mf_nb, ny);
            mf_pl = std::max(mf_pl, nx + ny);
        } else {
            rest.push_back(o);
        }
    }
    std::stable_sort(h_mf.begin(), h_mf.end(), [](const MfRgItem& x, const MfRgItem& y) {
        return (uint64_t)x.n_a * x.n_b > (uint64_t)y.n_a * y.n_b;
    });
    MfRgConfig mf = {};
    if (!h_mf.empty() && !mfma_rg_config(mf_na, mf_nb, mf_pl, &mf)) return MIDINT_EUNSUPPORTED;
    std::vector<RgItem> h_items;
    std::vector<RgOp> h_ops;
    uint32_t conv_cap = 0;
    midint_status st = ragged_pack(rest.data(), rest.size(), 0, 0, 0, h_items, h_ops, &conv_cap);
    if (st != MIDINT_OK) return st;

    midint_ragged_plan* p = new (std::nothrow) midint_ragged_plan;
    if (!p) return MIDINT_ENOMEM;
    p->device = 0;
    if (hipGetDevice(&p->device) != hipSuccess) {
        free_plan(p);
        return MIDINT_EHIP;
    }
    const int64_t th = (tuning && tuning->threads) ? tuning->threads : tune_get("ragged.threads", 256);
    p->threads = th == 512 ? 512 : 256;
    p->n_items = (uint32_t)h_items.size();
    p->n_mf_items = (uint32_t)h_mf.size();
    p->mf = mf;
    p->conv_cap = conv_cap;
    p->lds = ragged_lds_size(p->threads, p->conv_cap);
    p->A = d_a;
    p->B = d_b;
    p->R = d_r;

    hipError_t e = hipMalloc((void**)&p->d_counter, sizeof(uint32_t));
    if (e == hipSuccess && !h_items.empty()) e = hipMalloc((void**)&p->d_items, h_items.size() * sizeof(RgItem));
    if (e == hipSuccess && !h_ops.empty()) e = hipMalloc((void**)&p->d_ops, h_ops.size() * sizeof(RgOp));
    if (e == hipSuccess && !h_items.empty())
        e = hipMemcpy(p->d_items, h_items.data(), h_items.size() * sizeof(RgItem), hipMemcpyHostToDevice);
    if (e == hipSuccess && !h_ops.empty())
        e = hipMemcpy(p->d_ops, h_ops.data(), h_ops.size() * sizeof(RgOp), hipMemcpyHostToDevice);
    if (e == hipSuccess && !h_mf.empty()) e = hipMalloc((void**)&p->d_mf_items, h_mf.size() * sizeof(MfRgItem));
    if (e == hipSuccess && !h_mf.empty())
        e = hipMemcpy(p->d_mf_items, h_mf.data(), h_mf.size() * sizeof(MfRgItem), hipMemcpyHostToDevice);
    if (e == hipSuccess && !h_mf.empty() && mf.lds > 64 * 1024)
        e = hipFuncSetAttribute(mf.func, hipFuncAttributeMaxDynamicSharedMemorySize, (int)mf.lds);
    const void* kfn = ragged_kernel_ptr(p->threads);
    if (e == hipSuccess && p->lds > 64 * 1024)
        e = hipFuncSetAttribute(kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds);
    int per_cu = 0, cus = 0;
    if (e == hipSuccess)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, (int)p->threads, p->lds);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device);
    if (e != hipSuccess) {
        free_plan(p);
        return MIDINT_EHIP;
    }
    if (per_cu < 1) per_cu = 1;
    if (tuning && tuning->blocks_per_cu > 0 && (int)tuning->blocks_per_cu < per_cu) per_cu = (int)tuning->blocks_per_cu;
    p->grid = (uint32_t)std::max(1, std::min<int>(per_cu * cus, (int)std::max<uint32_t>(1, p->n_items)));

    // ---- graph: reset the work counter, the matrix-core products, then the
    // persistent kernel
    e = hipGraphCreate(&p->graph, 0);
    hipGraphNode_t n_reset = nullptr, n_mf = nullptr, n_main = nullptr;
    if (e == hipSuccess) {
        void* rargs[] = {(void*)&p->d_counter};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(ragged_reset_ptr());
        kp.gridDim = dim3(1);
        kp.blockDim = dim3(64);
        kp.sharedMemBytes = 0;
        kp.kernelParams = rargs;
        kp.extra = nullptr;
        e = hipGraphAddKernelNode(&n_reset, p->graph, nullptr, 0, &kp);
    }
    hipGraphNode_t n_prev = n_reset;
    if (e == hipSuccess && p->n_mf_items > 0) {
        void* margs[] = {(void*)&p->d_mf_items, (void*)&p->mf.pow8};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(p->mf.func);
        kp.gridDim = dim3(p->n_mf_items);
        kp.blockDim = dim3(p->mf.threads);
        kp.sharedMemBytes = (unsigned)p->mf.lds;
        kp.kernelParams = margs;
        kp.extra = nullptr;
        e = hipGraphAddKernelNode(&n_mf, p->graph, &n_reset, 1, &kp);
        // ragged.overlap = 1: the persistent kernel does not wait for the
        // matrix-core node (parallel branches after the reset)
        if (tune_get("ragged.overlap", 1) == 0) n_prev = n_mf;
    }
    if (e == hipSuccess && p->n_items > 0) {
        void* kargs[] = {(void*)&p->d_items, (void*)&p->d_ops, (void*)&p->n_items, (void*)&p->d_counter,
                         (void*)&p->A,       (void*)&p->B,     (void*)&p->R,       (void*)&p->conv_cap};
        hipKernelNodeParams kp = {};
        kp.func = const_cast<void*>(kfn);
        kp.gridDim = dim3(p->grid);
        kp.blockDim = dim3(p->threads);
        kp.sharedMemBytes = (unsigned)p->lds;
        kp.kernelParams = kargs;
        kp.extra = nullptr;
        e = hipGraphAddKernelNode(&n_main, p->graph, &n_prev, 1, &kp);
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
    if (p->n_items == 0 && p->n_mf_items == 0) return MIDINT_OK;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return MIDINT_EHIP;
    if (cur != p->device && hipSetDevice(p->device) != hipSuccess) return MIDINT_EHIP;
    return hipGraphLaunch(p->exec, (hipStream_t)stream) == hipSuccess ? MIDINT_OK : MIDINT_EHIP;
}

midint_status midint_ragged_plan_stats(const midint_ragged_plan* p, uint64_t* n_items, uint64_t* grid,
                                       uint64_t* lds_bytes) {
    if (!p) return MIDINT_EINVAL;
    if (n_items) *n_items = (uint64_t)p->n_items + p->n_mf_items;
    if (grid) *grid = p->grid;
    if (lds_bytes) *lds_bytes = p->lds;
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
