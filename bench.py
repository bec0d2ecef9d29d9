"""Benchmark of the midint HIP kernels on one MI355X.

Headline (``value``): big-integer multiplications per second at 8192 bits --
16384 pairs of 128-limb operands, full 256-limb products (classical
schoolbook, on the kernel the tuning file selects for 128 limbs), inputs
resident in HBM, one step = one batch. The same JSON line carries, under
``extra``:

  * add_65536x64   -- batched addition (HBM roofline) beside the streaming
                      ceiling of the same access pattern (r = a + b per limb,
                      no carries, same launch shape and buffers), and the
                      six-add composite (a6pb10Ker) on the same buffers;
  * ntt_1024x2048  -- NTT multiplication of 131072-bit integers;
  * composites     -- polyKerQ / polyFttKer equivalents with the reference's
                      Gu32ops metric (cuda/main.cu:317-351, 526-570);
  * ragged_50000   -- one persistent hipGraph launch vs plain launches per
                      power-of-two size class and per exact operand sizes;
  * serving        -- Poisson streams up to a saturating burst;
  * cpu            -- GMP (the library the reference's CPU path calls,
                      cuda/helper.h:146-170) on every workload, all host
                      threads used and one core.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
                       [--only all|core] [--calibrate]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
INT8_PEAK_TOPS = 5000.0  # MI355X dense int8 MFMA, 5.0 POP/s
MIB = 1 << 20
CPU_THREAD_CAP = 16  # host CPU share of one GPU on the measurement boxes


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def tuned_value(mi, key, m, default):
    """The launch parameter the library uses for operands of m limbs: the
    key at m's power-of-two class, else the global key, else `default`."""
    c = 1
    while c < m:
        c <<= 1
    for k in (f"{key}@{c}", key):
        try:
            return mi.tune_get(k)
        except Exception:
            pass
    return default


def mfma_work(m, wpi=0):
    """(useful, issued) int8 ops of one m x m-limb full product on the matrix
    cores: balanced base-256 digits, D = 8m + 1 per operand, useful = 2 D^2
    digit products; issued = 2 x 32768 MACs per v_mfma_i32_32x32x32_i8, one
    per column tile and d step of each pass (up to two tiles c0, c0 + wpi per
    pass over the union of their d ranges, as the kernel runs them)."""
    D = 8 * m + 1
    P = m // 4 + 1
    dmax = (D + 30) // 32
    N = (2 * m + 3) // 4
    NT = (N + 31) // 32
    if wpi <= 0:
        wpi = (NT + 2) // 4
    wpi = max(1, min(wpi, 16, NT))
    mfmas = 0
    for wi in range(wpi):
        for c0 in range(wi, NT, 2 * wpi):
            tiles = [c0 + k * wpi for k in range(2) if c0 + k * wpi < NT]
            lo = min(max(0, 32 * ct - P + 1) for ct in tiles)
            hi = max(min(dmax, 32 * ct + 31) for ct in tiles)
            mfmas += max(0, hi - lo + 1) * len(tiles)
    return 2 * D * D, 2 * 32768 * mfmas


def wave16_work(m, full=True, delta=48):
    """(useful, issued) int8 ops of one m x m-limb product on the
    one-product-per-wave kernel: one v_mfma_i32_16x16x64_i8 (2 x 16384 ops)
    per tile of 16 output blocks of 16 digits and per step k that has a
    nonzero term (a window of 64 digits of a starting at 16(I - 4k) - delta
    against digits of b, as the kernel's compile-time plan counts them)."""
    D = 8 * m + 1
    nblk = m if full else m // 2
    mfmas = 0
    for t in range(nblk // 16):
        for k in range(-8, D // 64 + 9):
            for n in range(16):
                I = 16 * t + n
                L = 16 * I - 64 * k - delta
                ilo, ihi = max(L, 0), min(L + 63, D - 1)
                if ilo <= ihi and ilo <= 16 * I + 15 and ihi + D - 1 >= 16 * I:
                    mfmas += 1
                    break
    return 2 * D * D, 2 * 16384 * mfmas


def gu32ops_of_mul(m):
    """The reference's numAd32OpsOfMultInst (cuda/helper.h:36-43) for an
    operand of m u64 limbs: 300 n log2 n with n = 2m u32 words."""
    n = 2 * m
    return 300 * n * int(math.log2(n))


def timed(torch, fn, steps, warmup):
    """Mean ms per call of fn() over `steps` calls after `warmup` untimed
    calls, HIP events on torch's current stream (the stream the kernels are
    launched on)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def to_dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x).view(np.int64)).to("cuda:0")


def measured_traffic(name):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC
    summary (profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)["kernels"][name]["hbm_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def profiled_valu_per_product(name):
    """wave64 VALU instructions per product of a kernel (SQ_INSTS_VALU of the
    committed rocprofv3 summary, profiles/traffic.json), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            return json.load(f)["kernels"][name]["valu_wave_instr_per_product"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_threads():
    """Host threads for the CPU baselines: the CPUs this process may run on,
    capped at one GPU's share of a measurement box."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover - non-Linux
        avail = os.cpu_count() or 1
    return max(1, min(CPU_THREAD_CAP, avail))


def rate(fn, units, budget_s):
    """Units per second of fn() repeated for about budget_s of wall time
    (at least twice; the first call is a warm-up)."""
    fn()
    reps, t0 = 0, time.perf_counter()
    while True:
        fn()
        reps += 1
        el = time.perf_counter() - t0
        if el >= budget_s and reps >= 2:
            return reps * units / el


def cpu_baselines(orc, inputs, threads, budget_s=1.2):
    """GMP on each workload with `threads` threads and with one, plus the
    oracle's plain-C restatement of the classical kernel on the headline."""
    out = {"threads": threads, "host_logical_cpus": os.cpu_count(), "library": "libgmp (mpn_add_n / mpn_mul / "
           "mpn_add), the library the reference's CPU path calls (cuda/helper.h:146-170, mpz_add / mpz_mul)"}
    if not orc.gmp_available():
        out["gmp"] = None
        return out
    a1, b1 = inputs["add"]
    a2, b2 = inputs["mul"]
    a3, b3 = inputs["ntt"]
    ops, ra, rb, rl = inputs["ragged"]
    # D2 sample: 2048 of the 16384 pairs; D3: 64 of the 1024 pairs
    k2, k3 = 2048, 64
    s2a, s2b = a2[: k2 * 128], b2[: k2 * 128]
    s3a, s3b = a3[: k3 * 2048], b3[: k3 * 2048]
    n_mul = sum(1 for o in ops if o[0] == 1)
    prep = orc.gmp_ragged_prepare(ops)
    rbuf = np.zeros(rl, dtype=np.uint64)
    for t, key in ((threads, "all"), (1, "one_core")):
        out[key] = {
            "add_65536x64_GB_per_s": rate(lambda: orc.gmp_add_batch(a1, b1, 64, threads=t), 3 * a1.size * 8,
                                          budget_s) / 1e9,
            "mul_16384x128_muls_per_s": rate(lambda: orc.gmp_mul_full(s2a, s2b, 128, threads=t), k2, budget_s),
            "mul_1024x2048_muls_per_s": rate(lambda: orc.gmp_mul_full(s3a, s3b, 2048, threads=t), k3, budget_s),
            "ragged_50000_ops_per_s": rate(lambda: orc.gmp_ragged(prep, ra, rb, rl, threads=t, out=rbuf), len(ops),
                                           budget_s),
        }
        out[key]["ragged_50000_muls_per_s"] = out[key]["ragged_50000_ops_per_s"] * n_mul / len(ops)
    out["port_mul_16384x128_muls_per_s"] = rate(lambda: orc.mul_full(s2a, s2b, 128, threads=threads), k2, budget_s)
    out["samples"] = {"add": "all 65536 pairs", "mul_16384x128": f"{k2} of the 16384 pairs",
                      "mul_1024x2048": f"{k3} of the 1024 pairs", "ragged": "all 50000 ops"}
    return out


# --------------------------------------------------------------------------- workloads
def bench_add(torch, mi, ctx, args):
    """65536 x 64-limb addition over 4 rotating 96 MiB buffer sets (past the
    256 MiB Infinity Cache), the streaming ceiling of the same access
    pattern, and the six-add composite."""
    from workloads import make_inputs

    na, ma = 65536, 64
    aa, ba = make_inputs("add_65536x64", mi.gen_limbs)
    sets = [(to_dev(torch, aa), to_dev(torch, ba), torch.empty(na * ma, dtype=torch.int64, device="cuda:0"))
            for _ in range(4)]
    it = [0]

    def nxt():
        x = sets[it[0] & 3]
        it[0] += 1
        return x

    def add_step():
        x, y, z = nxt()
        ctx.add(x, y, z, ma)

    def stream_step():
        x, y, z = nxt()
        mi.probe_stream(x, y, z)

    def add6_step():
        x, y, z = nxt()
        ctx.add6(x, y, z, ma)

    steps, warm = 2 * args.steps, max(4, args.warmup)
    # interleaved repeats so that the add and its ceiling see the same clocks
    ms_add, ms_ceil = [], []
    for _ in range(3):
        ms_add.append(timed(torch, add_step, steps, warm))
        ms_ceil.append(timed(torch, stream_step, steps, warm))
    ms_a, ms_c = min(ms_add), min(ms_ceil)
    ms_6 = timed(torch, add6_step, steps, warm)
    byts = 3 * na * ma * 8
    gbs, gbs_c, gbs_6 = (byts / (t * 1e-3) / 1e9 for t in (ms_a, ms_c, ms_6))
    log(f"add {na} x {ma}: {ms_a * 1e3:.2f} us ({gbs:.0f} GB/s); streaming ceiling {ms_c * 1e3:.2f} us "
        f"({gbs_c:.0f} GB/s); add6 {ms_6 * 1e3:.2f} us ({gbs_6:.0f} GB/s)")
    del sets
    return {
        "us_per_launch": ms_a * 1e3, "GB_per_s": gbs,
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": measured_traffic("add_65536x64")},
        "stream_ceiling": {"us_per_launch": ms_c * 1e3, "GB_per_s": gbs_c, "frac_of_peak": gbs_c / HBM_PEAK_GBS,
                           "add_frac_of_ceiling": gbs / gbs_c,
                           "what": "r = a + b per u64 limb, no carries, same 4 rotating buffer sets and the same "
                                   "launch shape (4-wave blocks, 2 rows of 128 limbs per wave, 16-byte "
                                   "non-temporal accesses); best of 3 interleaved runs of each"},
        "add6": {"us_per_launch": ms_6 * 1e3, "GB_per_s": gbs_6, "mil_instances_per_s": na / ms_6 * 1e-3,
                 "metric": "GB/s = 3 n m 8 / t, as the reference's SIX Additions line (cuda/main.cu:145-153)"},
        "note": "3*N*m*8 bytes per launch; 4 rotating buffer sets (384 MiB)",
    }


def bench_composites(torch, mi, ctx, args, d2, d3):
    """poly (polyKerQ) on the headline operands and poly_ntt (polyFttKer) on
    the NTT operands; Gu32ops = 4 n numAd32OpsOfMultInst(m) / t as the
    reference reports them (cuda/main.cu:340-346, 560-566)."""
    (da, db, n, m), (dan, dbn, nn, mn) = d2, d3
    r = torch.empty(n * m, dtype=torch.int64, device="cuda:0")
    ms_p = timed(torch, lambda: ctx.poly(da, db, r, m), max(5, args.steps // 4), 3)
    rn = torch.empty(nn * mn, dtype=torch.int64, device="cuda:0")
    ms_pn = timed(torch, lambda: ctx.poly_ntt(dan, dbn, rn, mn), max(3, args.steps // 40), 2)
    out = {}
    for key, ms, nb, mb in (("poly_16384x128", ms_p, n, m), ("poly_ntt_1024x2048", ms_pn, nn, mn)):
        out[key] = {"us_per_batch": ms * 1e3, "polys_per_s": nb / (ms * 1e-3),
                    "Gu32ops_per_s": 4 * nb * gu32ops_of_mul(mb) / (ms * 1e-3) / 1e9,
                    "algorithmic_hbm_bytes": 3 * nb * mb * 8}
        log(f"{key}: {ms * 1e3:.1f} us/batch")
    del r, rn
    return out


def bench_mul_sweep(torch, mi, ctx, args):
    """Classical full products at 256 .. 2048 limbs (2^21 limbs per operand
    batch, random limbs): the tuned kernel of each size and, for comparison,
    the 32 x 32 x 32 matrix-core kernel on the same operands (results checked
    equal)."""
    out = {}
    rng = np.random.default_rng(0x5EE9)
    for m in (256, 512, 1024, 2048):
        n = (1 << 21) // m
        a = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
        b = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
        da, db = to_dev(torch, a), to_dev(torch, b)
        dr = torch.empty(2 * n * m, dtype=torch.int64, device="cuda:0")
        kernel = ("16x16x64 one product per workgroup" if tuned_value(mi, "mul.m16", m, 1) != 0
                  else "32x32x32")
        steps = max(5, args.steps // 10)
        ms = timed(torch, lambda: ctx.mul_full(da, db, dr, m), steps, 3)
        tuned = dr.clone()
        key = f"mul.m16@{m}"
        try:
            old = mi.tune_get(key)
        except Exception:
            old = None
        mi.tune_set(key, 0)
        ms_old = timed(torch, lambda: ctx.mul_full(da, db, dr, m), steps, 3)
        same = bool(torch.equal(dr, tuned))
        mi.tune_set(key, old if old is not None else mi.tune_get("mul.m16"))
        D = 8 * m + 1
        out[f"mul_{n}x{m}"] = {"us_per_batch": ms * 1e3, "muls_per_s": n / (ms * 1e-3), "kernel": kernel,
                               "useful_TOPs": n * 2 * D * D / (ms * 1e-3) / 1e12,
                               "us_per_batch_32x32x32": ms_old * 1e3, "results_equal": same}
        log(f"mul_full {n} x {m}: {ms * 1e3:.1f} us ({kernel}); 32x32x32 kernel {ms_old * 1e3:.1f} us; equal {same}")
        del da, db, dr, tuned
    return out


def bench_ragged(torch, mi, ctx, args, rag):
    """Ragged batch (50000 ops): one persistent hipGraph launch vs plain
    fixed-size launches, one per (op, power-of-two size class) and one per
    exact (op, n_a, n_b), on zero-padded operands."""
    from workloads import size_class

    ops, ra, rb, rlimbs = rag
    dra, drb = to_dev(torch, ra), to_dev(torch, rb)
    drr = torch.zeros(rlimbs, dtype=torch.int64, device="cuda:0")
    n_mul = sum(1 for o in ops if o[0] == 1)
    plan = mi.RaggedPlan(ctx, ops, dra, drb, drr)
    rsteps = max(5, args.steps // 20)
    ms_rag = timed(torch, plan.launch, rsteps, 2)
    pstats = plan.stats()

    def grouped(key):
        groups = {}
        for i, (kind, na, nb, oa, ob, orr) in enumerate(ops):
            groups.setdefault(key(kind, na, nb), []).append(i)
        return groups

    def plain_launches(groups):
        """groups: [(kind, M, op indices)]. One packed padded buffer per
        operand; each group is a batch-major slice of it at its size M."""
        metas, tot_in, tot_out = [], 0, 0
        for kind, M, idx in groups:
            n_c = len(idx)
            metas.append((kind, M, idx, tot_in, tot_out))
            tot_in += n_c * M
            tot_out += n_c * (2 * M if kind == 1 else M)
        pa = np.zeros(tot_in, dtype=np.uint64)
        pb = np.zeros(tot_in, dtype=np.uint64)
        for kind, M, idx, off, _ in metas:
            for j, i in enumerate(idx):
                _, na, nb, oa, ob, _ = ops[i]
                pa[off + j * M: off + j * M + na] = ra[oa:oa + na]
                pb[off + j * M: off + j * M + nb] = rb[ob:ob + nb]
        x, y = to_dev(torch, pa), to_dev(torch, pb)
        z = torch.empty(tot_out, dtype=torch.int64, device="cuda:0")
        c = torch.empty(len(ops), dtype=torch.uint8, device="cuda:0")
        calls = []
        for kind, M, idx, off, oo in metas:
            n_c = len(idx)
            xa, yb = x[off:off + n_c * M], y[off:off + n_c * M]
            if kind == 1:
                calls.append((ctx.mul_full, (xa, yb, z[oo:oo + 2 * n_c * M], M)))
            else:
                calls.append((ctx.add, (xa, yb, z[oo:oo + n_c * M], M, c[:n_c])))

        def step():
            for f, a in calls:
                f(*a)

        return step, len(calls), (x, y, z, c)

    classes = grouped(lambda k, na, nb: (k, size_class(max(na, nb))))
    step_c, launches_c, keep_c = plain_launches([(k, M, idx) for (k, M), idx in sorted(classes.items())])
    ms_plain = timed(torch, step_c, rsteps, 2)
    del keep_c
    # exact sizes: the kernels are square, so each distinct (op, n_a, n_b)
    # is one launch at M = max(n_a, n_b)
    exact = grouped(lambda k, na, nb: (k, na, nb))
    step_e, launches_e, keep_e = plain_launches([(k, max(na, nb), idx) for (k, na, nb), idx in sorted(exact.items())])
    ms_exact = timed(torch, step_e, 2, 1)
    del keep_e
    log(f"ragged {len(ops)} ops: persistent {ms_rag:.3f} ms ({pstats}); plain per size class {ms_plain:.3f} ms "
        f"({launches_c} launches, {ms_plain / ms_rag:.2f}x); plain per exact size {ms_exact:.1f} ms "
        f"({launches_e} launches, {ms_exact / ms_rag:.1f}x)")
    del plan, dra, drb, drr
    return {"ops": len(ops), "muls": n_mul, "persistent_ms": ms_rag, "plain_per_size_ms": ms_plain,
            "speedup": ms_plain / ms_rag, "muls_per_s_persistent": n_mul / (ms_rag * 1e-3),
            "muls_per_s_plain": n_mul / (ms_plain * 1e-3), "plain_launches": launches_c,
            "plain_per_exact_size_ms": ms_exact, "plain_exact_launches": launches_e,
            "speedup_vs_exact": ms_exact / ms_rag, "plan": pstats,
            "note": "one hipGraph launch (the additions and the products whose shorter operand has at most 128 "
                    "limbs one op per wave, the larger products on matrix-core nodes, parallel branches) vs one "
                    "fixed-size launch per (op, power-of-two size class) and vs one launch per distinct "
                    "(op, n_a, n_b), each on operands zero-padded to the launch size; eager launches from Python "
                    "on one stream"}


def bench_serving(mi, ctx):
    """Serving loop: Poisson arrivals at three offered rates and a burst in
    which every request arrives at once (the saturated server)."""
    from workloads import make_requests

    reqs = make_requests(mi.gen_limbs, 3000)
    serving = []
    inf = float("inf")
    for qps, batched in ((1000, False), (16000, False), (64000, False), (inf, False), (64000, True), (inf, True)):
        srv = mi.Server(ctx, hbm_budget=4 << 30, streams=4)
        res = mi.serve_poisson(srv, reqs, qps=qps, seed=int(qps) if qps != inf else 1,
                               keep_results=False, batched=batched)
        st = srv.stats()
        srv.close()
        res.pop("results")
        res["batches"] = st["batches_launched"]
        res["mean_batch_ops"] = st["ops_in_batches"] / max(1, st["batches_launched"])
        res["deferred_admissions"] = st["deferred_admissions"]
        res["client"] = "batched (submit_many / result_many)" if batched else "one call per request"
        serving.append(res)
        log(f"serving {'burst' if qps == inf else f'qps {qps}'}{' batched' if batched else ''}: "
            f"p50 {res['p50_ms']:.2f} ms, p99 {res['p99_ms']:.2f} ms, {res['ops_per_s']:.0f} ops/s, "
            f"{res['batches']} batches, {res['mean_batch_ops']:.1f} ops/batch")
    return {"requests": len(reqs), "mix": "70% mul / 30% add, operands 8..4096 limbs log-uniform",
            "points": serving,
            "note": "offered_qps null = every request submitted at time 0 (saturation); latency = completion "
                    "- scheduled arrival on the host clock; client 'one call per request' drives the server "
                    "with one Python call per submit and per result, 'batched' with one call for the "
                    "requests due at once and one per poll for the completed ones"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=("all", "core"), default="all",
                    help="core: only the fixed-size workloads (mul, NTT, add with its ceiling and add6, poly, "
                         "poly_ntt), so that a profiler run sees each kernel from one workload")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 256 MiB int64 negation (2 x 256 MiB of HBM traffic per launch) x5, "
                         "the reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: every instance is independent, so N GPUs would each run "
              f"this same kernel on 1/N of the batch with no exchange; --gpus {args.gpus} "
              f"is not measured here.", file=sys.stderr)
        sys.exit(2)

    import torch

    from __graft_entry__ import load_package
    from workloads import make_inputs, make_ragged

    mi = load_package()
    mi.lib()
    ctx = mi.Context(0)

    # ---------------------------------------------------------- headline: mul
    n, m = 16384, 128
    a, b = make_inputs("mul_16384x128", mi.gen_limbs)
    da, db = to_dev(torch, a), to_dev(torch, b)
    dr = torch.empty(2 * n * m, dtype=torch.int64, device="cuda:0")
    log(f"mul_full {n} x {m} limbs: warmup {args.warmup}, steps {args.steps}")
    # Untimed lead-in: an idle MI355X needs on the order of 100 ms of load to
    # reach its working clocks (measured: 88 us per batch after 20 warm-up
    # steps, 78 us after 2000). Run the step for 250 ms first, then the W
    # warm-up steps, then the K timed steps.
    t_end = time.perf_counter() + 0.25
    while time.perf_counter() < t_end:
        for _ in range(50):
            ctx.mul_full(da, db, dr, m)
        torch.cuda.synchronize()
    ms = timed(torch, lambda: ctx.mul_full(da, db, dr, m), args.steps, args.warmup)
    muls_per_s = n / (ms * 1e-3)
    pp_per_launch = n * (2 * m) ** 2  # 32x32->64 partial products of the full product
    mad_rate, add_rate = mi.probe_valu(0)
    achieved_pp = pp_per_launch / (ms * 1e-3)
    on_mfma = tuned_value(mi, "mul.mfma", m, 0) != 0 and tuned_value(mi, "mul.digits", m, 0) == 0
    log(f"mul_full: {ms * 1e3:.1f} us/batch, {muls_per_s / 1e6:.2f} M muls/s "
        f"({'int8 matrix-core' if on_mfma else 'VALU'} kernel); probe v_mad_u64_u32 {mad_rate / 1e12:.2f} T/s")
    valu_roofline = {"bound": "valu", "achieved": achieved_pp / 1e12, "peak": mad_rate / 1e12,
                     "unit": "T partial-products/s (32x32->64 v_mad_u64_u32)",
                     "frac": achieved_pp / mad_rate, "traffic": measured_traffic("mul_full_16384x128"),
                     "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                     "peak_source": "midint_probe_valu on this GPU (issue rate of v_mad_u64_u32)"}
    if on_mfma:
        wave16 = m in (64, 128) and tuned_value(mi, "mul.wave16", m, 1) != 0
        useful, issued = wave16_work(m) if wave16 else mfma_work(m, tuned_value(mi, "mul.mfma_wpi", m, 0))
        tops = n * useful / (ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": tops, "peak": INT8_PEAK_TOPS, "unit": "TOP/s",
                    "frac": tops / INT8_PEAK_TOPS, "traffic": measured_traffic("mul_full_16384x128"),
                    "work": f"int8 ops per launch = {n} x 2 x {8 * m + 1}^2 (balanced base-256 digit products "
                            f"of the {64 * m}-bit operands); "
                            f"{'v_mfma_i32_16x16x64_i8, one product per wave' if wave16 else 'v_mfma_i32_32x32x32_i8'}; "
                            f"MFMA-issued ops incl. tile padding: "
                            f"{n * issued / (ms * 1e-3) / 1e12:.0f} TOP/s ({issued / useful:.2f}x useful)",
                    "traffic_source": "profiles/traffic.json (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated)",
                    "peak_source": "MI355X published dense int8 MFMA rate"}
        burst, sustained = mi.probe_mfma(0)
        issued_tops = n * issued / (ms * 1e-3) / 1e12
        roofline["mfma_ceiling"] = {
            "burst_TOPs": burst / 1e12, "sustained_TOPs": sustained / 1e12,
            "issued_TOPs": issued_tops, "issued_frac_of_sustained": issued_tops * 1e12 / sustained,
            "what": "midint_probe_mfma: dense v_mfma_i32_16x16x64_i8 chains with changing random operands on "
                    "every SIMD, no memory traffic (burst: short launches after idle; sustained: after 0.2 s "
                    "of back-to-back load)"}
        log(f"mfma ceiling (probe): burst {burst / 1e12:.0f} TOP/s, sustained {sustained / 1e12:.0f} TOP/s; "
            f"the kernel issues {issued_tops:.0f} TOP/s ({issued_tops * 1e12 / sustained:.2f} of sustained)")
    else:
        roofline = valu_roofline

    # ---------------------------------------------------------- NTT
    nn, mn = 1024, 2048
    an, bn = make_inputs("ntt_1024x2048", mi.gen_limbs)
    dan, dbn = to_dev(torch, an), to_dev(torch, bn)
    drn = torch.empty(2 * nn * mn, dtype=torch.int64, device="cuda:0")
    ctx.ntt_prepare(mn)
    ms_ntt = timed(torch, lambda: ctx.mul_ntt(dan, dbn, drn, mn), max(5, args.steps // 20), 2)
    ntt_pp = nn * 360448 * 4  # radix-2-equivalent modmuls per product (3 transforms + pointwise), 4 pp each
    log(f"mul_ntt {nn} x {mn}: {ms_ntt * 1e3:.1f} us/batch ({ms_ntt * 1e3 / nn:.3f} us per product)")
    # the NTT's own bound: VALU instruction issue. wave64 instructions per
    # product from the committed profile, the issue ceiling from the probe
    # (probe_add32 counts two ops per instruction and lane)
    ntt_valu = profiled_valu_per_product("mul_ntt_1024x2048")
    ntt_issue = None
    if ntt_valu:
        achieved_wi = ntt_valu * nn / (ms_ntt * 1e-3)
        peak_wi = add_rate / 2 / 64
        ntt_issue = {"bound": "valu-issue", "achieved": achieved_wi / 1e9, "peak": peak_wi / 1e9,
                     "unit": "G wave64 VALU instructions/s", "frac": achieved_wi / peak_wi,
                     "instr_per_product": ntt_valu,
                     "source": "instructions per product: profiles/traffic.json (rocprofv3 SQ_INSTS_VALU); peak: "
                               "midint_probe_valu's 32-bit VALU issue rate on this GPU"}
    del drn

    # ---------------------------------------------------------- add (HBM)
    add = bench_add(torch, mi, ctx, args)

    composites = ragged = serving = None
    rag = make_ragged(mi.gen_limbs) if (args.only == "all" or not args.no_cpu) else None
    composites = bench_composites(torch, mi, ctx, args, (da, db, n, m), (dan, dbn, nn, mn))
    mul_sweep = bench_mul_sweep(torch, mi, ctx, args)
    if args.only == "all":
        ragged = bench_ragged(torch, mi, ctx, args, rag)
        serving = bench_serving(mi, ctx)
    del dan, dbn

    if args.calibrate:
        src = torch.ones(32 * MIB, dtype=torch.int64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    # ---------------------------------------------------------- CPU baselines
    cpu, cpu_line = None, None
    if not args.no_cpu:
        import oracle

        aa, ba = make_inputs("add_65536x64", mi.gen_limbs)
        threads = cpu_threads()
        cpu = cpu_baselines(oracle, {"add": (aa, ba), "mul": (a, b), "ntt": (an, bn), "ragged": rag}, threads)
        if cpu.get("all"):
            cpu_line = {"value": cpu["all"]["mul_16384x128_muls_per_s"], "unit": "muls/s", "cores": threads,
                        "kind": "reference",
                        "sample": "2048 of the 16384 pairs (128 x 128 limbs, full product) with libgmp mpn_mul "
                                  "(the library the reference's CPU path calls, cuda/helper.h:146-170), repeated "
                                  f"for ~1.2 s on {threads} threads (OpenMP over pairs)",
                        "one_core": cpu["one_core"]["mul_16384x128_muls_per_s"],
                        "host_logical_cpus": os.cpu_count(),
                        "port_muls_per_s": cpu["port_mul_16384x128_muls_per_s"]}
        log(f"cpu baselines: {json.dumps(cpu)}")
    out = {
        "metric": "big-int muls/s at 8192-bit",
        "value": muls_per_s,
        "unit": "muls/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "dtype_note": ("exact u64 limb arithmetic; the products are computed through int8 matrix-core (MFMA) "
                       "digit products of balanced base-256 digits, every result bit-exact against GMP")
        if on_mfma else "exact u64 limb arithmetic on the VALU",
        "data": "synthetic (xoshiro256** limbs, seed 0x6D6964696E74 + 1)",
        "config": {"workload": "mul_16384x128: 16384 pairs of 8192-bit ints, full 256-limb products, "
                               + ("classical kernel on the int8 matrix cores (balanced 8-bit digits)" if on_mfma
                                  else "classical kernel (u32 words, VALU)"), "instances": n, "limbs": m},
        "roofline": roofline,
        "cpu_baseline": cpu_line,
        "extra": {
            "add_65536x64": add,
            "ntt_1024x2048": {"us_per_batch": ms_ntt * 1e3, "us_per_product": ms_ntt * 1e3 / nn,
                              "muls_per_s": nn / (ms_ntt * 1e-3),
                              "roofline": {"bound": "valu", "achieved": ntt_pp / (ms_ntt * 1e-3) / 1e12,
                                           "peak": mad_rate / 1e12,
                                           "unit": "T partial-products/s (360448 radix-2-equivalent modmuls x 4 "
                                                   "per product)",
                                           "frac": ntt_pp / (ms_ntt * 1e-3) / mad_rate,
                                           "traffic": measured_traffic("mul_ntt_1024x2048")},
                              "valu_issue": ntt_issue},
            "mul_valu_equivalent": valu_roofline,
            "composites": composites,
            "mul_sweep": mul_sweep,
            "ragged_50000": ragged,
            "probe": {"mad_u64_u32_per_s": mad_rate, "valu32_ops_per_s": add_rate},
            "serving": serving,
            "cpu": cpu,
        },
    }
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
