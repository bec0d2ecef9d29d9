"""Autotuner of the midint launch parameters (run on the target GPU).

For each power-of-two operand-size class it times every candidate launch
configuration of a kernel on a synthetic batch of that class and keeps the
fastest. A candidate is only timed after its output is bit-identical to the
output of the library's default configuration on the same inputs, so tuning
can never trade correctness for speed. The result is written as
``key@class value`` lines (see midint_tune_set) to tune/mi355x.txt, which
Context() applies when it is created (and C callers with midint_tune_load).

    python midint-arithmetic_amd/autotune.py [--out PATH] [--quick] [--only add,mul,ntt,ragged]

--only re-tunes the listed parts and keeps the other parts' lines of an
existing tuning file.
"""
from __future__ import annotations

import argparse
import importlib.util
import itertools
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _pkg():
    name = "midint_arithmetic_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _timed(torch, fn, steps, warm=3):
    for _ in range(warm):
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


def _set(mi, params, cls):
    for k, v in params.items():
        mi.tune_set(f"{k}@{cls}", v)


def _get(mi, key):
    """Value of a tuning key; an unset "key@C" reads as the global key."""
    try:
        return mi.tune_get(key)
    except Exception:
        return mi.tune_get(key.split("@")[0])


def grid_of(grid):
    """Every combination of a {key: [values]} grid, as a list of dicts."""
    return [dict(zip(grid.keys(), v)) for v in itertools.product(*grid.values())]


def tune_op(mi, torch, ctx, op, cls, cands, defaults, n, rng, log, sets=1, key_cls=None):
    """Times every configuration in `cands` (a list of {key: value} dicts) for
    operand size `cls` limbs on n instances of `op` ("add", "mul" or "ntt"),
    cycling through `sets` buffer sets so that additions stream from HBM
    rather than the Infinity Cache; returns the fastest params. The keys are
    set at class `key_cls` (default `cls`)."""
    m = cls
    key_cls = cls if key_cls is None else key_cls
    bufs = []
    for _ in range(sets):
        a = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
        b = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
        a[: m] = np.uint64(2**64 - 1)
        b[: m] = np.uint64(2**64 - 1)
        out_l = n * m if op == "add" else 2 * n * m
        bufs.append((torch.from_numpy(a.view(np.int64)).to("cuda:0"), torch.from_numpy(b.view(np.int64)).to("cuda:0"),
                     torch.empty(out_l, dtype=torch.int64, device="cuda:0")))
    it = [0]

    def run():
        da, db, dr = bufs[it[0] % sets]
        it[0] += 1
        if op == "mul":
            ctx.mul_full(da, db, dr, m)
        elif op == "ntt":
            ctx.mul_ntt(da, db, dr, m)
        else:
            ctx.add(da, db, dr, m)

    _set(mi, defaults, key_cls)
    it[0] = 0
    run()
    torch.cuda.synchronize()
    ref = bufs[0][2].clone()
    valid = []
    for params in cands:
        _set(mi, params, key_cls)
        bufs[0][2].zero_()
        it[0] = 0
        run()
        torch.cuda.synchronize()
        if not torch.equal(bufs[0][2], ref):
            log(f"  {op}@{cls} {params}: output differs from the default configuration -- rejected")
            continue
        valid.append(params)
    # two interleaved rounds over the candidates, the faster time of each: a
    # candidate timed while the clocks settle is not left behind by it
    times = [None] * len(valid)
    for _ in range(2):
        for i, params in enumerate(valid):
            _set(mi, params, key_cls)
            us = _timed(torch, run, 40 if op == "add" else 30) * 1e3
            times[i] = us if times[i] is None else min(times[i], us)
    rows = list(zip(valid, times))
    best, best_us = dict(defaults), None
    for params, us in rows:
        if best_us is None or us < best_us:
            best, best_us = params, us
    _set(mi, best, key_cls)
    log(f"  {op}@{key_cls}: best {best} {best_us:.2f} us over {len(rows)} candidates (n = {n})")
    return best, rows


def tune_ragged(mi, torch, ctx, log, classes=(1024, 512, 256, 128, 64, 32, 16, 8, 4)):
    """Ragged batch launch parameters on the 50000-op batch. First the
    global keys: one op per wave for the additions and the products with a
    short operand (ragged.wave = 1; their products on one kernel or two,
    ragged.wave_split), or the persistent kernels with a
    matrix-core threshold (shorter operand limbs) and whether the graph
    branches overlap (ragged.wave = 0); then, size class by size class
    (coordinate descent, largest class first), the persistent kernel's
    threads and columns per lane group (ragged.threads@C, ragged.q@C) and
    the matrix-core node's tiles (ragged.m16@C: 32 x 32 x 32 or
    16 x 16 x 64) and waves per workgroup (ragged.mfma_waves@C,
    ragged.m16_waves@C). Every candidate's output must equal the first
    plan's before it is timed, and a class setting is kept only if the batch
    gets at least 1 % faster. Returns the tuning-file lines."""
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
    from workloads import make_ragged

    ops, a, b, rl = make_ragged(mi.gen_limbs)
    da = torch.from_numpy(a.view(np.int64)).to("cuda:0")
    db = torch.from_numpy(b.view(np.int64)).to("cuda:0")
    state = {"ref": None}

    def trial(cand):
        for k, v in cand.items():
            mi.tune_set(k, v)
        dr = torch.zeros(rl, dtype=torch.int64, device="cuda:0")
        plan = mi.RaggedPlan(ctx, ops, da, db, dr)
        plan.launch()
        torch.cuda.synchronize()
        if state["ref"] is None:
            state["ref"] = dr.clone()
        elif not torch.equal(dr, state["ref"]):
            log(f"  ragged {cand}: output differs -- rejected")
            plan.close()
            return None
        ms = _timed(torch, plan.launch, 30)
        plan.close()
        return ms

    best, best_ms = {}, None
    globals_ = [{"ragged.wave": 1, "ragged.wave_split": sp, "ragged.mfma_min": 32, "ragged.overlap": 1}
                for sp in (0, 1)]
    globals_ += [{"ragged.wave": 0, "ragged.wave_split": 0, "ragged.mfma_min": mf_min, "ragged.overlap": ov}
                 for mf_min in (0, 32, 64, 128) for ov in (0, 1) if not (mf_min == 0 and ov == 1)]
    for cand in globals_:
        ms = trial(cand)
        if ms is None:
            continue
        log(f"  ragged {cand}: {ms:.3f} ms")
        if best_ms is None or ms < best_ms:
            best, best_ms = cand, ms
    for k, v in best.items():
        mi.tune_set(k, v)
    lines = [f"{k} {v}" for k, v in best.items()]

    # classes present in the batch, by where their ops go
    wave = best["ragged.wave"] == 1
    mf_min = best["ragged.mfma_min"]
    has_rg, has_mf = set(), set()
    for kind, na, nb, *_ in ops:
        c = 1
        while c < max(na, nb):
            c *= 2
        if wave:
            if kind == 1 and min(na, nb) > 128:
                has_mf.add(c)
        elif kind == 1 and mf_min > 0 and min(na, nb) >= mf_min:
            has_mf.add(c)
        else:
            has_rg.add(c)
    for c in classes:
        cands = []
        if c in has_rg:
            cands += [{f"ragged.q@{c}": q, f"ragged.threads@{c}": th} for q in (4, 8) for th in (256, 512)]
        if c in has_mf:
            cands += [{f"ragged.m16@{c}": 0, f"ragged.mfma_waves@{c}": w} for w in (1, 2, 4, 8)]
            cands += [{f"ragged.m16@{c}": 1, f"ragged.m16_waves@{c}": w} for w in (2, 4, 8)]
        for cand in cands:
            keep = {k: _get(mi, k) for k in cand}
            ms = trial(cand)
            if ms is not None and ms < 0.99 * best_ms:  # at least 1 %: below that is timing noise
                best_ms = ms
                log(f"  ragged {cand}: {ms:.3f} ms (kept)")
            else:
                for k, v in keep.items():
                    mi.tune_set(k, v)
        lines += [f"{k} {_get(mi, k)}" for k in sorted({k for cand in cands for k in cand})]
    log(f"  ragged: best {best_ms:.3f} ms")
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune", "mi355x.txt"))
    ap.add_argument("--quick", action="store_true", help="fewer classes and candidates")
    ap.add_argument("--only", default="add,mul,ntt,ragged", help="comma-separated parts to tune")
    ap.add_argument("--base", default=os.path.join(HERE, "tune", "mi355x.txt"),
                    help="with --only: the table whose other parts are kept")
    args = ap.parse_args(argv)
    parts = set(args.only.split(","))

    import torch

    mi = _pkg()
    ctx = mi.Context(0, tuned=False)
    rng = np.random.default_rng(7)
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    lines = ["# midint launch parameters tuned on " + torch.cuda.get_device_name(0),
             "# written by midint-arithmetic_amd/autotune.py " + time.strftime("%Y-%m-%d %H:%M:%S")]

    # an idle GPU reaches its working clocks only after some 100 ms of load
    lead = torch.zeros(1 << 24, dtype=torch.int64, device="cuda:0")
    t_end = time.perf_counter() + 0.3
    while time.perf_counter() < t_end:
        ctx.mul_full(lead, lead, torch.empty(2 << 24, dtype=torch.int64, device="cuda:0"), 1024)
        torch.cuda.synchronize()
    del lead

    if parts != {"add", "mul", "ntt", "ragged"} and os.path.exists(args.base):
        # keep the parts not re-tuned (and apply them: they shape the others' timing)
        with open(args.base) as f:
            for line in f:
                kv = line.split()
                if len(kv) == 2 and not line.startswith("#") and kv[0].split(".")[0] not in parts:
                    lines.append(line.strip())
                    mi.tune_set(kv[0], int(kv[1]))

    add_classes = [] if "add" not in parts else [16, 64, 256] if args.quick else [1, 2, 4, 8, 16, 32, 64, 128, 256]
    add_grid = {"add.rows": [1, 2, 4], "add.threads": [64, 128, 256, 512], "add.nt": [0, 1]}
    for cls in add_classes:
        # 32 MiB per operand, 4 rotating sets (384 MiB > the 256 MiB Infinity Cache)
        best, _ = tune_op(mi, torch, ctx, "add", cls, grid_of(add_grid), {"add.rows": 2, "add.threads": 256, "add.nt": 1},
                          (1 << 22) // cls, rng, log, sets=4)
        lines += [f"{k}@{cls} {v}" for k, v in best.items()]

    mul_classes = [] if "mul" not in parts else [64, 128, 512] if args.quick else [4, 8, 16, 32, 64, 128, 256, 512,
                                                                                     1024, 2048]
    # word kernel: Q x register park x prefetch x block size; digit kernel:
    # register park x prefetch; matrix-core kernel: waves per instance
    mul_cands = grid_of({"mul.mfma": [0], "mul.digits": [0], "mul.q": [8, 16], "mul.rpark": [0, 1],
                         "mul.prefetch": [0, 1], "mul.threads": [128, 256]})
    mul_cands += grid_of({"mul.mfma": [0], "mul.digits": [0], "mul.q": [4], "mul.rpark": [0], "mul.prefetch": [0, 1],
                          "mul.threads": [256]})
    mul_cands += grid_of({"mul.mfma": [0], "mul.digits": [1], "mul.q": [8], "mul.rpark": [0, 1],
                          "mul.prefetch": [0, 1], "mul.threads": [256]})
    # int8 matrix-core kernels (every candidate sets every key of the three,
    # so none inherits a previous candidate's class key): 32 x 32 x 32 tiles
    # with their waves per instance, 16 x 16 x 64 tiles one product per
    # workgroup with their waves per product, one product per wave (64 and
    # 128 limbs; elsewhere it falls through to the 16 x 16 x 64 kernel);
    # 0 = the kernel's own choice
    mx = {"mul.mfma": [1], "mul.digits": [0], "mul.q": [8], "mul.rpark": [1], "mul.prefetch": [0],
          "mul.threads": [256]}
    mul_cands += grid_of({**mx, "mul.wave16": [0], "mul.m16": [0], "mul.mfma_wpi": [0, 1, 2, 4, 8, 16],
                          "mul.m16_waves": [0]})
    mul_cands += grid_of({**mx, "mul.wave16": [0], "mul.m16": [1], "mul.mfma_wpi": [0],
                          "mul.m16_waves": [0, 1, 2, 4, 8, 16]})
    mul_cands += grid_of({**mx, "mul.wave16": [1], "mul.m16": [1], "mul.mfma_wpi": [0], "mul.m16_waves": [0]})
    mul_default = {"mul.mfma": 0, "mul.mfma_wpi": 0, "mul.digits": 0, "mul.q": 8, "mul.rpark": 1, "mul.prefetch": 0,
                   "mul.threads": 256, "mul.wave16": 1, "mul.m16": 1, "mul.m16_waves": 0}
    for cls in mul_classes:
        # about 2^30 partial products per launch, never fewer than 512 instances
        n = min(1 << 20, max(512, (1 << 30) // (2 * cls) ** 2))
        best, _ = tune_op(mi, torch, ctx, "mul", cls, mul_cands, mul_default, n, rng, log)
        lines += [f"{k}@{cls} {v}" for k, v in best.items()]

    # NTT values per thread, keyed by transform length L = 8m (2^11..2^14)
    ntt_classes = [] if "ntt" not in parts else [2048] if args.quick else [256, 512, 1024, 2048]
    for cls in ntt_classes:
        L = 8 * cls
        n = max(256, (1 << 21) // cls)
        best, _ = tune_op(mi, torch, ctx, "ntt", cls, grid_of({"ntt.vals": [16, 32]}), {"ntt.vals": 0}, n, rng, log,
                          key_cls=L)
        lines += [f"{k}@{L} {v}" for k, v in best.items()]

    if "ragged" in parts:
        lines += tune_ragged(mi, torch, ctx, log)

    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        f.write("\n".join(lines) + "\n")
    log(f"wrote {args.out}")
    ctx.close()


if __name__ == "__main__":
    main()
