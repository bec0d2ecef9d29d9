"""Ragged batches in one persistent graph launch: every op's result equals
the oracle's, alone and in the mix, and the plain per-size-class launches
(operands zero-padded to the next power of two, the reference's way of
serving mixed sizes) agree with it."""
import numpy as np
import pytest

from workloads import make_ragged, ragged_reference, size_class

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def ctx(mi, torch_dev):
    c = mi.Context(0)
    yield c
    c.close()


def dev(torch, x):
    return torch.from_numpy(np.ascontiguousarray(x).view(np.int64)).to("cuda:0")


def run_plan(mi, ctx, torch, ops, a, b, rl, threads=0, reps=1):
    da, db = dev(torch, a), dev(torch, b)
    dr = torch.zeros(max(rl, 1), dtype=torch.int64, device="cuda:0")
    plan = mi.RaggedPlan(ctx, ops, da, db, dr, threads=threads)
    for _ in range(reps):
        plan.launch()
    torch.cuda.synchronize()
    out = dr.cpu().numpy().view(np.uint64)[:rl]
    plan.close()
    return out


@pytest.mark.parametrize("wave", [0, 1], ids=["persistent", "by_wave"])
@pytest.mark.parametrize("threads", [256, 512])
def test_ragged_small_vs_oracle(mi, ctx, torch_dev, orc, threads, wave):
    ops, a, b, rl = make_ragged(mi.gen_limbs, n_ops=600, seed=11, lo_bits=6, hi_bits=14)
    saved = _set_keys(mi, {"ragged.wave": wave})
    try:
        got = run_plan(mi, ctx, torch_dev, ops, a, b, rl, threads=threads, reps=2)
    finally:
        _set_keys(mi, saved)
    want = ragged_reference(orc, ops, a, b, rl)
    assert np.array_equal(got, want), int(np.flatnonzero(got != want)[0])


def test_ragged_edge_sizes(mi, ctx, torch_dev, orc):
    MASK = np.uint64(0xFFFFFFFFFFFFFFFF)
    shapes = [(1, 1), (1, 1024), (1024, 1), (1024, 1024), (3, 5), (2, 2), (1023, 1021), (4, 4), (511, 513)]
    ops, a_parts, b_parts = [], [], []
    oa = ob = orr = 0
    rng = np.random.default_rng(3)
    for i, (na, nb) in enumerate(shapes * 2):
        kind = i % 2 if i >= len(shapes) else 1
        x = np.full(na, MASK, dtype=np.uint64) if i % 3 == 0 else rng.integers(0, 2**64, na, dtype=np.uint64)
        y = np.full(nb, MASK, dtype=np.uint64) if i % 3 == 0 else rng.integers(0, 2**64, nb, dtype=np.uint64)
        rlen = na + nb if kind == 1 else max(na, nb) + 1
        ops.append((kind, na, nb, oa, ob, orr))
        a_parts.append(x)
        b_parts.append(y)
        oa += na
        ob += nb
        orr += rlen
    a, b = np.concatenate(a_parts), np.concatenate(b_parts)
    got = run_plan(mi, ctx, torch_dev, ops, a, b, orr)
    assert np.array_equal(got, ragged_reference(orc, ops, a, b, orr))


def test_ragged_alone_equals_mixed(mi, ctx, torch_dev, orc):
    ops, a, b, rl = make_ragged(mi.gen_limbs, n_ops=300, seed=21, lo_bits=6, hi_bits=13)
    mixed = run_plan(mi, ctx, torch_dev, ops, a, b, rl)
    for i in (0, 7, 150, 299):
        kind, na, nb, oa, ob, orr = ops[i]
        rlen = na + nb if kind == 1 else max(na, nb) + 1
        alone = run_plan(mi, ctx, torch_dev, [(kind, na, nb, 0, 0, 0)], a[oa:oa + na], b[ob:ob + nb], rlen)
        assert np.array_equal(alone, mixed[orr:orr + rlen])


def test_ragged_full_workload(mi, ctx, torch_dev, orc):
    """The 50000-op batch: persistent launch vs oracle, and the plain
    per-size-class launches vs the same oracle (bit-exact)."""
    torch = torch_dev
    ops, a, b, rl = make_ragged(mi.gen_limbs)
    got = run_plan(mi, ctx, torch, ops, a, b, rl, reps=2)
    want = ragged_reference(orc, ops, a, b, rl)
    assert np.array_equal(got, want)
    # plain per-size launches (mul_full at class M, add at class M + carry)
    classes = {}
    for i, (kind, na, nb, oa, ob, orr) in enumerate(ops):
        classes.setdefault((kind, size_class(max(na, nb))), []).append(i)
    for (kind, M), idx in sorted(classes.items()):
        n = len(idx)
        pa = np.zeros((n, M), dtype=np.uint64)
        pb = np.zeros((n, M), dtype=np.uint64)
        for j, i in enumerate(idx):
            _, na, nb, oa, ob, _ = ops[i]
            pa[j, :na] = a[oa:oa + na]
            pb[j, :nb] = b[ob:ob + nb]
        da, db = dev(torch, pa), dev(torch, pb)
        if kind == 1:
            dr = torch.empty(2 * n * M, dtype=torch.int64, device="cuda:0")
            ctx.mul_full(da, db, dr, M)
            res = dr.cpu().numpy().view(np.uint64).reshape(n, 2 * M)
        else:
            dr = torch.empty(n * M, dtype=torch.int64, device="cuda:0")
            dc = torch.empty(n, dtype=torch.uint8, device="cuda:0")
            ctx.add(da, db, dr, M, carry=dc)
            res = np.concatenate([dr.cpu().numpy().view(np.uint64).reshape(n, M),
                                  dc.cpu().numpy().astype(np.uint64)[:, None]], axis=1)
        for j, i in enumerate(idx):
            _, na, nb, _, _, orr = ops[i]
            rlen = na + nb if kind == 1 else max(na, nb) + 1
            if kind == 0 and rlen - 1 < M:
                # carry sits at limb max(na,nb) of the M-limb sum, or in the carry byte
                assert np.array_equal(res[j, :rlen], want[orr:orr + rlen])
            else:
                assert np.array_equal(res[j, :rlen], want[orr:orr + rlen])


@pytest.mark.parametrize("m16", [0, 1, 2], ids=["tiles32", "tiles16", "balanced16"])
@pytest.mark.parametrize("mfma_min", [0, 1, 16])
def test_ragged_mfma_threshold(mi, ctx, torch_dev, orc, mfma_min, m16):
    """Multiplications whose shorter operand has at least ragged.mfma_min
    limbs run on the matrix-core graph node (one product per workgroup, the
    longer operand as the block operand), the rest on the persistent
    kernel: 0 sends none there, 1 sends every one (1 x 1 and 1 x 1024
    products included). The node on 32 x 32 x 32 tiles, on 16 x 16 x 64
    tiles (ragged.m16 = 1, the shorter operand as the windows), and split
    between the two (ragged.m16 = 2: products whose shorter operand has at
    least half the longer one's limbs on 16 x 16 x 64 tiles). One op per
    wave is off here."""
    ops, a, b, rl = make_ragged(mi.gen_limbs, n_ops=400, seed=31, lo_bits=6, hi_bits=16)
    saved = _set_keys(mi, {"ragged.mfma_min": mfma_min, "ragged.m16": m16, "ragged.wave": 0})
    try:
        got = run_plan(mi, ctx, torch_dev, ops, a, b, rl, reps=2)
    finally:
        _set_keys(mi, saved)
    want = ragged_reference(orc, ops, a, b, rl)
    assert np.array_equal(got, want), int(np.flatnonzero(got != want)[0])


RAGGED_CLASSES = [4, 8, 16, 32, 64, 128, 256, 512, 1024]


def _set_keys(mi, cfg):
    """Sets tuning keys (global or key@C); returns what restores them: a
    class key that was unset gets the global value (the same lookup)."""
    saved = {}
    for k, v in cfg.items():
        try:
            saved[k] = mi.tune_get(k)
        except Exception:
            saved[k] = mi.tune_get(k.split("@")[0])
        mi.tune_set(k, v)
    return saved


@pytest.mark.parametrize("seed", range(12))
def test_ragged_config_sweep(mi, ctx, torch_dev, orc, seed):
    """Random batches (op count, size range, add/mul mix) under random
    settings of every ragged tuning key: one op per wave for the additions
    and the products with a short operand or not, the matrix-core threshold, whether
    the persistent kernels run beside the matrix-core nodes or after them,
    and per size class the persistent kernel's threads and columns per lane
    group (Q) and the matrix-core node's waves per workgroup -- or one thread
    count for every class through the plan's own tuning argument, and which
    matrix-core products run on 16 x 16 x 64 tiles with how many waves. Two
    replays of the graph, each equal to the oracle."""
    rng = np.random.default_rng(0xA11 + seed)
    lo = int(rng.integers(6, 11))
    hi = int(rng.integers(lo + 1, 17))
    ops, a, b, rl = make_ragged(mi.gen_limbs, n_ops=int(rng.integers(50, 500)), seed=100 + seed,
                                mul_frac=float(rng.choice([0.3, 0.7, 1.0])), lo_bits=lo, hi_bits=hi)
    cfg = {"ragged.mfma_min": int(rng.choice([0, 1, 8, 32, 64])), "ragged.overlap": int(rng.integers(0, 2)),
           "ragged.wave": int(rng.integers(0, 2)), "ragged.wave_split": int(rng.integers(0, 2))}
    for c in RAGGED_CLASSES:
        cfg[f"ragged.q@{c}"] = int(rng.choice([4, 8]))
        cfg[f"ragged.threads@{c}"] = int(rng.choice([256, 512]))
        cfg[f"ragged.mfma_waves@{c}"] = int(rng.choice([1, 2, 4, 8]))
        cfg[f"ragged.m16@{c}"] = int(rng.choice([0, 1, 2]))
        cfg[f"ragged.m16_ratio@{c}"] = int(rng.choice([1, 2, 4]))
        cfg[f"ragged.m16_waves@{c}"] = int(rng.choice([1, 2, 4, 8]))
    threads = int(rng.choice([0, 256, 512]))
    saved = _set_keys(mi, cfg)
    try:
        got = run_plan(mi, ctx, torch_dev, ops, a, b, rl, threads=threads, reps=2)
    finally:
        _set_keys(mi, saved)
    want = ragged_reference(orc, ops, a, b, rl)
    assert np.array_equal(got, want), (cfg, threads, int(np.flatnonzero(got != want)[0]))


@pytest.mark.parametrize("q,threads,waves", [(8, 256, 1), (8, 512, 2), (4, 512, 8), (8, 256, 8)])
def test_ragged_uniform_class_keys(mi, ctx, torch_dev, orc, q, threads, waves):
    """Every class on one (Q, threads, waves) setting, with the matrix-core
    threshold off (every product on the persistent kernel) and on (one op per
    wave off)."""
    ops, a, b, rl = make_ragged(mi.gen_limbs, n_ops=800, seed=31, lo_bits=6, hi_bits=16)
    want = ragged_reference(orc, ops, a, b, rl)
    for mf_min in (0, 32):
        cfg = {"ragged.mfma_min": mf_min, "ragged.wave": 0}
        for c in RAGGED_CLASSES:
            cfg[f"ragged.q@{c}"] = q
            cfg[f"ragged.threads@{c}"] = threads
            cfg[f"ragged.mfma_waves@{c}"] = waves
        saved = _set_keys(mi, cfg)
        try:
            got = run_plan(mi, ctx, torch_dev, ops, a, b, rl, reps=2)
        finally:
            _set_keys(mi, saved)
        assert np.array_equal(got, want), (q, threads, waves, mf_min, int(np.flatnonzero(got != want)[0]))


@pytest.mark.parametrize("m16", [0, 1], ids=["tiles32", "tiles16"])
def test_ragged_mfma_edge_shapes(mi, ctx, torch_dev, orc, m16):
    """Every multiplication on the matrix-core node (32 x 32 x 32 or
    16 x 16 x 64 tiles): unbalanced and tiny shapes, all-ones and
    0x7F-byte operands (the longest digit carry chain)."""
    K7F = np.uint64(0x7F7F7F7F7F7F7F7F)
    MASK = np.uint64(0xFFFFFFFFFFFFFFFF)
    shapes = [(1, 1), (1, 1024), (1024, 1), (1024, 1024), (3, 5), (1023, 1021), (4, 4), (511, 513), (128, 129),
              (1000, 7), (64, 900)]
    ops, a_parts, b_parts = [], [], []
    oa = ob = orr = 0
    rng = np.random.default_rng(5)
    for i, (na, nb) in enumerate(shapes * 3):
        fill = (MASK, K7F, None)[i % 3]
        x = np.full(na, fill, dtype=np.uint64) if fill is not None else rng.integers(0, 2**64, na, dtype=np.uint64)
        y = np.full(nb, fill, dtype=np.uint64) if fill is not None else rng.integers(0, 2**64, nb, dtype=np.uint64)
        ops.append((1, na, nb, oa, ob, orr))
        a_parts.append(x)
        b_parts.append(y)
        oa += na
        ob += nb
        orr += na + nb
    a, b = np.concatenate(a_parts), np.concatenate(b_parts)
    saved = _set_keys(mi, {"ragged.mfma_min": 1, "ragged.m16": m16, "ragged.wave": 0})
    try:
        got = run_plan(mi, ctx, torch_dev, ops, a, b, orr)
    finally:
        _set_keys(mi, saved)
    assert np.array_equal(got, ragged_reference(orc, ops, a, b, orr))


def _packed(shapes, kinds, fills, seed):
    rng = np.random.default_rng(seed)
    ops, a_parts, b_parts = [], [], []
    oa = ob = orr = 0
    for (na, nb), kind, fill in zip(shapes, kinds, fills):
        x = np.full(na, fill, dtype=np.uint64) if fill is not None else rng.integers(0, 2**64, na, dtype=np.uint64)
        y = np.full(nb, fill, dtype=np.uint64) if fill is not None else rng.integers(0, 2**64, nb, dtype=np.uint64)
        ops.append((kind, na, nb, oa, ob, orr))
        a_parts.append(x)
        b_parts.append(y)
        oa += na
        ob += nb
        orr += na + nb if kind == 1 else max(na, nb) + 1
    return ops, np.concatenate(a_parts), np.concatenate(b_parts), orr


def test_ragged_wave_every_short_size(mi, ctx, torch_dev, orc):
    """One product per wave (ragged.wave = 1): every shorter-operand size
    1..128 limbs (every fragment count K = 1..17), as either operand, against
    longer operands of odd, even, row-multiple and maximal length (the rows
    of a staged through the ring, the carry digit in the next row), with
    all-ones, 0x7F-byte and random limbs; and additions of the same shapes.
    The products on one kernel and split by fragment count
    (ragged.wave_split)."""
    K7F = np.uint64(0x7F7F7F7F7F7F7F7F)
    MASK = np.uint64(0xFFFFFFFFFFFFFFFF)
    longs = [1, 2, 127, 128, 129, 255, 256, 383, 512, 640, 1000, 1023, 1024]
    shapes, kinds, fills = [], [], []
    for ns in range(1, 129):
        nl = max(ns, longs[ns % len(longs)])
        shp = (nl, ns) if ns % 2 else (ns, nl)
        for kind in (1, 1, 0):
            shapes.append(shp)
            kinds.append(kind)
            fills.append((MASK, K7F, None)[len(fills) % 3])
    for nl in longs:
        for ns in (1, 64, 127, 128):
            if ns <= nl:
                for kind in (1, 0):
                    shapes.append((nl, ns))
                    kinds.append(kind)
                    fills.append((None, MASK)[len(fills) % 2])
    ops, a, b, rl = _packed(shapes, kinds, fills, 17)
    want = ragged_reference(orc, ops, a, b, rl)
    for split in (0, 1):
        saved = _set_keys(mi, {"ragged.wave": 1, "ragged.wave_split": split})
        try:
            got = run_plan(mi, ctx, torch_dev, ops, a, b, rl, reps=2)
        finally:
            _set_keys(mi, saved)
        if not np.array_equal(got, want):
            bad = int(np.flatnonzero(got != want)[0])
            op = max(i for i, o in enumerate(ops) if o[5] <= bad)
            raise AssertionError(f"split {split}: first mismatch at limb {bad}, op {ops[op][:3]}")
