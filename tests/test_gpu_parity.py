"""Parity of the HIP kernels (called through the C ABI) with the oracle.

Small and ragged sizes are compared limb for limb with the oracle; the
fixed-size workloads are compared with the SHA-256 of GMP's results
(tests/golden/digests.json) and with size-independent properties.
Bit-exact everywhere: this is integer arithmetic.
"""
import contextlib
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASK = np.uint64(0xFFFFFFFFFFFFFFFF)
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "digests.json")


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x).view(np.uint8).tobytes()).hexdigest()


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


def host(t):
    return t.cpu().numpy().view(np.uint64)


def empty(torch, n):
    return torch.empty(max(n, 1), dtype=torch.int64, device="cuda:0")


def operands(seed, n, m, kind="mixed"):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    if kind == "mixed" and n >= 4:
        A = a.reshape(n, m)
        B = b.reshape(n, m)
        A[0] = MASK
        B[0] = 0
        B[0, 0] = 1  # all-ones + 1: full-length carry
        A[1] = MASK
        B[1] = MASK  # maximal operands
        A[2] = 0  # zero operand
        pool = np.array([0, MASK, MASK, 1], dtype=np.uint64)
        A[3] = pool[rng.integers(0, 4, size=m)]
        B[3] = pool[rng.integers(0, 4, size=m)]
    return a, b


ADD_SIZES = [1, 2, 3, 5, 7, 16, 31, 63, 64, 65, 127, 128, 129, 255, 256, 257, 511, 512, 1000, 2048, 4095, 4096]


@pytest.mark.parametrize("m", ADD_SIZES)
def test_add(ctx, torch_dev, orc, m):
    torch = torch_dev
    for n in (1, 5, 37):
        a, b = operands(m * 131 + n, n, m)
        da, db = dev(torch, a), dev(torch, b)
        dr = empty(torch, n * m)
        dc = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        ctx.add(da, db, dr, m, carry=dc)
        torch.cuda.synchronize()
        want, wc = orc.add(a, b, m)
        got = host(dr)[: n * m]
        assert np.array_equal(got, want), (m, n, int(np.flatnonzero(got != want)[0]))
        assert np.array_equal(dc.cpu().numpy(), wc), (m, n)


ADD_CFGS = [(r, t, nt) for r in (1, 2, 4, 8) for t in (64, 256, 1024) for nt in (0, 1)]


@pytest.mark.parametrize("cfg", ADD_CFGS, ids=lambda c: f"rows{c[0]}-threads{c[1]}-nt{c[2]}")
def test_add_layouts(mi, ctx, torch_dev, orc, cfg):
    """Every addition layout the tuning table can pick up to 256 limbs (rows
    of 128 limbs per wave, threads per block), with and without
    non-temporal loads. A layout whose rows do not cover an instance falls
    back to the default one."""
    rows, threads, nt = cfg
    torch = torch_dev
    for m in (1, 7, 64, 129, 256):
        with tuned(mi, m, {"add.rows": rows, "add.threads": threads, "add.nt": nt}):
            for n in (3, 41):
                a, b = operands(m * 17 + n + rows, n, m)
                dr = empty(torch, n * m)
                dc = torch.empty(n, dtype=torch.uint8, device="cuda:0")
                ctx.add(dev(torch, a), dev(torch, b), dr, m, carry=dc)
                torch.cuda.synchronize()
                want, wc = orc.add(a, b, m)
                got = host(dr)[: n * m]
                assert np.array_equal(got, want), (m, n, int(np.flatnonzero(got != want)[0]))
                assert np.array_equal(dc.cpu().numpy(), wc), (m, n)


@pytest.mark.parametrize("m", [1, 3, 64, 129, 1000, 4096])
def test_add_in_place_and_unaligned_vectors(ctx, torch_dev, orc, m):
    torch = torch_dev
    n = 9
    a, b = operands(m + 5, n, m)
    want, _ = orc.add(a, b, m)
    da, db = dev(torch, a), dev(torch, b)
    ctx.add(da, db, da, m)  # r == a
    torch.cuda.synchronize()
    assert np.array_equal(host(da), want)
    # 8-byte aligned but not 16-byte aligned views: scalar path
    big_a = dev(torch, np.concatenate([[np.uint64(0)], a]))
    big_b = dev(torch, np.concatenate([[np.uint64(0)], b]))
    big_r = empty(torch, n * m + 1)
    ctx.add(big_a[1:], big_b[1:], big_r[1:], m)
    torch.cuda.synchronize()
    assert np.array_equal(host(big_r)[1:], want)


@pytest.mark.parametrize("m", ADD_SIZES)
def test_add6(ctx, torch_dev, orc, m):
    """(6a + 10b) mod B^m in one addition: each limb's high part moves up one
    limb, across lanes, rows and waves, and stops at instance starts."""
    torch = torch_dev
    for n in (1, 5, 21):
        a, b = operands(m * 7 + n, n, m)
        want = orc.add6(a, b, m)
        da = dev(torch, a)
        dr = empty(torch, n * m)
        ctx.add6(da, dev(torch, b), dr, m)
        torch.cuda.synchronize()
        assert np.array_equal(host(dr), want), (m, n)
        # in place, and 8-byte aligned views (scalar loads and stores)
        ctx.add6(da, dev(torch, b), da, m)
        big_a = dev(torch, np.concatenate([[np.uint64(0)], a]))
        big_b = dev(torch, np.concatenate([[np.uint64(0)], b]))
        big_r = empty(torch, n * m + 1)
        ctx.add6(big_a[1:], big_b[1:], big_r[1:], m)
        torch.cuda.synchronize()
        assert np.array_equal(host(da), want), (m, n)
        assert np.array_equal(host(big_r)[1:], want), (m, n)


# Instances longer than one workgroup's rows (4096 limbs) run chunk by chunk
# with the carry kept between chunks: exact chunk multiples, one limb past
# them, odd lengths (scalar loads) and a length that needs many chunks.
LONG_ADD_SIZES = [4097, 8192, 8193, 16385, 32768, 100001]


@pytest.mark.parametrize("m", LONG_ADD_SIZES)
def test_add_long_instances(ctx, torch_dev, orc, m):
    torch = torch_dev
    n = 5
    a, b = operands(m * 3 + 1, n, m)
    # instance 4: a carry chain through every chunk boundary (all-ones plus 1
    # in its lowest limb, with a zero limb near the end to stop it there)
    A, B = a.reshape(n, m), b.reshape(n, m)
    A[4] = MASK
    B[4] = 0
    B[4, 0] = 1
    A[4, m - 3] = 0
    da, db = dev(torch, a), dev(torch, b)
    dr = empty(torch, n * m)
    dc = torch.empty(n, dtype=torch.uint8, device="cuda:0")
    ctx.add(da, db, dr, m, carry=dc)
    dr6 = empty(torch, n * m)
    ctx.add6(da, db, dr6, m)
    torch.cuda.synchronize()
    want, wc = orc.add(a, b, m)
    got = host(dr)
    assert np.array_equal(got, want), (m, int(np.flatnonzero(got != want)[0]))
    assert np.array_equal(dc.cpu().numpy(), wc), m
    assert np.array_equal(host(dr6), orc.add6(a, b, m)), m
    # 8-byte aligned only (scalar path) and in place
    big_a = dev(torch, np.concatenate([[np.uint64(0)], a]))
    big_b = dev(torch, np.concatenate([[np.uint64(0)], b]))
    ctx.add(big_a[1:], big_b[1:], big_a[1:], m)
    torch.cuda.synchronize()
    assert np.array_equal(host(big_a)[1:], want), m
    # add6 in place: each chunk's first limb takes the high part of the
    # previous chunk's last limb, which the previous chunk has overwritten
    ctx.add6(da, db, da, m)
    torch.cuda.synchronize()
    assert np.array_equal(host(da), orc.add6(a, b, m)), m


def test_add_long_strided(mi, ctx, torch_dev, orc):
    """Long instances at instance strides above their length (the gaps stay untouched)."""
    torch = torch_dev
    n, m, sa, sb, sr = 3, 9000, 9003, 9010, 9001
    rng = np.random.default_rng(77)
    A = rng.integers(0, 2**64, size=n * sa, dtype=np.uint64)
    B = rng.integers(0, 2**64, size=n * sb, dtype=np.uint64)
    R = np.full(n * sr, 0x5A5A5A5A5A5A5A5A, dtype=np.uint64)
    a = np.stack([A[i * sa:i * sa + m] for i in range(n)]).reshape(-1)
    b = np.stack([B[i * sb:i * sb + m] for i in range(n)]).reshape(-1)
    for op, want in (("add", orc.add(a, b, m)[0]), ("add6", orc.add6(a, b, m))):
        dr = dev(torch, R)
        ctx.strided(op, n, m, dev(torch, A), sa, dev(torch, B), sb, dr, sr)
        torch.cuda.synchronize()
        got = host(dr).reshape(n, sr)
        assert np.array_equal(got[:, :m].reshape(-1), want), op
        assert np.all(got[:, m:] == np.uint64(0x5A5A5A5A5A5A5A5A)), op


MUL_SIZES = [1, 2, 3, 4, 5, 8, 16, 17, 31, 32, 64, 100, 127, 128, 129, 200, 256, 512, 777, 1024, 2048]


# Classical-kernel configurations (tuning keys): u32 words with mad + carry
# at Q = 4, 8, 16 columns per lane group, the first group parked in LDS or
# in registers, with and without operand prefetch; 28-bit digits with one
# mad per product; balanced 8-bit digits on the int8 matrix cores (automatic
# and one wave per instance; 16 x 16 x 64 tiles with one product per
# workgroup at 1, 3 or the default number of waves; and the
# one-product-per-wave 16 x 16 x 64 kernel where it is compiled: 64 and 128
# limbs).
MUL_CFGS = [
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 4, "mul.rpark": 0, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 4, "mul.rpark": 0, "mul.prefetch": 1},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 8, "mul.rpark": 0, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 8, "mul.rpark": 0, "mul.prefetch": 1},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 8, "mul.rpark": 1, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 8, "mul.rpark": 1, "mul.prefetch": 1},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 16, "mul.rpark": 0, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 16, "mul.rpark": 1, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 0, "mul.q": 16, "mul.rpark": 1, "mul.prefetch": 1},
    {"mul.mfma": 0, "mul.digits": 1, "mul.rpark": 0, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 1, "mul.rpark": 1, "mul.prefetch": 0},
    {"mul.mfma": 0, "mul.digits": 1, "mul.rpark": 0, "mul.prefetch": 1},
    {"mul.mfma": 1, "mul.mfma_wpi": 0, "mul.wave16": 0, "mul.m16": 0},
    {"mul.mfma": 1, "mul.mfma_wpi": 1, "mul.wave16": 0, "mul.m16": 0},
    {"mul.mfma": 1, "mul.wave16": 0, "mul.m16": 1, "mul.m16_waves": 0},
    {"mul.mfma": 1, "mul.wave16": 0, "mul.m16": 1, "mul.m16_waves": 1},
    {"mul.mfma": 1, "mul.wave16": 0, "mul.m16": 1, "mul.m16_waves": 3},
    {"mul.mfma": 1, "mul.wave16": 1},
]


def _cfg_id(c):
    return "-".join(f"{k.split('.')[1]}{v}" for k, v in c.items())


def _size_class(m):
    c = 1
    while c < m:
        c <<= 1
    return c


@contextlib.contextmanager
def tuned(mi, m, cfg):
    """Sets `cfg` for the size class of m (class keys take precedence over
    the global ones, and the tuning file sets class keys). Afterwards every
    touched key holds its previous value, or the global value where the
    class key was unset (the same lookup result)."""
    saved = {}
    for k, v in cfg.items():
        key = f"{k}@{_size_class(m)}"
        try:
            saved[key] = mi.tune_get(key)
        except Exception:
            saved[key] = mi.tune_get(k)
        mi.tune_set(key, v)
    try:
        yield
    finally:
        for key, v in saved.items():
            mi.tune_set(key, v)


@pytest.mark.parametrize("cfg", MUL_CFGS, ids=_cfg_id)
@pytest.mark.parametrize("m", MUL_SIZES)
def test_mul_full_and_lo(mi, ctx, torch_dev, orc, m, cfg):
    """Every classical-kernel configuration against the oracle."""
    with tuned(mi, m, cfg):
        _check_mul(ctx, torch_dev, orc, m)


MFMA_PATTERN_SIZES = [1, 3, 4, 5, 64, 100, 128, 129, 256, 777]


@pytest.mark.parametrize("wpi", [0, 1, 2, "w16", "m16"])
@pytest.mark.parametrize("m", MFMA_PATTERN_SIZES)
def test_mul_mfma_digit_patterns(mi, ctx, torch_dev, orc, m, wpi):
    """The matrix-core kernel writes operands in balanced base-256 digits
    (byte + 128 with carries): bytes 0x7F make the carry run through the
    whole operand, 0x80 and 0xFF bytes sit at the digit range's ends, and
    mixed limbs break chains at random places."""
    torch = torch_dev
    rng = np.random.default_rng(m * 7 + {"w16": 3, "m16": 4}.get(wpi, wpi))
    k7f = np.uint64(0x7F7F7F7F7F7F7F7F)
    k80 = np.uint64(0x8080808080808080)
    pool = np.array([k7f, k80, MASK, 0, np.uint64(0x7F7F7F7F7F7F7F80)], dtype=np.uint64)
    rows = [
        (np.full(m, k7f), np.full(m, k7f)),
        (np.full(m, k80), np.full(m, MASK)),
        (np.full(m, MASK), np.full(m, k7f)),
        (np.full(m, k7f), rng.integers(0, 2**64, size=m, dtype=np.uint64)),
        (pool[rng.integers(0, 5, size=m)], pool[rng.integers(0, 5, size=m)]),
        (np.zeros(m, dtype=np.uint64), np.full(m, k80)),
    ]
    a = np.concatenate([r[0] for r in rows]).astype(np.uint64)
    b = np.concatenate([r[1] for r in rows]).astype(np.uint64)
    n = len(rows)
    cfg = ({"mul.mfma": 1, "mul.wave16": 1} if wpi == "w16" else
           {"mul.mfma": 1, "mul.wave16": 0, "mul.m16": 1} if wpi == "m16" else
           {"mul.mfma": 1, "mul.mfma_wpi": wpi, "mul.wave16": 0, "mul.m16": 0})
    with tuned(mi, m, cfg):
        da, db = dev(torch, a), dev(torch, b)
        rf = empty(torch, 2 * n * m)
        rl = empty(torch, n * m)
        ctx.mul_full(da, db, rf, m)
        ctx.mul_lo(da, db, rl, m)
        torch.cuda.synchronize()
    got = host(rf)
    want = orc.mul_full(a, b, m)
    assert np.array_equal(got, want), (m, int(np.flatnonzero(got != want)[0]))
    assert np.array_equal(host(rl), orc.mul_lo(a, b, m))


@pytest.mark.parametrize("seed", range(24))
def test_mul_mfma_layout_sweep(mi, ctx, torch_dev, orc, seed):
    """Random sizes, batch counts and launch layouts (instances per
    workgroup, waves per instance) of the matrix-core kernel. Workgroups
    whose rows cover exactly their products load and store without range
    checks; a partial last workgroup or an odd product length takes the
    checked path. Both agree with the oracle bit for bit. (Layouts beyond
    16 waves per workgroup run on the word kernel, which must agree too.)"""
    torch = torch_dev
    rng = np.random.default_rng(0x5EED + seed)
    m = int(rng.choice([int(rng.integers(1, 64)), int(rng.integers(64, 300)), int(rng.integers(300, 1100))]))
    n = int(rng.integers(1, 25))
    ipb = int(rng.choice([0, 1, 2, 3, 4, 5, 8]))
    wpi = int(rng.choice([0, 1, 2, 4]))
    a, b = operands(seed * 31 + m, n, m)
    with tuned(mi, m, {"mul.mfma": 1, "mul.mfma_wpi": wpi, "mul.mfma_ipb": ipb, "mul.wave16": 0, "mul.m16": 0}):
        da, db = dev(torch, a), dev(torch, b)
        rf = empty(torch, 2 * n * m)
        rl = empty(torch, n * m)
        ctx.mul_full(da, db, rf, m)
        ctx.mul_lo(da, db, rl, m)
        torch.cuda.synchronize()
    got = host(rf)
    want = orc.mul_full(a, b, m)
    assert np.array_equal(got, want), (m, n, ipb, wpi, int(np.flatnonzero(got != want)[0]))
    assert np.array_equal(host(rl), orc.mul_lo(a, b, m)), (m, n, ipb, wpi)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 9, 67])
@pytest.mark.parametrize("m", [64, 128])
def test_mul_wave16_batches(mi, ctx, torch_dev, orc, m, n):
    """The one-product-per-wave kernel: batch counts that leave waves of the
    last workgroup idle, instance strides above the operand and product
    lengths (odd strides: 8-byte alignment only, the gaps untouched), and the
    same products from the 32 x 32 x 32 kernel."""
    torch = torch_dev
    a, b = operands(m * 1000 + n, n, m)
    sa, sb = m + 3, m + 1
    A = np.full(n * sa, 0x3C3C3C3C3C3C3C3C, dtype=np.uint64)
    B = np.full(n * sb, 0x3C3C3C3C3C3C3C3C, dtype=np.uint64)
    for i in range(n):
        A[i * sa:i * sa + m] = a[i * m:(i + 1) * m]
        B[i * sb:i * sb + m] = b[i * m:(i + 1) * m]
    gap = np.uint64(0x5A5A5A5A5A5A5A5A)
    for op, pl, want in (("mul_full", 2 * m, orc.mul_full(a, b, m)), ("mul_lo", m, orc.mul_lo(a, b, m))):
        sr = pl + 5
        with tuned(mi, m, {"mul.mfma": 1, "mul.wave16": 1}):
            dr = dev(torch, np.full(n * sr, gap, dtype=np.uint64))
            ctx.strided(op, n, m, dev(torch, A), sa, dev(torch, B), sb, dr, sr)
            dc = empty(torch, n * pl)
            getattr(ctx, op)(dev(torch, a), dev(torch, b), dc, m)
            torch.cuda.synchronize()
        got = host(dr).reshape(n, sr)
        assert np.array_equal(got[:, :pl].reshape(-1), want), (op, m, n)
        assert np.all(got[:, pl:] == gap), (op, m, n)
        assert np.array_equal(host(dc), want), (op, m, n)
        with tuned(mi, m, {"mul.mfma": 1, "mul.wave16": 0}):
            do = empty(torch, n * pl)
            getattr(ctx, op)(dev(torch, a), dev(torch, b), do, m)
            torch.cuda.synchronize()
        assert np.array_equal(host(do), want), (op, m, n)


@pytest.mark.parametrize("seed", range(24))
def test_mul_mfma16_sweep(mi, ctx, torch_dev, orc, seed):
    """16 x 16 x 64 tiles with the tile plan at run time: random sizes
    (odd, straddling rows and tiles, up to 2048 limbs), batch counts and
    waves per product, full and low-half products, against the oracle and
    the 32 x 32 x 32 kernel."""
    torch = torch_dev
    rng = np.random.default_rng(0x16F + seed)
    m = int(rng.choice([int(rng.integers(1, 64)), int(rng.integers(64, 300)), int(rng.integers(300, 2049))]))
    n = int(rng.integers(1, 7))
    waves = int(rng.choice([0, 1, 2, 3, 5, 8, 16]))
    a, b = operands(seed * 37 + m, n, m)
    want_f, want_l = orc.mul_full(a, b, m), orc.mul_lo(a, b, m)
    with tuned(mi, m, {"mul.mfma": 1, "mul.wave16": 0, "mul.m16": 1, "mul.m16_waves": waves}):
        da, db = dev(torch, a), dev(torch, b)
        rf = empty(torch, 2 * n * m)
        rl = empty(torch, n * m)
        ctx.mul_full(da, db, rf, m)
        ctx.mul_lo(da, db, rl, m)
        torch.cuda.synchronize()
    got = host(rf)
    assert np.array_equal(got, want_f), (m, n, waves, int(np.flatnonzero(got != want_f)[0]))
    assert np.array_equal(host(rl), want_l), (m, n, waves)


def _check_mul(ctx, torch_dev, orc, m):
    torch = torch_dev
    n = 6 if m >= 1024 else 11
    a, b = operands(m * 17, n, m)
    da, db = dev(torch, a), dev(torch, b)
    rf = empty(torch, 2 * n * m)
    rl = empty(torch, n * m)
    ctx.mul_full(da, db, rf, m)
    ctx.mul_lo(da, db, rl, m)
    torch.cuda.synchronize()
    want = orc.mul_full(a, b, m)
    got = host(rf)
    assert np.array_equal(got, want), (m, int(np.flatnonzero(got != want)[0]))
    assert np.array_equal(host(rl), orc.mul_lo(a, b, m))


POLY_SIZES = [1, 2, 4, 16, 33, 64, 127, 128, 129, 300, 512, 1024, 2048]


@pytest.mark.parametrize("fused", [2, 1, 0], ids=["wave16", "fused", "chain"])
@pytest.mark.parametrize("m", POLY_SIZES)
def test_poly(mi, ctx, torch_dev, orc, m, fused):
    """polyKerQ: one instance per wave on the 16 x 16 x 64 kernel (64 and
    128 limbs), the one-kernel 32 x 32 x 32 version and the launch chain,
    batch counts that leave the last workgroup partial."""
    torch = torch_dev
    cfg = {"poly.fused": min(fused, 1), "mul.wave16": 1 if fused == 2 else 0}
    for n in ((13, 1) if m < 1024 else (5, 1)):
        a, b = operands(m * 3 + 1 + n, n, m)
        dr = empty(torch, n * m)
        with tuned(mi, m, cfg):
            ctx.poly(dev(torch, a), dev(torch, b), dr, m)
            torch.cuda.synchronize()
        got, want = host(dr)[: n * m], orc.poly(a, b, m)
        assert np.array_equal(got, want), (m, n, fused, int(np.flatnonzero(got != want)[0]))


@pytest.mark.parametrize("m,wpi,ipb", [(64, 1, 1), (64, 1, 8), (128, 1, 2), (128, 2, 4), (256, 1, 4), (256, 2, 2),
                                       (512, 1, 2), (512, 4, 1), (1024, 1, 1), (1024, 2, 2), (2048, 2, 1)])
def test_poly_fused_layouts(mi, ctx, torch_dev, orc, m, wpi, ipb):
    """The fused poly kernel under the matrix-core layout keys (waves per
    instance, instances per workgroup; the kernel halves the instances per
    workgroup when its LDS buffers do not fit), including the 8-row
    normalisation."""
    torch = torch_dev
    n = 7
    a, b = operands(m * 5 + wpi + 3 * ipb, n, m)
    dr = empty(torch, n * m)
    with tuned(mi, m, {"poly.fused": 1, "mul.mfma_wpi": wpi, "mul.mfma_ipb": ipb, "mul.wave16": 0}):
        ctx.poly(dev(torch, a), dev(torch, b), dr, m)
        torch.cuda.synchronize()
    assert np.array_equal(host(dr), orc.poly(a, b, m)), (m, wpi, ipb)


POLY_NTT_SIZES = [1, 2, 3, 64, 100, 127, 128, 129, 255, 256, 511, 512, 1000, 1024, 2047, 2048]


@pytest.mark.parametrize("vals", [16, 32])
@pytest.mark.parametrize("m", POLY_NTT_SIZES)
def test_poly_ntt(mi, ctx, torch_dev, orc, m, vals):
    """NTT composite (polyFttKer, one kernel): same closed form as the
    classical one, with 16 and 32 values per thread (32 needs transforms of
    2^11 points or more, 256 limbs and up)."""
    torch = torch_dev
    n = 3 if m >= 1024 else 9
    a, b = operands(m * 11 + 5, n, m)
    dr = empty(torch, n * m)
    L = 1 << max(10, (8 * m - 1).bit_length())
    with tuned(mi, L, {"ntt.vals": vals}):
        ctx.poly_ntt(dev(torch, a), dev(torch, b), dr, m)
        torch.cuda.synchronize()
    got, want = host(dr), orc.poly(a, b, m)
    assert np.array_equal(got, want), (m, vals, int(np.flatnonzero(got != want)[0]))
    if m <= midint_max_mul(ctx):
        dc = empty(torch, n * m)
        ctx.poly(dev(torch, a), dev(torch, b), dc, m)
        torch.cuda.synchronize()
        assert np.array_equal(host(dc), host(dr))


def test_scratch_users_on_two_streams(ctx, torch_dev, orc):
    """poly, poly_ntt and the repacking u32 path share the context scratch;
    calls on two streams (default priority) must each give their own exact
    result. Both batches are large enough to overlap on the device if they
    were not ordered."""
    torch = torch_dev
    n, m = 2048, 64
    a1, b1 = operands(901, n, m)
    a2, b2 = operands(902, n, m)
    d = [dev(torch, x) for x in (a1, b1, a2, b2)]
    r = [empty(torch, n * m) for _ in range(4)]
    m32 = 2 * m - 1  # odd word count: the repacking path
    w1 = a1.view(np.uint32)[: n * m32].copy()
    w2 = b1.view(np.uint32)[: n * m32].copy()
    dw1 = torch.from_numpy(w1.view(np.int32)).to("cuda:0")
    dw2 = torch.from_numpy(w2.view(np.int32)).to("cuda:0")
    rw = torch.empty(n * m32, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        ctx.poly(d[0], d[1], r[0], m, stream=s1)
        ctx.poly(d[2], d[3], r[1], m, stream=s2)
        ctx.poly_ntt(d[2], d[3], r[2], m, stream=s1)
        ctx.mul_lo_u32(dw1, dw2, rw, m32, n, stream=s2)
        ctx.poly_ntt(d[0], d[1], r[3], m, stream=s2)
    torch.cuda.synchronize()
    want1, want2 = orc.poly(a1, b1, m), orc.poly(a2, b2, m)
    assert np.array_equal(host(r[0]), want1)
    assert np.array_equal(host(r[1]), want2)
    assert np.array_equal(host(r[2]), want2)
    assert np.array_equal(host(r[3]), want1)
    # u32 words: product mod 2^(32 m32) from the u64 oracle on zero-extended operands
    za = np.zeros(n * m, dtype=np.uint64)
    zb = np.zeros(n * m, dtype=np.uint64)
    za.view(np.uint32).reshape(n, 2 * m)[:, :m32] = w1.reshape(n, m32)
    zb.view(np.uint32).reshape(n, 2 * m)[:, :m32] = w2.reshape(n, m32)
    want_w = orc.mul_lo(za, zb, m).view(np.uint32).reshape(n, 2 * m)[:, :m32].reshape(-1)
    assert np.array_equal(rw.cpu().numpy().view(np.uint32), want_w)


def midint_max_mul(ctx):
    from __graft_entry__ import load_package

    return load_package().lib().midint_max_limbs_mul()


# transform lengths 2^6 .. 2^14 points (L >= 8m): every length at least once
NTT_SIZES = [1, 2, 3, 8, 16, 24, 64, 100, 128, 256, 300, 512, 1000, 1024, 2048]


@pytest.mark.parametrize("m", NTT_SIZES)
def test_mul_ntt(ctx, torch_dev, orc, m):
    torch = torch_dev
    n = 3 if m >= 1024 else 7
    a, b = operands(m * 29, n, m)
    dr = empty(torch, 2 * n * m)
    ctx.mul_ntt(dev(torch, a), dev(torch, b), dr, m)
    torch.cuda.synchronize()
    want = orc.mul_full(a, b, m)
    got = host(dr)
    assert np.array_equal(got, want), (m, int(np.flatnonzero(got != want)[0]))


@pytest.mark.parametrize("vals", [16, 32])
@pytest.mark.parametrize("m", [256, 512, 1024, 2048])
def test_mul_ntt_values_per_thread(mi, ctx, torch_dev, orc, m, vals):
    """Both register-pass layouts of the 2^11 .. 2^14-point kernels (16 or
    32 values per thread, ntt.vals), whichever the tuning table picks."""
    torch = torch_dev
    n = 3
    a, b = operands(m * 31 + vals, n, m)
    with tuned(mi, 8 * m, {"ntt.vals": vals}):
        dr = empty(torch, 2 * n * m)
        ctx.mul_ntt(dev(torch, a), dev(torch, b), dr, m)
        torch.cuda.synchronize()
    got = host(dr)
    want = orc.mul_full(a, b, m)
    assert np.array_equal(got, want), (m, vals, int(np.flatnonzero(got != want)[0]))


@pytest.mark.parametrize("m", [128, 512, 2048])
def test_mul_ntt_lds_kernel(mi, ctx, torch_dev, orc, m):
    """ntt.fast = 0: 2^10 .. 2^14-point products on the all-in-LDS kernel
    (one radix-2 stage per barrier) instead of the register passes."""
    torch = torch_dev
    n = 3
    a, b = operands(m * 37, n, m)
    old = mi.tune_get("ntt.fast")
    mi.tune_set("ntt.fast", 0)
    try:
        dr = empty(torch, 2 * n * m)
        ctx.mul_ntt(dev(torch, a), dev(torch, b), dr, m)
        torch.cuda.synchronize()
    finally:
        mi.tune_set("ntt.fast", old)
    got = host(dr)
    want = orc.mul_full(a, b, m)
    assert np.array_equal(got, want), (m, int(np.flatnonzero(got != want)[0]))


def test_ntt_conv_kat(ctx, torch_dev, orc):
    """cuda/main.cu:637-652: the reference's FFT known-answer test (raw
    convolution), padded to the smallest supported transform (64 points)."""
    torch = torch_dev
    a = [11400, 28374, 23152, 9576, 29511, 20787, 13067, 14015]
    b = [30268, 20788, 8033, 15446, 26275, 11619, 2494, 7016]
    ref = [345055200, 1095807432, 1382179648, 1175142886, 2016084656, 2555168834, 2179032777, 1990011337,
           1860865174, 1389799087, 942120918, 778961552, 341270975, 126631482, 98329240, 0]
    A = np.zeros(64, dtype=np.uint64)
    B = np.zeros(64, dtype=np.uint64)
    A[:8] = a
    B[:8] = b
    dr = empty(torch, 64)
    ctx.ntt_conv(dev(torch, A), dev(torch, B), dr, 6)
    torch.cuda.synchronize()
    got = host(dr)
    assert [int(v) for v in got[:16]] == ref and not got[16:].any()


@pytest.mark.parametrize("lg", [6, 10, 14])
def test_ntt_conv_random(ctx, torch_dev, orc, lg):
    torch = torch_dev
    L = 1 << lg
    rng = np.random.default_rng(lg)
    n = 2
    a = rng.integers(0, orc.GL_P, size=n * L, dtype=np.uint64)
    b = rng.integers(0, orc.GL_P, size=n * L, dtype=np.uint64)
    dr = empty(torch, n * L)
    ctx.ntt_conv(dev(torch, a), dev(torch, b), dr, lg)
    torch.cuda.synchronize()
    got = host(dr)
    for i in range(n):
        s = slice(i * L, (i + 1) * L)
        assert np.array_equal(got[s], orc.ntt_conv(a[s], b[s]))


def test_errors(ctx, torch_dev, mi):
    torch = torch_dev
    L = mi.lib()
    x = empty(torch, 64)
    h = ctx.handle
    p = x.data_ptr()
    assert L.midint_add(h, 1, 0, p, p, p, None, None) == 1          # m == 0
    assert L.midint_add(h, 0, 8, None, None, None, None, None) == 0  # empty batch
    assert L.midint_add(h, 1, 8, p + 4, p, p, None, None) == 1      # misaligned
    assert L.midint_mul_full(h, 1, 4097, p, p, p, None) == 2         # too large
    assert L.midint_mul_ntt(h, 1, 1 << 20, p, p, p, None) == 2
    assert L.midint_add(h, 1, L.midint_max_limbs_add() + 1, p, p, p, None, None) == 2  # rejected before launch


# ---------------------------------------------------------------- full size
def _golden(name):
    with open(GOLD) as f:
        return json.load(f)["workloads"][name]


def test_workload_add_digest(ctx, torch_dev, mi):
    from workloads import make_inputs

    torch = torch_dev
    a, b = make_inputs("add_65536x64", mi.gen_limbs)
    da, db = dev(torch, a), dev(torch, b)
    dr = empty(torch, a.size)
    dc = torch.empty(65536, dtype=torch.uint8, device="cuda:0")
    ctx.add(da, db, dr, 64, carry=dc)
    torch.cuda.synchronize()
    r = host(dr)
    assert sha(r) == _golden("add_65536x64")["result"]
    c = dc.cpu().numpy()
    assert c[::100].all()  # forced all-ones + 1 rows carry out
    # (a + b) - b == a  (property at full size, via a second addition with -b)
    negb = (~b.reshape(-1, 64)).reshape(-1).copy()
    one = np.zeros_like(b)
    one.reshape(-1, 64)[:, 0] = 1
    t = empty(torch, a.size)
    ctx.add(dev(torch, negb), dev(torch, one), t, 64)  # -b mod B^64
    back = empty(torch, a.size)
    ctx.add(dr, t, back, 64)
    torch.cuda.synchronize()
    assert np.array_equal(host(back), a)


def test_workload_mul_digest(ctx, torch_dev, mi):
    from workloads import make_inputs

    torch = torch_dev
    a, b = make_inputs("mul_16384x128", mi.gen_limbs)
    da, db = dev(torch, a), dev(torch, b)
    dr = empty(torch, 2 * a.size)
    ctx.mul_full(da, db, dr, 128)
    torch.cuda.synchronize()
    assert sha(host(dr)) == _golden("mul_16384x128")["result"]
    # commutativity and agreement of the two algorithms at full size
    dr2 = empty(torch, 2 * a.size)
    ctx.mul_ntt(db, da, dr2, 128)
    torch.cuda.synchronize()
    assert torch.equal(dr, dr2)


def test_workload_ntt_digest(ctx, torch_dev, mi):
    from workloads import make_inputs

    torch = torch_dev
    a, b = make_inputs("ntt_1024x2048", mi.gen_limbs)
    da, db = dev(torch, a), dev(torch, b)
    dr = empty(torch, 2 * a.size)
    ctx.mul_ntt(da, db, dr, 2048)
    torch.cuda.synchronize()
    assert sha(host(dr)) == _golden("ntt_1024x2048")["result"]
    # classical and NTT products agree on a slice of the workload
    k = 16
    rc = empty(torch, 2 * k * 2048)
    ctx.mul_full(da[: k * 2048], db[: k * 2048], rc, 2048)
    torch.cuda.synchronize()
    assert torch.equal(rc, dr[: 2 * k * 2048])


def test_host_dropins(mi, torch_dev, orc):
    n, m = 33, 16
    a, b = operands(77, n, m)
    r, us = mi.gpu_add_host(a, b, m, runs=3)
    assert np.array_equal(r, orc.add(a, b, m)[0]) and us > 0
    r, _ = mi.gpu_mul_host(a, b, m, full=True, runs=2)
    assert np.array_equal(r, orc.mul_full(a, b, m))
    r, _ = mi.gpu_mul_host(a, b, m, full=False, runs=0)
    assert np.array_equal(r, orc.mul_lo(a, b, m))
    r, _ = mi.gpu_mul_fft_host(a, b, m, runs=1)
    assert np.array_equal(r, orc.mul_full(a, b, m))
