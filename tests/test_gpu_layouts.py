"""External layouts through the C ABI: instance strides (array-of-structs),
CGBN's instance_t arrays (cuda/cgbn-tests/cgbn-kers.cu.h:5-9, {a, b, sum}
per instance) and the Futhark entry layout [m][ipb][4n]u64
(futhark-clean/big-add.fut:124-128). Expected values from the oracle on
batch-major copies; words outside the written fields must be untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SENT = np.uint64(0x5A5A5A5A5A5A5A5A)


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


def spread(x, n, m, stride):
    """Batch-major (n, m) -> buffer with instance stride `stride`, gaps = SENT."""
    buf = np.full(n * stride, SENT, dtype=np.uint64)
    buf.reshape(n, stride)[:, :m] = x.reshape(n, m)
    return buf


def expected(orc, op, a, b, m):
    if op == "add":
        return orc.add(a, b, m)[0]
    if op == "add6":
        return orc.add6(a, b, m)
    if op == "mul_lo":
        return orc.mul_lo(a, b, m)
    if op == "mul_full":
        return orc.mul_full(a, b, m)
    return orc.poly(a, b, m)


@pytest.mark.parametrize("op", ["add", "add6", "mul_lo", "mul_full", "poly"])
@pytest.mark.parametrize("m", [1, 3, 64, 100, 300])
def test_strided_ops(ctx, torch_dev, orc, op, m):
    torch = torch_dev
    n = 11
    rng = np.random.default_rng(m * 17 + len(op))
    a = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    a[:m] = np.uint64(0xFFFFFFFFFFFFFFFF)
    b[:m] = 0
    b[0] = 1
    out = 2 * m if op == "mul_full" else m
    sa, sb, sr = m + 2, m + 5, out + 3
    da, db = dev(torch, spread(a, n, m, sa)), dev(torch, spread(b, n, m, sb))
    dr = dev(torch, np.full(n * sr, SENT, dtype=np.uint64))
    carry = torch.zeros(n, dtype=torch.uint8, device="cuda:0") if op == "add" else None
    ctx.strided(op, n, m, da, sa, db, sb, dr, sr, carry=carry)
    torch.cuda.synchronize()
    got = host(dr).reshape(n, sr)
    assert np.array_equal(got[:, :out].reshape(-1), expected(orc, op, a, b, m))
    assert (got[:, out:] == SENT).all()
    if op == "add":
        assert np.array_equal(carry.cpu().numpy(), orc.add(a, b, m)[1])


def cgbn_expected(orc, op, a32, b32, n, w):
    """Oracle on the zero-extended operands, truncated to w words."""
    m = (w + 1) // 2
    def ext(x):
        pad = np.zeros((n, 2 * m), dtype=np.uint32)
        pad[:, :w] = x.reshape(n, w)
        return pad.reshape(-1).view(np.uint64)
    pa, pb = ext(a32), ext(b32)
    name = {0: "add", 1: "add6", 2: "mul_lo", 3: "poly"}[op]
    r = expected(orc, name, pa, pb, m)
    return r.reshape(n, m).view(np.uint32)[:, :w]


@pytest.mark.parametrize("op", [0, 1, 2, 3])
@pytest.mark.parametrize("bits", [64, 96, 1024, 2048, 3200, 3232, 4096])
def test_cgbn_instances(mi, ctx, torch_dev, orc, op, bits):
    torch = torch_dev
    n = 7
    w = bits // 32
    rng = np.random.default_rng(bits + op)
    inst = rng.integers(0, 2**32, size=(n, 3, w), dtype=np.uint32)
    inst[0, 0, :] = 0xFFFFFFFF
    inst[0, 1, :] = 0
    inst[0, 1, 0] = 1
    a32, b32 = inst[:, 0, :].copy(), inst[:, 1, :].copy()
    flat = inst.reshape(-1)
    buf = torch.from_numpy(flat.view(np.int32)).to("cuda:0")
    ctx.cgbn(op, buf, n, bits)
    torch.cuda.synchronize()
    got = buf.cpu().numpy().view(np.uint32).reshape(n, 3, w)
    assert np.array_equal(got[:, 0, :], a32) and np.array_equal(got[:, 1, :], b32)
    assert np.array_equal(got[:, 2, :], cgbn_expected(orc, op, a32, b32, n, w))


def test_futhark_entry_layout(ctx, torch_dev, orc):
    """[mg][ipb][4n]u64 row-major is batch-major: instance (g, i) at (g*ipb + i) * 4n."""
    torch = torch_dev
    mg, ipb, n4 = 64, 4, 16
    rng = np.random.default_rng(5)
    ass = rng.integers(0, 2**64, size=(mg, ipb, 4 * n4), dtype=np.uint64)
    bss = rng.integers(0, 2**64, size=(mg, ipb, 4 * n4), dtype=np.uint64)
    r = torch.empty(mg * ipb * 4 * n4, dtype=torch.int64, device="cuda:0")
    ctx.add(dev(torch, ass.reshape(-1)), dev(torch, bss.reshape(-1)), r, 4 * n4)
    torch.cuda.synchronize()
    got = host(r).reshape(mg, ipb, 4 * n4)
    want = orc.add(ass.reshape(-1), bss.reshape(-1), 4 * n4)[0].reshape(mg, ipb, 4 * n4)
    assert np.array_equal(got, want)


def test_layout_errors(mi, ctx, torch_dev):
    L = mi.lib()
    x = torch_dev.zeros(64, dtype=torch_dev.int64, device="cuda:0")
    p = x.data_ptr()
    assert L.midint_add_strided(ctx.handle, 2, 8, p, 7, p, 8, p, 8, None, None) == 1   # stride < m
    assert L.midint_mul_full_strided(ctx.handle, 2, 8, p, 8, p, 8, p, 15, None) == 1   # stride < 2m
    assert L.midint_cgbn(ctx.handle, 0, 1, 48, p, None) == 1                          # bits % 32
    assert L.midint_cgbn(ctx.handle, 9, 1, 64, p, None) == 1                          # unknown op
