"""u32-limb entry points (the reference's U32bits base) through the C ABI.
Expected values: the oracle on the operands zero-extended to u64 limbs,
truncated to m32 words (every operation commutes with truncation mod
2^(32 m32)); carries from the bit above the kept words. Covers odd and even
word counts, 8-byte aligned pointers (in-place u64 path) and 4-byte aligned
ones (repack path)."""
import numpy as np
import pytest

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


def to_u64(x32, n, m32):
    m = (m32 + 1) // 2
    pad = np.zeros((n, 2 * m), dtype=np.uint32)
    pad[:, :m32] = x32.reshape(n, m32)
    return pad.reshape(-1).view(np.uint64), m


def words(x64, n, limbs, w32):
    return x64.reshape(n, limbs).view(np.uint32).reshape(n, 2 * limbs)[:, :w32]


def device_u32(torch, x32, offset):
    """x32 on the device at a 4-byte offset of `offset` words from an aligned base."""
    buf = torch.zeros(x32.size + 2, dtype=torch.int32, device="cuda:0")
    buf[offset:offset + x32.size] = torch.from_numpy(x32.view(np.int32))
    return buf, buf[offset:offset + x32.size]


def run(torch, fn, n, m32, a32, b32, out_w32, offset, with_carry=False):
    _, da = device_u32(torch, a32, offset)
    _, db = device_u32(torch, b32, offset)
    rbuf, dr = device_u32(torch, np.zeros(n * out_w32, dtype=np.uint32), offset)
    carry = torch.zeros(n, dtype=torch.uint8, device="cuda:0") if with_carry else None
    if with_carry:
        fn(da, db, dr, m32, n, carry=carry)
    else:
        fn(da, db, dr, m32, n)
    torch.cuda.synchronize()
    out = dr.cpu().numpy().view(np.uint32).reshape(n, out_w32)
    return out, (carry.cpu().numpy() if with_carry else None)


SIZES = [1, 2, 3, 7, 64, 127, 128, 255, 1001]


@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("m32", SIZES)
def test_u32_add_add6(ctx, torch_dev, orc, m32, offset):
    n = 9
    rng = np.random.default_rng(m32 * 3 + offset)
    a32 = rng.integers(0, 2**32, size=n * m32, dtype=np.uint32)
    b32 = rng.integers(0, 2**32, size=n * m32, dtype=np.uint32)
    a32[:m32] = 0xFFFFFFFF
    b32[:m32] = 0
    b32[0] = 1  # instance 0: all-ones + 1 -> 0, carry 1
    pa, m = to_u64(a32, n, m32)
    pb, _ = to_u64(b32, n, m32)
    got, carry = run(torch_dev, ctx.add_u32, n, m32, a32, b32, m32, offset, with_carry=True)
    s64, c64 = orc.add(pa, pb, m)
    full = words(s64, n, m, 2 * m)
    # odd m32: the carry is the bit above the kept words; even m32: the
    # carry-out of the m-limb sum
    want_carry = full[:, m32] & 1 if 2 * m > m32 else c64
    assert np.array_equal(got, full[:, :m32])
    assert np.array_equal(carry, want_carry.astype(np.uint8))
    assert carry[0] == 1 and not got[0].any()
    got6, _ = run(torch_dev, ctx.add6_u32, n, m32, a32, b32, m32, offset)
    assert np.array_equal(got6, words(orc.add6(pa, pb, m), n, m, m32))


@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("m32", SIZES)
def test_u32_mul_poly(ctx, torch_dev, orc, m32, offset):
    n = 5
    rng = np.random.default_rng(m32 * 5 + offset)
    a32 = rng.integers(0, 2**32, size=n * m32, dtype=np.uint32)
    b32 = rng.integers(0, 2**32, size=n * m32, dtype=np.uint32)
    a32[:m32] = 0xFFFFFFFF
    b32[:m32] = 0xFFFFFFFF
    pa, m = to_u64(a32, n, m32)
    pb, _ = to_u64(b32, n, m32)
    full = words(orc.mul_full(pa, pb, m), n, 2 * m, 2 * m32)
    got, _ = run(torch_dev, ctx.mul_full_u32, n, m32, a32, b32, 2 * m32, offset)
    assert np.array_equal(got, full)
    lo, _ = run(torch_dev, ctx.mul_lo_u32, n, m32, a32, b32, m32, offset)
    assert np.array_equal(lo, full[:, :m32])
    gp, _ = run(torch_dev, ctx.poly_u32, n, m32, a32, b32, m32, offset)
    assert np.array_equal(gp, words(orc.poly(pa, pb, m), n, m, m32))


def test_u32_errors(mi, ctx, torch_dev):
    L = mi.lib()
    x = torch_dev.zeros(16, dtype=torch_dev.int32, device="cuda:0")
    p = x.data_ptr()
    assert L.midint_add_u32(ctx.handle, 1, 0, p, p, p, None, None) == 1          # m32 == 0
    assert L.midint_mul_lo_u32(ctx.handle, 1, 4, p + 2, p, p, None) == 1         # not 4-byte aligned
    assert L.midint_mul_full_u32(ctx.handle, 1, 4097, p, p, p, None) == 2        # above 2048 limbs
