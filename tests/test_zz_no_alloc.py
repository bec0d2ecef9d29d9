"""midint_mul_lo_u32_noalloc (static per-device scratch, no allocation)
against midint_mul_lo_u32 (context scratch) through the C ABI: two calls in
a row, the second smaller (the array keeps the first call's contents), the
u64-layout case that needs no scratch, and the size limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ETOOBIG = 7
SCRATCH_LIMBS = 6 << 20


@pytest.fixture(scope="module")
def torch_dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _words(seed, n, m32):
    rng = np.random.default_rng(seed)
    w = rng.integers(0, 2**32, size=n * m32, dtype=np.uint64).astype(np.uint32)
    w[:m32] = 0xFFFFFFFF  # instance 0: all-ones operand
    return w


def _dev32(torch, w, offset=0):
    """int32 device buffer holding w at word `offset` (offset 1: 4-byte but
    not 8-byte aligned), and the view that starts there."""
    buf = torch.zeros(w.size + offset, dtype=torch.int32, device="cuda:0")
    buf[offset:] = torch.from_numpy(w.view(np.int32)).to("cuda:0")
    return buf, buf.data_ptr() + 4 * offset


def _run(torch, L, ctx, fn, n, m32, wa, wb, offset):
    da, pa = _dev32(torch, wa, offset)
    db, pb = _dev32(torch, wb, offset)
    dr, pr = _dev32(torch, np.zeros(n * m32, dtype=np.uint32), offset)
    st = fn(ctx.handle, n, m32, ctypes.c_void_p(pa), ctypes.c_void_p(pb), ctypes.c_void_p(pr), None)
    assert st == 0, (fn.__name__, st)
    torch.cuda.synchronize()
    return dr[offset:].cpu().numpy().view(np.uint32).copy()


def _want(orc, wa, wb, n, m32):
    m = (m32 + 1) // 2
    za = np.zeros((n, 2 * m), dtype=np.uint32)
    zb = np.zeros((n, 2 * m), dtype=np.uint32)
    za[:, :m32] = wa.reshape(n, m32)
    zb[:, :m32] = wb.reshape(n, m32)
    p = orc.mul_lo(za.reshape(-1).view(np.uint64), zb.reshape(-1).view(np.uint64), m)
    return p.view(np.uint32).reshape(n, 2 * m)[:, :m32].reshape(-1)


@pytest.mark.parametrize("offset", [0, 1])
def test_noalloc_matches_context_scratch_path(mi, torch_dev, orc, offset):
    torch = torch_dev
    L = mi.lib()
    ctx = mi.Context(0)
    try:
        # (n, m32): the first call fills most of the static array, the second
        # is smaller; odd word counts (repack), then an even one (u64 layout
        # at offset 0, repack at offset 1)
        for call, (n, m32) in enumerate([(3000, 257), (700, 65), (500, 130)]):
            assert 3 * n * ((m32 + 1) // 2) <= SCRATCH_LIMBS
            wa, wb = _words(100 * call + offset, n, m32), _words(100 * call + offset + 1, n, m32)
            got = _run(torch, L, ctx, L.midint_mul_lo_u32_noalloc, n, m32, wa, wb, offset)
            ref = _run(torch, L, ctx, L.midint_mul_lo_u32, n, m32, wa, wb, offset)
            assert np.array_equal(got, ref), (call, n, m32, offset)
            assert np.array_equal(got, _want(orc, wa, wb, n, m32)), (call, n, m32, offset)
    finally:
        ctx.close()


def test_noalloc_size_limit(mi, torch_dev):
    torch = torch_dev
    L = mi.lib()
    ctx = mi.Context(0)
    try:
        x = torch.zeros(64, dtype=torch.int32, device="cuda:0")
        p = ctypes.c_void_p(x.data_ptr())
        # 3 * 100000 * 128 limbs exceed the array: rejected before any launch
        assert L.midint_mul_lo_u32_noalloc(ctx.handle, 100000, 255, p, p, p, None) == ETOOBIG
        assert L.midint_strerror(ETOOBIG) == b"request exceeds the fixed-size static scratch"
        assert L.midint_mul_lo_u32_noalloc(ctx.handle, 1, 0, p, p, p, None) == 1
        assert L.midint_mul_lo_u32_noalloc(ctx.handle, 1, 4097, p, p, p, None) == 2  # above 2048 limbs
    finally:
        ctx.close()
