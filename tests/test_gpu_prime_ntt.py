"""NTT over the reference's primes (p32 = 3*2^30+1, p64 = 29*2^57+1) and
Goldilocks through the C ABI: the reference's transform (root from getOmega,
cuda/ker-fft-mul.cu.h:191-199) against the futhark-clean/fft-mul.fut:356-378
known-answer vectors and the oracle's restatement of fft (:304-348); true
NTTs (primitive roots) against the oracle and their inverse; exact products
against GMP. Bit-exact."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = {0: 0xFFFFFFFF00000001, 1: 3221225473, 2: 4179340454199820289}
G = {0: 7, 1: 13, 2: 21}
TWO_ADIC = {0: 32, 1: 30, 2: 57}


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
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.uint64)).view(np.int64)).to("cuda:0")


def host(t):
    return t.cpu().numpy().view(np.uint64)


def root(prime, lg, reference):
    p = P[prime]
    if reference:
        return pow(G[prime], 1 << (TWO_ADIC[prime] - lg), p)
    return pow(G[prime], (p - 1) >> lg, p)


def transform(torch, ctx, prime, lg, xs, inverse=False, reference=False):
    x = dev(torch, np.concatenate(xs))
    y = torch.empty_like(x)
    ctx.ntt_transform(x, y, lg, prime=prime, inverse=inverse, reference_root=reference)
    torch.cuda.synchronize()
    return host(y).reshape(len(xs), -1)


def test_reference_transform_kat_p32(ctx, torch_dev):
    x1 = [11400, 28374, 23152, 9576, 29511, 20787, 13067, 14015] + [0] * 8
    y1 = [149882, 2275842870, 1182190460, 2383281722, 1032368839, 2591113531, 1041148619, 1666456763,
          4378, 379675358, 2299045438, 2902395118, 2188866018, 1972597972, 1919993985, 1934855231]
    x2 = [30268, 20788, 8033, 15446, 26275, 11619, 2494, 7016] + [0] * 8
    y2 = [121939, 1466303512, 2586179827, 2215991939, 1959268324, 3184250004, 1699548738, 3026849220,
          12201, 3020035466, 549183621, 389018726, 1262049181, 799363395, 1607554732, 2004557247]
    out = transform(torch_dev, ctx, 1, 4, [x1, x2], reference=True)
    assert [int(v) for v in out[0]] == y1
    assert [int(v) for v in out[1]] == y2


@pytest.mark.parametrize("prime", [1, 2])
@pytest.mark.parametrize("lg", [1, 5, 6, 11, 14])
def test_reference_transform_vs_oracle(ctx, torch_dev, orc, prime, lg):
    rng = np.random.default_rng(prime * 100 + lg)
    p = P[prime]
    xs = [rng.integers(0, p, size=1 << lg, dtype=np.uint64) for _ in range(3)]
    xs[1][:] = p - 1
    out = transform(torch_dev, ctx, prime, lg, xs, reference=True)
    w = root(prime, lg, True)
    for x, y in zip(xs, out):
        assert np.array_equal(y, orc.ntt(x, w, p))


@pytest.mark.parametrize("prime", [0, 1, 2])
@pytest.mark.parametrize("lg", [3, 6, 10, 14])
def test_primitive_transform_and_inverse(ctx, torch_dev, orc, prime, lg):
    rng = np.random.default_rng(prime * 1000 + lg)
    p = P[prime]
    w = root(prime, lg, False)
    assert pow(w, 1 << (lg - 1), p) == p - 1  # primitive 2^lg-th root
    xs = [rng.integers(0, p, size=1 << lg, dtype=np.uint64) for _ in range(2)]
    xs.append(np.full(1 << lg, p - 1, dtype=np.uint64))
    fw = transform(torch_dev, ctx, prime, lg, xs)
    for x, y in zip(xs, fw):
        assert np.array_equal(y, orc.ntt(x, w, p))
    back = transform(torch_dev, ctx, prime, lg, list(fw), inverse=True)
    for x, z in zip(xs, back):
        assert np.array_equal(z, x)


@pytest.mark.parametrize("prime,m", [(1, 1), (1, 3), (1, 64), (1, 100), (1, 1024),
                                     (2, 1), (2, 5), (2, 128), (2, 1000), (2, 2048), (0, 77)])
def test_mul_ntt_prime_vs_gmp(ctx, torch_dev, orc, prime, m):
    n = 3 if m >= 1000 else 6
    rng = np.random.default_rng(m * 7 + prime)
    a = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=n * m, dtype=np.uint64)
    a[:m] = np.uint64(0xFFFFFFFFFFFFFFFF)
    b[:m] = np.uint64(0xFFFFFFFFFFFFFFFF)
    r = torch_dev.empty(2 * n * m, dtype=torch_dev.int64, device="cuda:0")
    ctx.mul_ntt_prime(dev(torch_dev, a), dev(torch_dev, b), r, m, prime)
    torch_dev.cuda.synchronize()
    assert np.array_equal(host(r), orc.mul_full(a, b, m))


def test_prime_errors(mi, ctx, torch_dev):
    L = mi.lib()
    x = torch_dev.zeros(64, dtype=torch_dev.int64, device="cuda:0")
    # the reference root is only defined for the reference's primes
    assert L.midint_ntt_transform(ctx.handle, 0, 2, 1, 6, x.data_ptr(), x.data_ptr(), None) == 1
    assert L.midint_ntt_transform(ctx.handle, 7, 0, 1, 6, x.data_ptr(), x.data_ptr(), None) == 1
    assert L.midint_ntt_transform(ctx.handle, 1, 0, 1, 15, x.data_ptr(), x.data_ptr(), None) == 2
    # p32 products are exact up to 1024 limbs (8-bit coefficients, 2^14 points)
    assert L.midint_mul_ntt_prime(ctx.handle, 1, 1, 1025, x.data_ptr(), x.data_ptr(), x.data_ptr(), None) == 2
