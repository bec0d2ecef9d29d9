"""Pins the CPU oracle before anything is checked against it.

Each oracle routine is checked against something independent of its own
code: libgmp's documented mpn functions (the library the reference validates
with, cuda/helper.h:146-170), Python's exact integers, the reference's own
known-answer vectors (cuda/main.cu:637-652, futhark-clean/fft-mul.fut:356-406)
and closed forms. A planted mistake in the oracle (dropped carry, wrong
twiddle sign, transposed operand) fails at least one of these.
"""
import numpy as np
import pytest

B = 1 << 64
MASK = B - 1


def ints(x, m):
    from oracle import to_int

    return [to_int(x[i * m:(i + 1) * m]) for i in range(len(x) // m)]


def rand_limbs(rng, n):
    return rng.integers(0, 2**64, size=n, dtype=np.uint64)


def carry_heavy(rng, n, m):
    """Limbs drawn mostly from {0, 2^64-1, 1} so carry chains run long."""
    pool = np.array([0, MASK, MASK, MASK, 1, MASK - 1], dtype=np.uint64)
    return pool[rng.integers(0, len(pool), size=n * m)]


@pytest.mark.parametrize("m", [1, 2, 3, 4, 7, 16, 64, 129])
def test_add_vs_python_and_gmp(orc, m):
    rng = np.random.default_rng(m)
    n = 37
    a = np.concatenate([rand_limbs(rng, n * m), carry_heavy(rng, n, m)])
    b = np.concatenate([rand_limbs(rng, n * m), carry_heavy(rng, n, m)])
    r, c = orc.add(a, b, m)
    for x, y, z, cc in zip(ints(a, m), ints(b, m), ints(r, m), c):
        s = x + y
        assert z == s % (B**m)
        assert cc == s >> (64 * m)
    if orc.gmp_available():
        rg, cg = orc.gmp_add(a, b, m)
        assert np.array_equal(r, rg) and np.array_equal(c, cg)


def test_add_all_ones_plus_one(orc):
    """The 1% rows of the addition workload: (B^m - 1) + 1 = 0, carry 1."""
    for m in (1, 8, 64):
        a = np.full(m, MASK, dtype=np.uint64)
        b = np.zeros(m, dtype=np.uint64)
        b[0] = 1
        r, c = orc.add(a, b, m)
        assert not r.any() and c[0] == 1


@pytest.mark.parametrize("m", [1, 2, 5, 8, 33])
def test_mul_vs_python_and_gmp(orc, m):
    rng = np.random.default_rng(100 + m)
    n = 19
    a = np.concatenate([rand_limbs(rng, n * m), np.full(m, MASK, dtype=np.uint64)])
    b = np.concatenate([rand_limbs(rng, n * m), np.full(m, MASK, dtype=np.uint64)])
    rf = orc.mul_full(a, b, m)
    rl = orc.mul_lo(a, b, m)
    for x, y, zf, zl in zip(ints(a, m), ints(b, m), ints(rf, 2 * m), ints(rl, m)):
        assert zf == x * y
        assert zl == (x * y) % (B**m)
    if orc.gmp_available():
        assert np.array_equal(rf, orc.gmp_mul_full(a, b, m))


def test_mul_closed_forms(orc):
    m = 6
    ones = np.full(m, MASK, dtype=np.uint64)
    # (B^m - 1)^2 mod B^m = 1
    lo = orc.mul_lo(ones, ones, m)
    assert lo[0] == 1 and not lo[1:].any()
    one = np.zeros(m, dtype=np.uint64)
    one[0] = 1
    a = np.arange(1, m + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    assert np.array_equal(orc.mul_lo(a, one, m), a)
    assert not orc.mul_full(a, np.zeros(m, dtype=np.uint64), m).any()


@pytest.mark.parametrize("na,nb", [(1, 1), (3, 9), (17, 4), (40, 40)])
def test_unbalanced_vs_gmp(orc, na, nb):
    rng = np.random.default_rng(na * 1000 + nb)
    a, b = rand_limbs(rng, na), rand_limbs(rng, nb)
    from oracle import to_int

    assert to_int(orc.mul_unbal(a, b)) == to_int(a) * to_int(b)
    assert to_int(orc.add_unbal(a, b)) == to_int(a) + to_int(b)
    if orc.gmp_available():
        assert np.array_equal(orc.mul_unbal(a, b), orc.gmp_mul(a, b))


@pytest.mark.parametrize("m", [1, 4, 16])
def test_add6_and_poly_closed_forms(orc, m):
    rng = np.random.default_rng(7 * m)
    n = 11
    a = np.concatenate([rand_limbs(rng, n * m), carry_heavy(rng, n, m)])
    b = np.concatenate([rand_limbs(rng, n * m), carry_heavy(rng, n, m)])
    r6 = orc.add6(a, b, m)
    rp = orc.poly(a, b, m)
    mod = B**m
    for x, y, z6, zp in zip(ints(a, m), ints(b, m), ints(r6, m), ints(rp, m)):
        assert z6 == (6 * x + 10 * y) % mod
        assert zp == ((x * x + y) * (y * y + y) + x * y) % mod


# ---------------------------------------------------------------- NTT
def test_ntt_forward_kat_p32(orc):
    """futhark-clean/fft-mul.fut:356-378: forward NTT over p32 = 3*2^30+1 with
    w = 13^(2^(30-4)) (getOmega, cuda/ker-fft-mul.cu.h:191-199)."""
    p = orc.P32
    w = pow(13, 1 << 26, p)
    x1 = [11400, 28374, 23152, 9576, 29511, 20787, 13067, 14015] + [0] * 8
    y1 = [149882, 2275842870, 1182190460, 2383281722, 1032368839, 2591113531,
          1041148619, 1666456763, 4378, 379675358, 2299045438, 2902395118,
          2188866018, 1972597972, 1919993985, 1934855231]
    x2 = [30268, 20788, 8033, 15446, 26275, 11619, 2494, 7016] + [0] * 8
    y2 = [121939, 1466303512, 2586179827, 2215991939, 1959268324, 3184250004,
          1699548738, 3026849220, 12201, 3020035466, 549183621, 389018726,
          1262049181, 799363395, 1607554732, 2004557247]
    assert [int(v) for v in orc.ntt(x1, w, p)] == y1
    assert [int(v) for v in orc.ntt(x2, w, p)] == y2


def test_ntt_conv_kat(orc):
    """cuda/main.cu:637-652 (= fft-mul.fut:390-406): the raw integer
    convolution, prime-independent, so it pins the Goldilocks path too."""
    a = [11400, 28374, 23152, 9576, 29511, 20787, 13067, 14015] + [0] * 8
    b = [30268, 20788, 8033, 15446, 26275, 11619, 2494, 7016] + [0] * 8
    ref = [345055200, 1095807432, 1382179648, 1175142886, 2016084656, 2555168834,
           2179032777, 1990011337, 1860865174, 1389799087, 942120918, 778961552,
           341270975, 126631482, 98329240, 0]
    assert [int(v) for v in orc.ntt_conv(a, b)] == ref


def test_goldilocks_roots(orc):
    p = orc.GL_P
    for lg in (1, 2, 4, 10, 14, 15, 32):
        n = 1 << lg
        w = orc.gl_root(n)
        assert pow(w, n, p) == 1 and pow(w, n // 2, p) == p - 1
    # the 4th root is 2^48: radix-4 butterflies multiply by a shift
    assert orc.gl_root(4) == 1 << 48
    assert pow(2, 96, p) == p - 1


def test_ntt_matches_dft_and_inverts(orc):
    p = orc.GL_P
    rng = np.random.default_rng(3)
    for n in (2, 8, 32, 64):
        x = [int(v) for v in rng.integers(0, 2**63, size=n, dtype=np.uint64)]
        w = orc.gl_root(n)
        dft = [sum(x[j] * pow(w, j * k, p) for j in range(n)) % p for k in range(n)]
        X = [int(v) for v in orc.ntt(x, w, p)]
        assert X == dft
        back = [int(v) for v in orc.ntt(X, pow(w, p - 2, p), p)]
        ninv = pow(n, p - 2, p)
        assert [v * ninv % p for v in back] == x


def test_ntt_conv_vs_schoolbook(orc):
    rng = np.random.default_rng(5)
    n = 64
    a = [int(v) for v in rng.integers(0, 1 << 16, size=n // 2)] + [0] * (n // 2)
    b = [int(v) for v in rng.integers(0, 1 << 16, size=n // 2)] + [0] * (n // 2)
    want = [sum(a[i] * b[k - i] for i in range(k + 1)) for k in range(n)]
    assert [int(v) for v in orc.ntt_conv(a, b)] == want


@pytest.mark.parametrize("m", [1, 2, 3, 16, 100])
def test_mul_ntt_vs_python_and_gmp(orc, m):
    rng = np.random.default_rng(900 + m)
    n = 3
    a = np.concatenate([rand_limbs(rng, n * m), np.full(m, MASK, dtype=np.uint64)])
    b = np.concatenate([rand_limbs(rng, n * m), np.full(m, MASK, dtype=np.uint64)])
    r = orc.mul_ntt(a, b, m)
    for x, y, z in zip(ints(a, m), ints(b, m), ints(r, 2 * m)):
        assert z == x * y
    if orc.gmp_available():
        assert np.array_equal(r, orc.gmp_mul_full(a, b, m))


# ---------------------------------------------------------------- generators
def test_splitmix64_published_vector(orc):
    # splitmix64 with state 0: first output 0xE220A8397B1DCDAF (Vigna's
    # reference implementation).
    assert orc.splitmix64_stream(0, 1)[0] == 0xE220A8397B1DCDAF


def _xoshiro_py(seed, count):
    st = [seed]

    def sm():
        st[0] = (st[0] + 0x9E3779B97F4A7C15) & MASK
        z = st[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    s = [sm() for _ in range(4)]
    rotl = lambda x, k: ((x << k) | (x >> (64 - k))) & MASK
    out = []
    for _ in range(count):
        out.append((rotl((s[1] * 5) & MASK, 7) * 9) & MASK)
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = rotl(s[3], 45)
    return out


def test_xoshiro_matches_restatement(orc):
    seed = 0x6D6964696E74
    assert [int(v) for v in orc.gen_limbs(seed, 50)] == _xoshiro_py(seed, 50)


def test_ref_rand_generator(orc):
    """ourMkRandom (cuda/helper.h:99-114) with glibc's default seed 1: the
    first rand() is 1804289383, and bit 0 of every word is 0."""
    w = orc.gen_ref_rand(1, 2, 8, 6)
    lo = (1804289383 * 2) & 0xFFFFFFFF
    assert int(w[0]) & 0xFFFF == lo & 0xFFFF
    assert not (w & 1).any()
    assert not w[6:8].any() and not w[14:16].any()


def test_gmp_ragged_matches_op_by_op(orc):
    """The GMP ragged baseline (mpn_mul / mpn_add op by op) equals the
    restatement's unbalanced products and sums, operands in either order."""
    if not orc.gmp_available():
        pytest.skip("libgmp.so.10 not installed")
    from workloads import make_ragged, ragged_reference

    ops, a, b, rl = make_ragged(orc.gen_limbs, n_ops=400, seed=11)
    assert np.array_equal(orc.gmp_ragged(ops, a, b, rl, threads=4), ragged_reference(orc, ops, a, b, rl))
