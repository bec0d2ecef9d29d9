"""The fixed-size workloads of the benchmark and the golden fixtures.

  add_65536x64     : 65536 pairs of 4096-bit integers (64 u64 limbs), limbs
                     uniform from xoshiro256** seeded 0x6D6964696E74 ("midint"),
                     one stream for a then b; pairs i % 100 == 0 set to
                     a = 2^4096 - 1, b = 1 (full-length carry chains)
  mul_16384x128    : 16384 pairs of 8192-bit integers (128 limbs), seed + 1,
                     full 256-limb products
  ntt_1024x2048    : 1024 pairs of 131072-bit integers (2048 limbs), seed + 1,
                     full 4096-limb products through the NTT

`gen` is a function (seed, count) -> uint64 array; the library's
midint_gen_limbs and the oracle's orc_gen_limbs are the same generator
(checked in test_abi_cpu.py).
"""
import numpy as np

SEED = 0x6D6964696E74

WORKLOADS = {
    "add_65536x64": {"op": "add", "n": 65536, "m": 64, "seed": SEED},
    "mul_16384x128": {"op": "mul_full", "n": 16384, "m": 128, "seed": SEED + 1},
    "ntt_1024x2048": {"op": "mul_ntt", "n": 1024, "m": 2048, "seed": SEED + 1},
}


def make_inputs(name, gen):
    w = WORKLOADS[name]
    n, m = w["n"], w["m"]
    s = gen(w["seed"], 2 * n * m)
    a = np.ascontiguousarray(s[: n * m])
    b = np.ascontiguousarray(s[n * m:])
    if w["op"] == "add":
        a2 = a.reshape(n, m)
        b2 = b.reshape(n, m)
        a2[::100, :] = np.uint64(0xFFFFFFFFFFFFFFFF)
        b2[::100, :] = np.uint64(0)
        b2[::100, 0] = np.uint64(1)
    return a, b


# ---------------------------------------------------------------- ragged batch
def make_ragged(gen, n_ops=50000, seed=SEED + 2, mul_frac=0.7, lo_bits=8, hi_bits=16):
    """Ragged batch: each op is a multiplication with probability mul_frac,
    else an addition; each operand size is max(4, round(2^U(lo_bits, hi_bits)
    / 64)) limbs, drawn independently per operand (256..65536 bits for the
    defaults). Kinds and sizes come from numpy's PCG64 seeded with `seed`,
    limbs from `gen(seed, count)`.

    Returns (ops, a, b, r_limbs) where ops is a list of
    (kind, n_a, n_b, off_a, off_b, off_r) with kind 0 = add, 1 = mul, a and b
    the packed operand buffers, r_limbs the size of the packed result."""
    rng = np.random.default_rng(seed)
    kinds = (rng.random(n_ops) < mul_frac).astype(np.uint32)
    sizes = np.maximum(4, np.rint(2.0 ** rng.uniform(lo_bits, hi_bits, size=(n_ops, 2)) / 64)).astype(np.int64)
    na, nb = sizes[:, 0], sizes[:, 1]
    off_a = np.concatenate([[0], np.cumsum(na)[:-1]])
    off_b = np.concatenate([[0], np.cumsum(nb)[:-1]])
    rl = np.where(kinds == 1, na + nb, np.maximum(na, nb) + 1)
    off_r = np.concatenate([[0], np.cumsum(rl)[:-1]])
    limbs = gen(seed, int(na.sum() + nb.sum()))
    a = np.ascontiguousarray(limbs[: na.sum()])
    b = np.ascontiguousarray(limbs[na.sum():])
    ops = [(int(kinds[i]), int(na[i]), int(nb[i]), int(off_a[i]), int(off_b[i]), int(off_r[i]))
           for i in range(n_ops)]
    return ops, a, b, int(rl.sum())


def ragged_reference(orc, ops, a, b, r_limbs):
    """Per-op oracle results in the packed result layout."""
    r = np.zeros(r_limbs, dtype=np.uint64)
    for kind, na, nb, oa, ob, orr in ops:
        x, y = a[oa:oa + na], b[ob:ob + nb]
        if kind == 1:
            r[orr:orr + na + nb] = orc.mul_unbal(x, y)
        else:
            r[orr:orr + max(na, nb) + 1] = orc.add_unbal(x, y)
    return r


def size_class(n):
    """Power-of-two size class (>= 4 limbs) of the plain per-size launches."""
    c = 4
    while c < n:
        c *= 2
    return c


# ---------------------------------------------------------------- serving stream
def make_requests(gen, n, seed=SEED + 3, mul_frac=0.7, lo_bits=9, hi_bits=18):
    """Serving requests: kind mul with probability mul_frac, operand sizes
    max(8, round(2^U(lo_bits, hi_bits) / 64)) limbs (512..262144 bits for the
    defaults), independent per operand; limbs from gen(seed, count)."""
    rng = np.random.default_rng(seed)
    kinds = (rng.random(n) < mul_frac).astype(int)
    sizes = np.maximum(8, np.rint(2.0 ** rng.uniform(lo_bits, hi_bits, size=(n, 2)) / 64)).astype(np.int64)
    limbs = gen(seed, int(sizes.sum()))
    out, pos = [], 0
    for i in range(n):
        na, nb = int(sizes[i, 0]), int(sizes[i, 1])
        a = np.ascontiguousarray(limbs[pos:pos + na])
        pos += na
        b = np.ascontiguousarray(limbs[pos:pos + nb])
        pos += nb
        out.append((int(kinds[i]), a, b))
    return out


def request_reference(orc, kind, a, b):
    return orc.mul_unbal(a, b) if kind == 1 else orc.add_unbal(a, b)
