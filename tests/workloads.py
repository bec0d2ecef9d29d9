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
