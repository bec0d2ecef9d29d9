"""Writes the golden fixtures under tests/golden/ from libgmp (the library the
reference validates against, cuda/helper.h:146-170) through the oracle's
loader of GMP's documented mpn functions.

  small_vectors.json : explicit inputs and GMP outputs for m in {1,2,3,4,8,16}
                       (add with carry-out, full product), plus edge cases
  digests.json       : SHA-256 of GMP's outputs on the three fixed-size
                       workloads (addition 65536 x 64 limbs, classical
                       multiplication 16384 x 128 limbs -> 256, NTT
                       multiplication 1024 x 2048 limbs -> 4096)

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle  # noqa: E402
from workloads import WORKLOADS, make_inputs  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
MASK = (1 << 64) - 1


def sha(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x).view(np.uint8).tobytes()).hexdigest()


def small_vectors():
    rng = np.random.default_rng(20240601)
    cases = []
    for m in (1, 2, 3, 4, 8, 16):
        for kind in ("random", "ones", "carry"):
            if kind == "random":
                a = rng.integers(0, 2**64, size=m, dtype=np.uint64)
                b = rng.integers(0, 2**64, size=m, dtype=np.uint64)
            elif kind == "ones":
                a = np.full(m, MASK, dtype=np.uint64)
                b = np.full(m, MASK, dtype=np.uint64)
            else:
                a = np.full(m, MASK, dtype=np.uint64)
                b = np.zeros(m, dtype=np.uint64)
                b[0] = 1
            s = np.empty(m, dtype=np.uint64)
            c = int(oracle.lib().orc_gmp_add_n(s.ctypes.data_as(oracle._u64p), a.ctypes.data_as(oracle._u64p),
                                               b.ctypes.data_as(oracle._u64p), m))
            p = oracle.gmp_mul(a, b)
            cases.append({"m": m, "kind": kind, "a": [str(int(v)) for v in a], "b": [str(int(v)) for v in b],
                          "sum": [str(int(v)) for v in s], "carry": c, "prod": [str(int(v)) for v in p]})
    return cases


def main():
    assert oracle.gmp_available(), "libgmp.so.10 is needed to write the fixtures"
    with open(os.path.join(OUT, "small_vectors.json"), "w") as f:
        json.dump({"source": "libgmp 6.2.1 mpn_add_n / mpn_mul", "cases": small_vectors()}, f, indent=0)

    digests = {}
    for name, w in WORKLOADS.items():
        a, b = make_inputs(name, oracle.gen_limbs)
        n, m = w["n"], w["m"]
        if w["op"] == "add":
            r = oracle.gmp_add_batch(a, b, m, threads=os.cpu_count())
        else:
            r = oracle.gmp_mul_full(a, b, m, threads=os.cpu_count())
        digests[name] = {"n": n, "m": m, "op": w["op"], "seed": w["seed"], "input_a": sha(a),
                         "input_b": sha(b), "result": sha(r), "result_limbs": int(r.size)}
        print(name, digests[name]["result"])
    with open(os.path.join(OUT, "digests.json"), "w") as f:
        json.dump({"source": "libgmp 6.2.1 mpn_add_n / mpn_mul on the workloads of tests/workloads.py",
                   "workloads": digests}, f, indent=1)


if __name__ == "__main__":
    main()
