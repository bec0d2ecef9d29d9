"""The oracle against the committed GMP fixtures (tests/golden/)."""
import hashlib
import json
import os

import numpy as np
import pytest

from workloads import WORKLOADS, make_inputs

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(x):
    return hashlib.sha256(np.ascontiguousarray(x).view(np.uint8).tobytes()).hexdigest()


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_small_vectors(orc):
    for c in load("small_vectors.json")["cases"]:
        m = c["m"]
        a = np.array([int(v) for v in c["a"]], dtype=np.uint64)
        b = np.array([int(v) for v in c["b"]], dtype=np.uint64)
        s, carry = orc.add(a, b, m)
        assert [int(v) for v in s] == [int(v) for v in c["sum"]] and int(carry[0]) == c["carry"]
        assert [int(v) for v in orc.mul_full(a, b, m)] == [int(v) for v in c["prod"]]
        assert [int(v) for v in orc.mul_ntt(a, b, m)] == [int(v) for v in c["prod"]]
        assert [int(v) for v in orc.mul_lo(a, b, m)] == [int(v) for v in c["prod"]][:m]


def test_workload_inputs_digest(orc):
    d = load("digests.json")["workloads"]
    for name in WORKLOADS:
        a, b = make_inputs(name, orc.gen_limbs)
        assert sha(a) == d[name]["input_a"] and sha(b) == d[name]["input_b"], name


def test_oracle_add_workload_digest(orc):
    d = load("digests.json")["workloads"]["add_65536x64"]
    a, b = make_inputs("add_65536x64", orc.gen_limbs)
    r, c = orc.add(a, b, 64)
    assert sha(r) == d["result"]
    # the forced rows (i % 100 == 0) are all-ones + 1: zero with carry out
    assert not r.reshape(-1, 64)[::100].any() and c[::100].all()


def test_oracle_mul_workload_digest(orc):
    d = load("digests.json")["workloads"]["mul_16384x128"]
    a, b = make_inputs("mul_16384x128", orc.gen_limbs)
    r = orc.mul_full(a, b, 128, threads=os.cpu_count() or 1)
    assert sha(r) == d["result"]


def test_oracle_ntt_on_workload_sample(orc):
    """Oracle NTT product on the first instances of the 2048-limb workload
    equals GMP's product (the full digest is checked on the GPU)."""
    if not orc.gmp_available():
        pytest.skip("libgmp.so.10 not present")
    a, b = make_inputs("ntt_1024x2048", orc.gen_limbs)
    m = 2048
    k = 3
    r = orc.mul_ntt(a[: k * m], b[: k * m], m)
    assert np.array_equal(r, orc.gmp_mul_full(a[: k * m], b[: k * m], m))
