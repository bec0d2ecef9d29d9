"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/midint.h declares, reports errors without a GPU call, and its
host-side generators agree with the oracle's."""
import ctypes
import os
import subprocess

import numpy as np
import pytest


def test_exports_every_declared_symbol(mi):
    declared = mi.declared_symbols()
    assert len(declared) >= 20
    L = mi.lib()
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_declared_matches_binding_table(mi):
    # every declared function has a ctypes signature in the binding layer
    assert set(mi.declared_symbols()) == set(mi._SIGS)


def test_nm_dynamic_symbols(mi):
    out = subprocess.run(["nm", "-D", "--defined-only", mi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for s in mi.declared_symbols():
        assert s in names, s


def test_strerror_and_version(mi):
    L = mi.lib()
    assert L.midint_strerror(0) == b"success"
    assert L.midint_strerror(2) == b"operand size not supported"
    assert L.midint_strerror(99) == b"unknown status"
    assert b"gfx950" in L.midint_version()


def test_argument_errors_need_no_gpu(mi):
    L = mi.lib()
    # NULL context and m == 0 are rejected before any HIP call
    assert L.midint_add(None, 4, 8, None, None, None, None, None) == 1
    assert L.midint_mul_full(None, 4, 0, None, None, None, None) == 1
    assert L.midint_init(None, 0) == 1
    assert L.midint_destroy(None) == 1
    assert L.midint_gpu_add_host(0, 4, 0, None, None, None, 0, None) == 1


def test_size_limits(mi):
    L = mi.lib()
    assert L.midint_max_limbs_add() >= 4096
    assert L.midint_max_limbs_mul() >= 2048
    assert L.midint_max_limbs_ntt() >= 2048


def test_generators_match_oracle(mi, orc):
    seed = 0x6D6964696E74
    assert np.array_equal(mi.gen_limbs(seed, 1000), orc.gen_limbs(seed, 1000))
    assert np.array_equal(mi.gen_ref_rand(1, 3, 16, 16), orc.gen_ref_rand(1, 3, 16, 16))
    assert np.array_equal(mi.gen_ref_rand(7, 2, 8, 5), orc.gen_ref_rand(7, 2, 8, 5))


def test_harness_binary_built(mi):
    assert os.access(os.path.join(mi.HERE, "bin", "midint-harness"), os.X_OK)


def test_tuning_table(mi, tmp_path):
    L = mi.lib()
    old = mi.tune_get("add.rows")
    mi.tune_set("add.rows", 4)
    assert mi.tune_get("add.rows") == 4
    mi.tune_set("add.rows@64", 1)            # per size class
    assert mi.tune_get("add.rows@64") == 1
    assert L.midint_tune_set(b"add.rows@63", 1) == 1   # class must be a power of two
    assert L.midint_tune_set(b"no.such.key", 1) == 6
    f = tmp_path / "t.txt"
    f.write_text("# comment\nmul.q@128 4\nadd.nt 0\n")
    assert mi.tune_load(str(f))
    assert mi.tune_get("mul.q@128") == 4 and mi.tune_get("add.nt") == 0
    f.write_text("bogus.key 3\n")
    assert L.midint_tune_load(str(f).encode()) == 6
    mi.tune_set("add.rows", old)
    mi.tune_set("add.nt", 1)


def test_shipped_tuning_table_loads(mi):
    """Every key of the shipped table is one the library knows (an unknown
    key fails the load), including the matrix-core keys and the ragged
    graph's global and per-size-class keys."""
    path = os.path.join(mi.HERE, "tune", "mi355x.txt")
    keys = [ln.split()[0] for ln in open(path) if ln.strip() and not ln.startswith("#")]
    assert {"mul.mfma@128", "ragged.mfma_min", "ragged.overlap", "ragged.q@1024", "ragged.threads@1024",
            "ragged.mfma_waves@1024", "ntt.vals@16384"} <= set(keys)
    saved = {k: mi.tune_get(k) for k in ("ragged.mfma_min", "ragged.overlap", "mul.mfma_ipb")}
    assert mi.tune_load(path)
    for k, v in saved.items():
        mi.tune_set(k, v)


def test_new_entry_points_reject_bad_arguments(mi):
    """Argument checks of the u32, strided, CGBN, prime-NTT and poly_ntt entry
    points run before any HIP call."""
    L = mi.lib()
    # NULL context
    assert L.midint_add_u32(None, 1, 4, None, None, None, None, None) == 1
    assert L.midint_mul_full_strided(None, 1, 4, None, 0, None, 0, None, 0, None) == 1
    assert L.midint_cgbn(None, 0, 1, 64, None, None) == 1
    assert L.midint_ntt_transform(None, 0, 0, 1, 6, None, None, None) == 1
    assert L.midint_mul_ntt_prime(None, 1, 1, 4, None, None, None, None) == 1
    assert L.midint_poly_ntt(None, 1, 4, None, None, None, None) == 1
    # n == 0 is a no-op only with a context; m == 0 is always an error
    assert L.midint_poly_ntt(None, 0, 0, None, None, None, None) == 1


def test_declared_prime_and_layout_constants(mi):
    assert mi.PRIMES[mi.PRIME_P32] == 3 * 2**30 + 1
    assert mi.PRIMES[mi.PRIME_P64] == 29 * 2**57 + 1
    assert mi.PRIMES[mi.PRIME_GOLDILOCKS] == 2**64 - 2**32 + 1
    text = open(mi.HEADER).read()
    for name in ("MIDINT_PRIME_P32 = 1", "MIDINT_NTT_REFERENCE_ROOT = 2", "MIDINT_CGBN_POLY = 3"):
        assert name in text
