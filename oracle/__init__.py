"""CPU oracle for the midint HIP library -- TEST INFRASTRUCTURE ONLY.

numpy front end over ``liboracle.so`` (built from ``midint_oracle.c`` by
``oracle/Makefile``). Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product
library never does.

Arrays are numpy ``uint64`` with batch-major layout: instance i of an
n x m batch occupies ``x[i*m:(i+1)*m]``, limb 0 least significant
(reference cuda/ker-helpers.cu.h:12-19, cuda/helper.h:116).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

GL_P = 0xFFFFFFFF00000001
P32 = 3221225473  # reference FftPrime32::p (cuda/ker-fft-mul.cu.h:69)

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_sz = ctypes.c_size_t


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_add_n": (ctypes.c_uint32, [_u64p, _u64p, _u64p, _sz]),
            "orc_add_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p, _u8p]),
            "orc_add6_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p]),
            "orc_mul_full": (None, [_u64p, _u64p, _u64p, _sz]),
            "orc_mul_lo": (None, [_u64p, _u64p, _u64p, _sz]),
            "orc_mul_full_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p]),
            "orc_mul_full_batch_mt": (None, [_sz, _sz, _u64p, _u64p, _u64p, ctypes.c_int]),
            "orc_mul_lo_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p]),
            "orc_poly_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p]),
            "orc_mul_unbal": (None, [_u64p, _u64p, _sz, _u64p, _sz]),
            "orc_add_unbal": (None, [_u64p, _u64p, _sz, _u64p, _sz]),
            "orc_mulmod": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "orc_powmod": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64]),
            "orc_ntt": (None, [_u64p, _sz, ctypes.c_uint64, ctypes.c_uint64]),
            "orc_gl_root": (ctypes.c_uint64, [_sz]),
            "orc_ntt_conv": (None, [_u64p, _u64p, _u64p, _sz]),
            "orc_mul_ntt": (None, [_u64p, _u64p, _u64p, _sz]),
            "orc_mul_ntt_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p]),
            "orc_splitmix64": (ctypes.c_uint64, [_u64p]),
            "orc_gen_limbs": (None, [ctypes.c_uint64, _u64p, _sz]),
            "orc_gen_ref_rand": (None, [ctypes.c_uint32, _sz, _sz, _sz, _u32p]),
            "orc_gmp_available": (ctypes.c_int, []),
            "orc_gmp_add_n": (ctypes.c_uint64, [_u64p, _u64p, _u64p, _sz]),
            "orc_gmp_mul": (None, [_u64p, _u64p, _sz, _u64p, _sz]),
            "orc_gmp_add_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p, ctypes.c_int]),
            "orc_gmp_mul_full_batch": (None, [_sz, _sz, _u64p, _u64p, _u64p, ctypes.c_int]),
            "orc_gmp_ragged": (None, [_sz, _u32p, _u32p, _u32p, _u64p, _u64p, _u64p, _u64p, _u64p, _u64p,
                                      ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p64(x: np.ndarray):
    assert x.dtype == np.uint64 and x.flags.c_contiguous
    return x.ctypes.data_as(_u64p)


def _c(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


# ---------------------------------------------------------------- arithmetic
def add(a, b, m):
    """(a + b) mod B^m per instance; returns (r, carry_out)."""
    a, b = _c(a), _c(b)
    n = a.size // m
    r = np.empty_like(a)
    c = np.empty(n, dtype=np.uint8)
    lib().orc_add_batch(n, m, _p64(a), _p64(b), _p64(r), c.ctypes.data_as(_u8p))
    return r, c


def add6(a, b, m):
    a, b = _c(a), _c(b)
    r = np.empty_like(a)
    lib().orc_add6_batch(a.size // m, m, _p64(a), _p64(b), _p64(r))
    return r


def mul_full(a, b, m, threads=1):
    a, b = _c(a), _c(b)
    n = a.size // m
    r = np.empty(2 * a.size, dtype=np.uint64)
    if threads > 1:
        lib().orc_mul_full_batch_mt(n, m, _p64(a), _p64(b), _p64(r), threads)
    else:
        lib().orc_mul_full_batch(n, m, _p64(a), _p64(b), _p64(r))
    return r


def mul_lo(a, b, m):
    a, b = _c(a), _c(b)
    r = np.empty_like(a)
    lib().orc_mul_lo_batch(a.size // m, m, _p64(a), _p64(b), _p64(r))
    return r


def poly(a, b, m):
    a, b = _c(a), _c(b)
    r = np.empty_like(a)
    lib().orc_poly_batch(a.size // m, m, _p64(a), _p64(b), _p64(r))
    return r


def mul_unbal(a, b):
    a, b = _c(a), _c(b)
    r = np.empty(a.size + b.size, dtype=np.uint64)
    lib().orc_mul_unbal(_p64(r), _p64(a), a.size, _p64(b), b.size)
    return r


def add_unbal(a, b):
    a, b = _c(a), _c(b)
    r = np.empty(max(a.size, b.size) + 1, dtype=np.uint64)
    lib().orc_add_unbal(_p64(r), _p64(a), a.size, _p64(b), b.size)
    return r


def mul_ntt(a, b, m):
    a, b = _c(a), _c(b)
    r = np.empty(2 * a.size, dtype=np.uint64)
    lib().orc_mul_ntt_batch(a.size // m, m, _p64(a), _p64(b), _p64(r))
    return r


def ntt(x, w, p):
    x = _c(x).copy()
    lib().orc_ntt(_p64(x), x.size, w, p)
    return x


def gl_root(n):
    return int(lib().orc_gl_root(n))


def ntt_conv(a, b):
    a, b = _c(a), _c(b)
    out = np.empty_like(a)
    lib().orc_ntt_conv(_p64(a), _p64(b), _p64(out), a.size)
    return out


def mulmod(a, b, p):
    return int(lib().orc_mulmod(a, b, p))


def powmod(a, e, p):
    return int(lib().orc_powmod(a, e, p))


# ---------------------------------------------------------------- generators
def gen_limbs(seed, count):
    out = np.empty(count, dtype=np.uint64)
    lib().orc_gen_limbs(seed, _p64(out), count)
    return out


def splitmix64_stream(seed, count):
    st = ctypes.c_uint64(seed)
    return [int(lib().orc_splitmix64(ctypes.byref(st))) for _ in range(count)]


def gen_ref_rand(seed, n, m_u32, nz):
    out = np.empty(n * m_u32, dtype=np.uint32)
    lib().orc_gen_ref_rand(seed, n, m_u32, nz, out.ctypes.data_as(_u32p))
    return out


# ---------------------------------------------------------------- GMP
def gmp_available() -> bool:
    return bool(lib().orc_gmp_available())


def gmp_add(a, b, m):
    a, b = _c(a), _c(b)
    n = a.size // m
    r = np.empty_like(a)
    c = np.empty(n, dtype=np.uint8)
    for i in range(n):
        s = slice(i * m, (i + 1) * m)
        c[i] = lib().orc_gmp_add_n(r[s].ctypes.data_as(_u64p), a[s].ctypes.data_as(_u64p),
                                   b[s].ctypes.data_as(_u64p), m)
    return r, c


def gmp_mul_full(a, b, m, threads=1):
    a, b = _c(a), _c(b)
    r = np.empty(2 * a.size, dtype=np.uint64)
    lib().orc_gmp_mul_full_batch(a.size // m, m, _p64(a), _p64(b), _p64(r), threads)
    return r


def gmp_add_batch(a, b, m, threads=1):
    a, b = _c(a), _c(b)
    r = np.empty_like(a)
    lib().orc_gmp_add_batch(a.size // m, m, _p64(a), _p64(b), _p64(r), threads)
    return r


def gmp_mul(a, b):
    a, b = _c(a), _c(b)
    r = np.empty(a.size + b.size, dtype=np.uint64)
    lib().orc_gmp_mul(_p64(r), _p64(a), a.size, _p64(b), b.size)
    return r


# ---------------------------------------------------------------- helpers
def to_int(limbs) -> int:
    """Integer value of a little-endian u64 limb vector."""
    v = 0
    for x in reversed([int(t) for t in np.asarray(limbs, dtype=np.uint64)]):
        v = (v << 64) | x
    return v


def from_int(v: int, m: int) -> np.ndarray:
    out = np.empty(m, dtype=np.uint64)
    for i in range(m):
        out[i] = v & 0xFFFFFFFFFFFFFFFF
        v >>= 64
    return out


def gmp_ragged(ops, a, b, r_limbs, threads=1, out=None):
    """A ragged batch [(kind, n_a, n_b, off_a, off_b, off_r)] op by op with
    GMP (mpn_mul / mpn_add), results in the packed layout. `ops` may be the
    prepared tuple of gmp_ragged_prepare and `out` a result buffer of
    r_limbs limbs to reuse (both keep conversions out of timed loops)."""
    if not isinstance(ops, tuple):
        ops = gmp_ragged_prepare(ops)
    kind, na, nb, oa, ob, orr = ops
    a, b = _c(a), _c(b)
    r = np.zeros(r_limbs, dtype=np.uint64) if out is None else out
    lib().orc_gmp_ragged(kind.size, kind.ctypes.data_as(_u32p), na.ctypes.data_as(_u32p),
                         nb.ctypes.data_as(_u32p), _p64(oa), _p64(ob), _p64(orr), _p64(a), _p64(b), _p64(r),
                         threads)
    return r


def gmp_ragged_prepare(ops):
    """Column arrays of a ragged op list for gmp_ragged."""
    ops = np.asarray(ops, dtype=np.uint64).reshape(-1, 6)
    kind = np.ascontiguousarray(ops[:, 0].astype(np.uint32))
    na = np.ascontiguousarray(ops[:, 1].astype(np.uint32))
    nb = np.ascontiguousarray(ops[:, 2].astype(np.uint32))
    oa, ob, orr = (np.ascontiguousarray(ops[:, k]) for k in (3, 4, 5))
    return kind, na, nb, oa, ob, orr
