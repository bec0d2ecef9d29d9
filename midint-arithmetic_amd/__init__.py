"""Python view of libmidint's C ABI (include/midint.h).

This is a thin ctypes layer used by the test-suite and bench.py; the product
is the C ABI itself. Device buffers are torch tensors (dtype int64/uint64,
contiguous, on a HIP device): torch provides allocation and streams only.

The directory name is not a valid Python identifier; load it with
``load_package()`` from ``__graft_entry__`` or ``importlib``:

    pkg = importlib.util.spec_from_file_location(
        "midint_arithmetic_amd", ".../midint-arithmetic_amd/__init__.py",
        submodule_search_locations=[".../midint-arithmetic_amd"])

Every call goes to the HIP kernels; there is no CPU fallback. If the shared
library is missing, ``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmidint.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "midint.h")

OK = 0
STATUS = {0: "OK", 1: "EINVAL", 2: "EUNSUPPORTED", 3: "EHIP", 4: "ENOMEM", 5: "EBUSY", 6: "ENOTFOUND"}

_lib = None
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_dblp = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "midint_strerror": (ctypes.c_char_p, [_int]),
    "midint_version": (ctypes.c_char_p, []),
    "midint_last_hip_error": (_int, []),
    "midint_init": (_int, [ctypes.POINTER(_vp), _int]),
    "midint_destroy": (_int, [_vp]),
    "midint_max_limbs_add": (_u32, []),
    "midint_max_limbs_mul": (_u32, []),
    "midint_max_limbs_ntt": (_u32, []),
    "midint_add": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    "midint_add6": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_lo": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_full": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_poly": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_ntt": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_ntt_conv": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_ntt_transform": (_int, [_vp, _u32, _u32, _u32, _u32, _vp, _vp, _vp]),
    "midint_poly_ntt": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_add_strided": (_int, [_vp, _u32, _u32, _vp, _u64, _vp, _u64, _vp, _u64, _vp, _vp]),
    "midint_add6_strided": (_int, [_vp, _u32, _u32, _vp, _u64, _vp, _u64, _vp, _u64, _vp]),
    "midint_mul_lo_strided": (_int, [_vp, _u32, _u32, _vp, _u64, _vp, _u64, _vp, _u64, _vp]),
    "midint_mul_full_strided": (_int, [_vp, _u32, _u32, _vp, _u64, _vp, _u64, _vp, _u64, _vp]),
    "midint_poly_strided": (_int, [_vp, _u32, _u32, _vp, _u64, _vp, _u64, _vp, _u64, _vp]),
    "midint_cgbn": (_int, [_vp, _u32, _u32, _u32, _vp, _vp]),
    "midint_add_u32": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    "midint_add6_u32": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_lo_u32": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_full_u32": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_poly_u32": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_mul_lo_u32_noalloc": (_int, [_vp, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_probe_stream": (_int, [_u64, _vp, _vp, _vp, _vp]),
    "midint_mul_ntt_prime": (_int, [_vp, _u32, _u32, _u32, _vp, _vp, _vp, _vp]),
    "midint_ntt_prepare": (_int, [_vp, _u32]),
    "midint_gpu_add_host": (_int, [_int, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_add6_host": (_int, [_int, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_poly_host": (_int, [_int, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_poly_fft_host": (_int, [_int, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_mul_host": (_int, [_int, _u32, _u32, _int, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_mul_fft_host": (_int, [_int, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_gpu_mul_fft_prime_host": (_int, [_int, _u32, _u32, _u32, _vp, _vp, _vp, _u32, _dblp]),
    "midint_probe_valu": (_int, [_int, _dblp, _dblp]),
    "midint_probe_mac": (_int, [_int, _dblp]),
    "midint_probe_mfma": (_int, [_int, _dblp, _dblp]),
    "midint_max_limbs_ragged": (_u32, []),
    "midint_ragged_plan_create": (_int, [_vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "midint_ragged_launch": (_int, [_vp, _vp]),
    "midint_ragged_plan_stats": (_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "midint_ragged_plan_destroy": (_int, [_vp]),
    "midint_server_create": (_int, [_vp, _vp, ctypes.POINTER(_vp)]),
    "midint_server_submit": (_int, [_vp, _u32, _vp, _u32, _vp, _u32, ctypes.POINTER(ctypes.c_uint64)]),
    "midint_server_poll": (_int, [_vp, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]),
    "midint_server_result": (_int, [_vp, ctypes.c_uint64, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                    _dblp]),
    "midint_server_submit_many": (_int, [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp, _vp, _vp]),
    "midint_server_result_many": (_int, [_vp, ctypes.c_size_t, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp]),
    "midint_server_stats": (_int, [_vp, _vp]),
    "midint_server_destroy": (_int, [_vp]),
    "midint_tune_set": (_int, [ctypes.c_char_p, ctypes.c_int64]),
    "midint_tune_get": (_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "midint_tune_load": (_int, [ctypes.c_char_p]),
    "midint_gen_limbs": (None, [ctypes.c_uint64, _vp, ctypes.c_size_t]),
    "midint_gen_ref_rand": (None, [_u32, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t, _vp]),
}


class MidintError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        name = STATUS.get(status, str(status))
        extra = ""
        if status == 3 and _lib is not None:
            extra = f" (hip error {_lib.midint_last_hip_error()})"
        super().__init__(f"{what}: {name}{extra}")


def lib():
    """The loaded libmidint.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it first (make -C {HERE})")
        # torch bundles its own HIP runtime (SONAME libamdhip64.so.7, loaded by
        # torch under the unversioned name). Loading torch first makes
        # libmidint's dependency bind to that same runtime instead of a second
        # copy from /opt/rocm, which would not see the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header: str = HEADER):
    """Function names declared in include/midint.h."""
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(midint_[a-z0-9_]+)\s*\(", text)))


def check(status: int, what: str) -> None:
    if status != OK:
        raise MidintError(status, what)


def _ptr(t):
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        if not t.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)


def _stream(stream):
    if stream is None:
        try:
            import torch

            return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        except Exception:  # pragma: no cover - no torch / no device
            return None
    if isinstance(stream, int):
        return ctypes.c_void_p(stream)
    return ctypes.c_void_p(stream.cuda_stream)


PRIME_GOLDILOCKS, PRIME_P32, PRIME_P64 = 0, 1, 2
PRIMES = {PRIME_GOLDILOCKS: 0xFFFFFFFF00000001, PRIME_P32: 3221225473, PRIME_P64: 4179340454199820289}
NTT_INVERSE, NTT_REFERENCE_ROOT = 1, 2
CGBN_ADD, CGBN_ADD6, CGBN_MUL, CGBN_POLY = 0, 1, 2, 3


class Context:
    """Owns a midint_ctx (device binding, NTT twiddle cache, scratch)."""

    def __init__(self, device: int = 0, tuned: bool = True):
        self._h = ctypes.c_void_p()
        check(lib().midint_init(ctypes.byref(self._h), device), "midint_init")
        self.device = device
        if tuned:
            tune_load()

    def close(self):
        if self._h:
            lib().midint_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ ops
    def add(self, a, b, r, m, carry=None, stream=None):
        n = a.numel() // m
        check(lib().midint_add(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _ptr(carry), _stream(stream)),
              "midint_add")

    def add6(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_add6(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_add6")

    def mul_lo(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_mul_lo(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_mul_lo")

    def mul_full(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_mul_full(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_mul_full")

    def poly(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_poly(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_poly")

    def mul_ntt(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_mul_ntt(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_mul_ntt")

    def poly_ntt(self, a, b, r, m, stream=None):
        n = a.numel() // m
        check(lib().midint_poly_ntt(self._h, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_poly_ntt")

    def ntt_conv(self, a, b, r, lg, stream=None):
        n = a.numel() >> lg
        check(lib().midint_ntt_conv(self._h, n, lg, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_ntt_conv")

    # strided layouts: instance i of a at a + i*sa limbs (0: batch-major)
    def strided(self, op, n, m, a, sa, b, sb, r, sr, carry=None, stream=None):
        """op in 'add', 'add6', 'mul_lo', 'mul_full', 'poly'; a, b, r are
        device buffers (anything with data_ptr()) already offset to
        instance 0 of each operand."""
        fn = getattr(lib(), f"midint_{op}_strided")
        args = [self._h, n, m, _ptr(a), sa, _ptr(b), sb, _ptr(r), sr]
        if op == "add":
            args.append(_ptr(carry))
        check(fn(*args, _stream(stream)), f"midint_{op}_strided")

    def cgbn(self, op, instances, n, bits, stream=None):
        """CGBN instance_t array (a, b, sum per instance); op in CGBN_*."""
        check(lib().midint_cgbn(self._h, op, n, bits, _ptr(instances), _stream(stream)), "midint_cgbn")

    # u32-limb layout: a, b, r hold n * m32 u32 words (int32 tensors or any
    # buffer with data_ptr()); counts are in words
    def add_u32(self, a, b, r, m32, n, carry=None, stream=None):
        check(lib().midint_add_u32(self._h, n, m32, _ptr(a), _ptr(b), _ptr(r), _ptr(carry), _stream(stream)),
              "midint_add_u32")

    def add6_u32(self, a, b, r, m32, n, stream=None):
        check(lib().midint_add6_u32(self._h, n, m32, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_add6_u32")

    def mul_lo_u32(self, a, b, r, m32, n, stream=None):
        check(lib().midint_mul_lo_u32(self._h, n, m32, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_mul_lo_u32")

    def mul_full_u32(self, a, b, r, m32, n, stream=None):
        check(lib().midint_mul_full_u32(self._h, n, m32, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_mul_full_u32")

    def poly_u32(self, a, b, r, m32, n, stream=None):
        check(lib().midint_poly_u32(self._h, n, m32, _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_poly_u32")

    def ntt_transform(self, x, y, lg, prime=0, inverse=False, reference_root=False, stream=None):
        """Batched 2^lg-point transform over PRIME_* (see midint_ntt_transform)."""
        n = x.numel() >> lg
        flags = (NTT_INVERSE if inverse else 0) | (NTT_REFERENCE_ROOT if reference_root else 0)
        check(lib().midint_ntt_transform(self._h, prime, flags, n, lg, _ptr(x), _ptr(y), _stream(stream)),
              "midint_ntt_transform")

    def mul_ntt_prime(self, a, b, r, m, prime, stream=None):
        n = a.numel() // m
        check(lib().midint_mul_ntt_prime(self._h, prime, n, m, _ptr(a), _ptr(b), _ptr(r), _stream(stream)),
              "midint_mul_ntt_prime")

    def ntt_prepare(self, m):
        check(lib().midint_ntt_prepare(self._h, m), "midint_ntt_prepare")

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------- ragged
OP_ADD = 0
OP_MUL = 1


class Op(ctypes.Structure):
    """midint_op: one operation of a ragged batch."""
    _fields_ = [("kind", ctypes.c_uint32), ("n_a", ctypes.c_uint32), ("n_b", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("off_a", ctypes.c_uint64), ("off_b", ctypes.c_uint64),
                ("off_r", ctypes.c_uint64)]


class RaggedTuning(ctypes.Structure):
    _fields_ = [("threads", ctypes.c_uint32), ("blocks_per_cu", ctypes.c_uint32)]


class RaggedPlan:
    """midint_ragged_plan over fixed device buffers (torch tensors)."""

    def __init__(self, ctx, ops, a, b, r, threads=0, blocks_per_cu=0):
        arr = (Op * len(ops))()
        for i, (kind, na, nb, oa, ob, orr) in enumerate(ops):
            arr[i] = Op(kind, na, nb, 0, oa, ob, orr)
        self._keep = (a, b, r)
        tun = RaggedTuning(threads, blocks_per_cu)
        self._h = ctypes.c_void_p()
        check(lib().midint_ragged_plan_create(ctx.handle, arr, len(ops), _ptr(a), _ptr(b), _ptr(r),
                                              ctypes.byref(tun), ctypes.byref(self._h)), "midint_ragged_plan_create")

    def launch(self, stream=None):
        check(lib().midint_ragged_launch(self._h, _stream(stream)), "midint_ragged_launch")

    def stats(self):
        x, y, z = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().midint_ragged_plan_stats(self._h, ctypes.byref(x), ctypes.byref(y), ctypes.byref(z)),
              "midint_ragged_plan_stats")
        return {"items": x.value, "grid": y.value, "lds_bytes": z.value}

    def close(self):
        if self._h:
            lib().midint_ragged_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- serving
class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("max_batch_bytes", ctypes.c_uint64),
                ("streams", ctypes.c_uint32), ("max_batch_ops", ctypes.c_uint32)]


class ServerCounters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "queued", "inflight_batches",
                                                 "batches_launched", "ops_in_batches", "pool_bytes_in_use",
                                                 "pool_bytes_total", "deferred_admissions")]


class Server:
    """midint_server: continuous batching of single add / mul requests."""

    def __init__(self, ctx, hbm_budget=0, max_batch_bytes=0, streams=0, max_batch_ops=0):
        cfg = ServerConfig(hbm_budget, max_batch_bytes, streams, max_batch_ops)
        self._h = ctypes.c_void_p()
        self._ids = (ctypes.c_uint64 * 65536)()
        check(lib().midint_server_create(ctx.handle, ctypes.byref(cfg), ctypes.byref(self._h)),
              "midint_server_create")

    def submit(self, kind, a, b):
        rid = ctypes.c_uint64()
        check(lib().midint_server_submit(self._h, kind, _ptr(a), a.size, _ptr(b), b.size, ctypes.byref(rid)),
              "midint_server_submit")
        return rid.value

    def submit_many(self, reqs):
        """Queues [(kind, a, b)] in one call; returns their ids."""
        return self.submit_prepared(_Batch(reqs), 0, len(reqs))

    def submit_prepared(self, batch, lo, hi):
        """Queues requests lo..hi-1 of a _Batch in one call; returns their ids."""
        import numpy as np

        n = hi - lo
        ids = np.empty(n, dtype=np.uint64)
        if n:
            check(lib().midint_server_submit_many(self._h, n, batch.kinds[lo:].ctypes.data,
                                                  batch.a_ptr[lo:].ctypes.data, batch.n_a[lo:].ctypes.data,
                                                  batch.b_ptr[lo:].ctypes.data, batch.n_b[lo:].ctypes.data,
                                                  ids.ctypes.data), "midint_server_submit_many")
        return ids

    def result_many(self, ids, out, offsets):
        """Copies the results of `ids` (uint64 array) into the uint64 array
        `out` at `offsets`; returns their latencies (seconds)."""
        import numpy as np

        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        lat = np.empty(len(ids), dtype=np.float64)
        if len(ids):
            check(lib().midint_server_result_many(self._h, len(ids), ids.ctypes.data, out.ctypes.data,
                                                  offsets.ctypes.data, out.size, None, lat.ctypes.data),
                  "midint_server_result_many")
        return lat

    def poll(self):
        n = ctypes.c_size_t()
        check(lib().midint_server_poll(self._h, self._ids, len(self._ids), ctypes.byref(n)), "midint_server_poll")
        return list(self._ids[: n.value])

    def result(self, rid, n_limbs):
        import numpy as np

        out = np.empty(n_limbs, dtype=np.uint64)
        got = ctypes.c_size_t()
        lat = ctypes.c_double()
        check(lib().midint_server_result(self._h, rid, _ptr(out), n_limbs, ctypes.byref(got), ctypes.byref(lat)),
              "midint_server_result")
        return out[: got.value], lat.value

    def stats(self):
        c = ServerCounters()
        check(lib().midint_server_stats(self._h, ctypes.byref(c)), "midint_server_stats")
        return {n: getattr(c, n) for n, _ in ServerCounters._fields_}

    def close(self):
        if self._h:
            lib().midint_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Batch:
    """Requests [(kind, a, b)] as the arrays midint_server_submit_many takes
    (the operand arrays must stay alive while it is in use)."""

    def __init__(self, reqs):
        import numpy as np

        self.reqs = reqs
        self.kinds = np.array([k for k, _, _ in reqs], dtype=np.uint32)
        self.a_ptr = np.array([a.ctypes.data for _, a, _ in reqs], dtype=np.uint64)
        self.b_ptr = np.array([b.ctypes.data for _, _, b in reqs], dtype=np.uint64)
        self.n_a = np.array([a.size for _, a, _ in reqs], dtype=np.uint32)
        self.n_b = np.array([b.size for _, _, b in reqs], dtype=np.uint32)


def serve_poisson(server, requests, qps, seed=0, keep_results=True, time_limit_s=120.0, batched=False):
    """Drives `server` with `requests` [(kind, a, b)] arriving as a Poisson
    stream of rate `qps` (qps = inf: every request arrives at time 0, a
    burst that saturates the server). Latency of a request = completion -
    scheduled arrival (host monotonic clock; includes time waiting in the
    queue). Returns dict(p50_ms, p99_ms, mean_ms, ops_per_s, offered_qps
    (None for a burst), results{index: limbs}).
    batched: the requests due at once go in with one submit_many call and
    the completed ones come back with one result_many call per poll (the
    operand pointer arrays and the result buffer are prepared before the
    clock starts), so a burst measures the server rather than one
    interpreted call per request."""
    import math
    import time

    import numpy as np

    rng = np.random.default_rng(seed)
    n = len(requests)
    burst = math.isinf(qps)
    arrivals = np.zeros(n) if burst else np.cumsum(rng.exponential(1.0 / qps, size=n))
    out_len = [(len(a) + len(b)) if k == 1 else max(len(a), len(b)) + 1 for k, a, b in requests]
    id2idx, lat, results = {}, np.zeros(n), {}
    submit_late = np.zeros(n)
    if batched:
        batch = _Batch(requests)
        out_off = np.concatenate([[0], np.cumsum(out_len)[:-1]]).astype(np.uint64)
        out_buf = np.empty(int(sum(out_len)), dtype=np.uint64)
        out_buf.fill(0)  # touch the pages before the clock starts
    i = done = 0
    t0 = time.perf_counter()
    last = t0
    while done < n:
        now = time.perf_counter() - t0
        if now > time_limit_s:
            raise RuntimeError(f"serve_poisson: {done}/{n} done after {time_limit_s} s")
        if batched:
            j = i
            while j < n and arrivals[j] <= now:
                j += 1
            if j > i:
                ids = server.submit_prepared(batch, i, j)
                t_sub = time.perf_counter() - t0
                for q, rid in enumerate(ids.tolist()):
                    id2idx[rid] = i + q
                submit_late[i:j] = t_sub - arrivals[i:j]
                i = j
            got = server.poll()
            if got:
                idx = np.array([id2idx.pop(rid) for rid in got], dtype=np.int64)
                l_c = server.result_many(np.array(got, dtype=np.uint64), out_buf, out_off[idx])
                lat[idx] = submit_late[idx] + l_c
                done += len(got)
                last = time.perf_counter()
            continue
        while i < n and arrivals[i] <= now:
            k, a, b = requests[i]
            rid = server.submit(k, a, b)
            id2idx[rid] = i
            submit_late[i] = (time.perf_counter() - t0) - arrivals[i]
            i += 1
        for rid in server.poll():
            j = id2idx.pop(rid)
            r, l_c = server.result(rid, out_len[j])
            lat[j] = submit_late[j] + l_c
            if keep_results:
                results[j] = r
            done += 1
            last = time.perf_counter()
    if batched and keep_results:
        results = {q: out_buf[int(out_off[q]):int(out_off[q]) + out_len[q]].copy() for q in range(n)}
    span = last - (t0 + arrivals[0])
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "ops_per_s": n / span,
            "offered_qps": None if burst else qps, "requests": n,
            "results": results}


# ---------------------------------------------------------------- host side
def gen_limbs(seed: int, count: int):
    import numpy as np

    out = np.empty(count, dtype=np.uint64)
    lib().midint_gen_limbs(seed, _ptr(out), count)
    return out


def gen_ref_rand(seed: int, n: int, m_u32: int, nz: int):
    import numpy as np

    out = np.empty(n * m_u32, dtype=np.uint32)
    lib().midint_gen_ref_rand(seed, n, m_u32, nz, _ptr(out))
    return out


def tune_set(key: str, value: int) -> None:
    check(lib().midint_tune_set(key.encode(), int(value)), f"midint_tune_set({key})")


TUNE_FILE = os.path.join(HERE, "tune", "mi355x.txt")


def tune_load(path: str = TUNE_FILE) -> bool:
    """Applies a tuning file if it exists; returns whether it was applied."""
    if not os.path.exists(path):
        return False
    check(lib().midint_tune_load(path.encode()), "midint_tune_load")
    return True


def tune_get(key: str) -> int:
    v = ctypes.c_int64()
    check(lib().midint_tune_get(key.encode(), ctypes.byref(v)), f"midint_tune_get({key})")
    return v.value


def probe_stream(a, b, r, stream=None):
    """r = a + b per u64 limb without carries (the add's streaming ceiling)."""
    check(lib().midint_probe_stream(a.numel(), _ptr(a), _ptr(b), _ptr(r), _stream(stream)), "midint_probe_stream")


def probe_valu(device=0):
    """(v_mad_u64_u32 per second, 32-bit adds per second) measured on device."""
    mad = ctypes.c_double(0.0)
    add = ctypes.c_double(0.0)
    check(lib().midint_probe_valu(device, ctypes.byref(mad), ctypes.byref(add)), "midint_probe_valu")
    return mad.value, add.value


def probe_mac(device=0):
    """Partial products per second of the mad + addc accumulate pattern."""
    v = ctypes.c_double(0.0)
    check(lib().midint_probe_mac(device, ctypes.byref(v)), "midint_probe_mac")
    return v.value


def probe_mfma(device=0):
    """(burst, sustained) int8 ops per second of dense 16x16x64 MFMA chains."""
    burst = ctypes.c_double(0.0)
    sus = ctypes.c_double(0.0)
    check(lib().midint_probe_mfma(device, ctypes.byref(burst), ctypes.byref(sus)), "midint_probe_mfma")
    return burst.value, sus.value


def gpu_add_host(a, b, m, runs=0, device=0):
    """Host-pointer drop-in of the reference gpuAdd (cuda/main.cu:39-162)."""
    import numpy as np

    r = np.empty_like(a)
    us = ctypes.c_double(0.0)
    check(lib().midint_gpu_add_host(device, a.size // m, m, _ptr(a), _ptr(b), _ptr(r), runs, ctypes.byref(us)),
          "midint_gpu_add_host")
    return r, us.value


def gpu_mul_host(a, b, m, full=True, runs=0, device=0):
    """Host-pointer drop-in of the reference gpuMultiply (cuda/main.cu:221-356)."""
    import numpy as np

    r = np.empty(a.size * (2 if full else 1), dtype=np.uint64)
    us = ctypes.c_double(0.0)
    check(lib().midint_gpu_mul_host(device, a.size // m, m, int(full), _ptr(a), _ptr(b), _ptr(r), runs,
                                    ctypes.byref(us)), "midint_gpu_mul_host")
    return r, us.value


def gpu_mul_fft_host(a, b, m, runs=0, device=0):
    """Host-pointer drop-in of the reference gpuMulFFT (cuda/main.cu:415-579)."""
    import numpy as np

    r = np.empty(a.size * 2, dtype=np.uint64)
    us = ctypes.c_double(0.0)
    check(lib().midint_gpu_mul_fft_host(device, a.size // m, m, _ptr(a), _ptr(b), _ptr(r), runs,
                                        ctypes.byref(us)), "midint_gpu_mul_fft_host")
    return r, us.value
