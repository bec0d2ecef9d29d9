"""Serving loop: results of a Poisson stream under continuous batching equal
the oracle's, requests give the same result alone and in the mix, and the
HBM budget defers (never drops) requests."""
import numpy as np
import pytest

from workloads import make_requests, request_reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(mi):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = mi.Context(0)
    yield c
    c.close()


def test_stream_matches_oracle(mi, ctx, orc):
    reqs = make_requests(mi.gen_limbs, 400, seed=5)
    srv = mi.Server(ctx, hbm_budget=256 << 20, streams=4)
    out = mi.serve_poisson(srv, reqs, qps=4000, seed=1)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["queued"] == 0
    assert st["batches_launched"] > 1
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out["results"][j], request_reference(orc, k, a, b)), (j, k, a.size, b.size)


def test_saturated_burst_batches_and_matches(mi, ctx, orc):
    """Every request arrives at once (offered rate above what the server
    sustains): micro-batches hold many requests, and each result equals the
    oracle's and the result of the same request served alone."""
    reqs = make_requests(mi.gen_limbs, 600, seed=21)
    srv = mi.Server(ctx, hbm_budget=512 << 20, streams=4)
    out = mi.serve_poisson(srv, reqs, qps=float("inf"), seed=7)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    assert st["ops_in_batches"] / st["batches_launched"] > 4
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out["results"][j], request_reference(orc, k, a, b)), (j, k, a.size, b.size)
    for j in (0, 299, 599):
        solo = mi.Server(ctx, streams=1)
        alone = mi.serve_poisson(solo, [reqs[j]], qps=1000, seed=3)["results"][0]
        solo.close()
        assert np.array_equal(alone, out["results"][j])


@pytest.mark.parametrize("qps", [float("inf"), 30000.0], ids=["burst", "stream"])
def test_batched_submit_and_results(mi, ctx, orc, qps):
    """The batched entry points (midint_server_submit_many / _result_many):
    a burst and a Poisson stream driven through them give every request the
    oracle's result and the result of the same request served alone through
    the single-request calls."""
    reqs = make_requests(mi.gen_limbs, 500, seed=33)
    srv = mi.Server(ctx, hbm_budget=512 << 20, streams=4)
    out = mi.serve_poisson(srv, reqs, qps=qps, seed=11, batched=True)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["queued"] == 0
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out["results"][j], request_reference(orc, k, a, b)), (j, k, a.size, b.size)
    for j in (0, 250, 499):
        solo = mi.Server(ctx, streams=1)
        alone = mi.serve_poisson(solo, [reqs[j]], qps=1000, seed=3)["results"][0]
        solo.close()
        assert np.array_equal(alone, out["results"][j])


def test_batched_entry_points_reject_without_side_effects(mi, ctx):
    """submit_many checks every request before queueing any; result_many
    rejects unknown, unfinished and repeated ids before copying or releasing
    any result."""
    import ctypes

    srv = mi.Server(ctx, streams=2)
    good = (1, np.arange(1, 9, dtype=np.uint64), np.arange(1, 5, dtype=np.uint64))
    bad = (1, np.ones(5000, dtype=np.uint64), np.ones(4, dtype=np.uint64))  # above the 4096-limb limit
    with pytest.raises(mi.MidintError):
        srv.submit_many([good, bad])
    assert srv.stats()["submitted"] == 0
    ids = srv.submit_many([good, good])
    done = []
    while len(done) < 2:
        done += srv.poll()
    out = np.zeros(64, dtype=np.uint64)
    with pytest.raises(mi.MidintError):  # the same id twice
        srv.result_many(np.array([ids[0], ids[0]], dtype=np.uint64), out, [0, 16])
    with pytest.raises(mi.MidintError):  # an unknown id
        srv.result_many(np.array([ids[0], 999999], dtype=np.uint64), out, [0, 16])
    lat = srv.result_many(np.array(ids, dtype=np.uint64), out, [0, 16])  # both still there
    assert len(lat) == 2 and (lat >= 0).all()
    want = int("".join(f"{int(x):016x}" for x in good[1][::-1]), 16) * int(
        "".join(f"{int(x):016x}" for x in good[2][::-1]), 16)
    for off in (0, 16):
        got = sum(int(out[off + q]) << (64 * q) for q in range(12))
        assert got == want
    srv.close()
    del ctypes


def test_alone_equals_mixed(mi, ctx, orc):
    reqs = make_requests(mi.gen_limbs, 120, seed=9)
    srv = mi.Server(ctx, streams=3)
    mixed = mi.serve_poisson(srv, reqs, qps=20000, seed=2)["results"]
    srv.close()
    for j in (0, 1, 17, 63, 119):
        k, a, b = reqs[j]
        solo = mi.Server(ctx, streams=1)
        alone = mi.serve_poisson(solo, [reqs[j]], qps=1000, seed=3)["results"][0]
        solo.close()
        assert np.array_equal(alone, mixed[j])


def test_edge_sizes_every_route(mi, ctx, orc):
    """Both operands at the route boundaries: ragged (<= 1024 limbs), add
    classes 2048 / 4096 (including exactly 2048 and 4096, where the carry is
    the kernel's carry-out), NTT at <= 2048 and the four-way split above."""
    MASK = np.uint64(0xFFFFFFFFFFFFFFFF)
    rng = np.random.default_rng(4)
    shapes = [(1, 1), (1024, 1024), (1025, 3), (2048, 2048), (2049, 17), (4096, 4096), (4096, 1), (3000, 2500),
              (1, 4096), (2047, 2049)]
    reqs = []
    for kind in (0, 1):
        for na, nb in shapes:
            for fill in ("ones", "rand"):
                a = np.full(na, MASK, dtype=np.uint64) if fill == "ones" else rng.integers(0, 2**64, na, dtype=np.uint64)
                b = np.full(nb, MASK, dtype=np.uint64) if fill == "ones" else rng.integers(0, 2**64, nb, dtype=np.uint64)
                reqs.append((kind, a, b))
    srv = mi.Server(ctx, streams=2)
    out = mi.serve_poisson(srv, reqs, qps=50000, seed=4)["results"]
    srv.close()
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out[j], request_reference(orc, k, a, b)), (j, k, a.size, b.size)


def test_budget_defers_not_drops(mi, ctx, orc):
    """A pool that holds about one big request at a time still completes all."""
    reqs = make_requests(mi.gen_limbs, 60, seed=13, lo_bits=16, hi_bits=18)
    srv = mi.Server(ctx, hbm_budget=2 << 20, max_batch_bytes=1 << 20, streams=4)
    out = mi.serve_poisson(srv, reqs, qps=100000, seed=5)
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs)
    assert st["deferred_admissions"] > 0
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out["results"][j], request_reference(orc, k, a, b))


@pytest.mark.parametrize("wave,mfma_min,m16", [(0, 0, 0), (0, 1, 0), (1, 32, 0), (1, 32, 1), (1, 32, 2)])
def test_ragged_route_kernels(mi, ctx, orc, wave, mfma_min, m16):
    """Requests with both operands up to 1024 limbs run inside the
    micro-batch on the ragged plans' kernels: with ragged.wave = 1 the
    additions and the products whose shorter operand has at most 128 limbs
    one op per wave, the other products on the matrix-core kernels whose
    tiles ragged.m16@C picks per size class (0: 32 x 32 x 32, 1: 16 x 16 x
    64, 2: by shape); with ragged.wave = 0 the products whose shorter operand
    has at least ragged.mfma_min limbs on the matrix-core kernel (0: none,
    1: all) and the rest on the persistent kernel. Results equal the
    oracle's every way."""
    reqs = make_requests(mi.gen_limbs, 150, seed=17)
    keys = {"ragged.wave": wave, "ragged.mfma_min": mfma_min}
    for c in (256, 512, 1024):
        keys[f"ragged.m16@{c}"] = m16
    saved = {}
    for k, v in keys.items():
        try:
            saved[k] = mi.tune_get(k)
        except Exception:
            saved[k] = mi.tune_get(k.split("@")[0])
        mi.tune_set(k, v)
    try:
        srv = mi.Server(ctx, streams=2)
        out = mi.serve_poisson(srv, reqs, qps=20000, seed=6)["results"]
        srv.close()
    finally:
        for k, v in saved.items():
            mi.tune_set(k, v)
    for j, (k, a, b) in enumerate(reqs):
        assert np.array_equal(out[j], request_reference(orc, k, a, b)), (j, k, a.size, b.size)
