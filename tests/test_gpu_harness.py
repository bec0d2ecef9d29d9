"""The C harness (host/harness.c, the reference main.cu's runAdditions /
runNaiveMuls / runFFTMuls flow on the C ABI) end to end on the GPU: every
size of every sweep, and the composite timed after each base operation
(six additions, polynomial, NTT polynomial), validated against GMP, as
validateExact does."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_harness_sweeps_validate(mi):
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = os.path.join(mi.HERE, "bin", "midint-harness")
    out = subprocess.run([exe, str(1 << 20)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    lines = [json.loads(x) for x in out.stdout.splitlines() if x.startswith("{")]
    ops = {x["op"] for x in lines}
    assert ops == {"add", "add6", "mul_lo", "poly", "mul_ntt", "poly_ntt", "mul_ntt_p32"}, ops
    assert len(lines) == 9 * 6 + 8
    if "libgmp.so.10 not found" in out.stdout:
        pytest.skip("libgmp.so.10 is not installed on this machine: the harness ran every sweep but "
                    "could not validate the results against GMP")
    assert all(x["valid"] is True for x in lines), [x for x in lines if x["valid"] is not True]
