"""The differentiable extraction wrapped in torch.autograd.Function (the test
harness's wrapper of bmf_forward_diff / bmf_backward_diff; the library takes
plain device pointers), checked against the FP64 oracle.

The wrapper runs in its own process (tests/autograd_harness.py), where torch
is imported before the library so that both share torch's HIP runtime; the
rest of the GPU suite stays free of torch.

Bar: the forward is the CPU extraction's mesh bit for bit; for losses
composed in torch (linear in the positions, and the mean squared distance to
a sphere) the density gradient that reaches torch is within 1e-5
(norm-relative) of the FP64 adjoint applied to the same dL/dp; one small
gradient step that flips no label lowers the loss.
"""
import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

HARNESS = Path(__file__).resolve().parent / "autograd_harness.py"


@pytest.fixture(scope="module")
def result():
    pytest.importorskip("torch")
    p = subprocess.run([sys.executable, "-u", str(HARNESS)], capture_output=True, text=True, timeout=600)
    print(p.stderr[-2000:])
    assert p.returncode == 0, p.stderr[-4000:]
    line = [ln for ln in p.stdout.splitlines() if ln.startswith("{")][-1]
    r = json.loads(line)
    print(r)
    return r


def test_forward_is_the_cpu_extraction(result):
    assert result["cuda"]
    assert result["linear_forward_bit_exact"] and result["sphere_forward_bit_exact"]


@pytest.mark.parametrize("kind", ["linear", "sphere"])
def test_gradient_matches_fp64_adjoint(result, kind):
    assert result[f"{kind}_rel_vs_fp64_adjoint"] <= 1e-5
    assert result[f"{kind}_nonzero"] > 0


def test_gradient_step_lowers_the_loss(result):
    assert result["step_loss_after"] < result["step_loss_before"]
