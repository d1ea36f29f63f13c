"""The reference's own templates (compiled from its unmodified headers into
oracle/_ref/dropin_check) instantiated with the HIP KernelType, and the
reference's lbfgs() driven by wdl_evaluate_both: results against the same
templates on the reference CPU kernels."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "dropin_check")


@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/dropin_check not built")
def test_reference_templates_with_hip_kernel_type():
    print("dropin_check", flush=True)
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    line = [l for l in r.stdout.splitlines() if l.startswith("DROPIN ")]
    assert line, r.stdout[-2000:] + r.stderr[-2000:]
    res = json.loads(line[-1][len("DROPIN "):])
    print(res)
    assert r.returncode == 0 and res["ok"], res
