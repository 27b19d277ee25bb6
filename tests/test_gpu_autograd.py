"""torch.autograd wrapper of the HIP path (fwihip.autograd), in its own
process: see tests/autograd_case.py."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def test_autograd_gradient_matches_oracle_and_fd():
    # the parent does not touch the GPU: the child process is the first user
    r = subprocess.run([sys.executable, "-u", os.path.join(HERE, "autograd_case.py")], capture_output=True,
                       text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-3000:])
    if "no GPU" in r.stdout:
        pytest.skip("no GPU")
    assert r.returncode == 0
    assert "autograd ok" in r.stdout
