"""GPU: the C-ABI driven from a plain C program (tests/c_abi_example.c,
built in-tree by the build step): compute_q against Darcy's law on a linear
pressure field, compute_P back to the hydrostatic column."""
import os
import subprocess

import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu


def test_c_host_program():
    exe = os.path.join(ROOT, "transportse_amd", "c_abi_example")
    assert os.path.exists(exe), "build first: python -m transportse_amd.build"
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0
    assert p.stdout.count(": ok") == 2
