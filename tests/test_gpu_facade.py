"""The C++ API end to end on the GPU: the compiled example builds an MPHF
over string keys through MPHFbuilder (the C ABI underneath) and checks that
it is a bijection onto [0, n), for both of the reference's typedefs."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "phobic-gpu_amd", "examples", "bin", "example")


@pytest.mark.parametrize("kind", ["fast", "small"])
def test_example_builds_a_bijection(kind):
    if not os.path.exists(EXAMPLE):
        pytest.fail(f"{EXAMPLE} not built (run __graft_entry__.build())")
    args = [EXAMPLE, "300000", "7.0", "2048"] + (["small"] if kind == "small" else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "valid for 300000" in r.stdout
    bpk = float(r.stdout.split("total_bits=")[1].split()[0])
    assert 1.5 < bpk < 4.0
