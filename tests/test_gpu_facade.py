"""The C++ API end to end on the GPU: the compiled examples build MPHFs
through MPHFbuilder (the C ABI underneath) over string keys, for both of the
reference's typedefs, and over every other key type the reference's hasher
takes, and check that each is a bijection onto [0, n)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "phobic-gpu_amd", "examples", "bin", "example")
KEY_TYPES = os.path.join(ROOT, "phobic-gpu_amd", "examples", "bin", "key_types")


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
    # the reference's timer labels in its order (mphf_builder.h:138-146,
    # 229-290; printed by benchmark.cpp:81-90 as RESULT columns)
    result = next(line for line in r.stdout.splitlines() if line.startswith("RESULT "))
    cols = dict(tok.split("=", 1) for tok in result.split()[1:] if "=" in tok)
    labels = ["initial_hash", "allocation", "input_transfer", "setup_commands", "GPU_bucket_sizes",
              "GPU_bucket_sorting", "GPU_partition_offsets", "GPU_apply_partition_offsets", "GPU_key_redistribution",
              "GPU_search", "GPU_memory_map", "result_transfer", "encoding"]
    order = [tok.split("=", 1)[0] for tok in result.split()[1:] if tok.split("=", 1)[0] in labels]
    assert order == labels
    assert all(float(cols[k]) >= 0 for k in labels)
    assert float(cols["GPU_search"]) > 0 and float(cols["GPU_bucket_sizes"]) > 0


def test_every_key_type_builds_a_bijection():
    """MPHFbuilder::build over uint64_t, std::pair<uint64_t, uint64_t> and Key
    (hashed as the reference's xxhash does, hasher.hpp:28-66), and over
    pre-hashed Key values with nohash; each MPHF is a bijection."""
    if not os.path.exists(KEY_TYPES):
        pytest.fail(f"{KEY_TYPES} not built (run __graft_entry__.build())")
    r = subprocess.run([KEY_TYPES, "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "initialHash matches the known answer" in r.stdout
    for name in ("uint64_t", "pair<uint64_t,uint64_t>", "Key (xxhash)", "Key (nohash)"):
        assert f"{name} valid for 200000" in r.stdout
