"""The C++ API's host half (MPHF::setData encoders, serialisation, queries;
phobic-gpu_amd/include) against the oracle -- CPU only.  Pilots and offsets
come from the oracle; the compiled encode_driver runs FastQueryMphf on them."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phobic-gpu_amd", "examples", "bin", "encode_driver")


def run_driver(tmp_path, pil, off, n, lam, P, seed, qkeys=None, small=False):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} not built (run __graft_entry__.build())")
    np_ = off.size - 1
    B = po.sizes(max(n, 1), lam, P)[1]
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, np_, B, P], dtype="<u8").tobytes())
        f.write(np.array([lam], dtype="<f8").tobytes())
        f.write(np.array([seed], dtype="<u8").tobytes())
        f.write(np.ascontiguousarray(pil, dtype="<u4").tobytes())
        f.write(np.ascontiguousarray(off, dtype="<u4").tobytes())
    out = tmp_path / "out.bin"
    args = [DRIVER] + (["--small"] if small else []) + [str(inp), str(out)]
    if qkeys is not None:
        qk, qo = tmp_path / "q.bin", tmp_path / "qo.bin"
        with open(qk, "wb") as f:
            f.write(np.array([qkeys.size], dtype="<u8").tobytes())
            f.write(np.ascontiguousarray(qkeys, dtype="<u8").tobytes())
        args += [str(qk), str(qo)]
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    data = open(out, "rb").read()
    pos = np.fromfile(qo, dtype="<u4") if qkeys is not None else None
    return data, pos, r.stdout


@pytest.mark.parametrize("small", [False, True], ids=["FastQueryMphf", "SmallSpaceMphf"])
@pytest.mark.parametrize("n,lam,P,seed,kseed", [
    (1_000_000, 7.0, 2048, 0, 1),
    (300_000, 8.0, 2500, 0, 2),
    (200_000, 7.5, 2048, 42, 3),
    (50_000, 6.0, 2048, 0, 4),
    (5_000, 3.0, 64, 0, 5),
    (1, 7.0, 2048, 0, 6),
])
def test_encoded_bytes_and_queries_match_oracle(tmp_path, n, lam, P, seed, kseed, small):
    """The serialised bytes equal the oracle's encoding, the MPHF loaded back
    from them serialises identically and answers the oracle's positions, and
    the bits-per-key figure matches."""
    keys = po.splitmix64(kseed, n)
    hk = po.hash_u64(keys, seed)
    st, pil, off = po.build_keys(hk, lam, P)
    assert st == 0
    data, pos, out = run_driver(tmp_path, pil, off, n, lam, P, seed, keys, small=small)
    ref, ref_bpk = po.encode_mphf(pil, off, lam, P, n, seed=seed, encoding="rice" if small else "compact",
                                  with_bits_per_key=True)
    assert data == ref
    assert np.array_equal(pos, po.query_keys(hk, pil, off, lam, P))
    assert np.unique(pos).size == n and int(pos.max()) < n
    bpk = float(out.split("bits_per_key=")[1].split()[0])
    assert abs(bpk - ref_bpk) <= 1e-3 * ref_bpk


@pytest.mark.parametrize("small", [False, True])
def test_empty_build_serialises(tmp_path, small):
    off = np.zeros(1, np.uint32)
    data, _, _ = run_driver(tmp_path, np.zeros(0, np.uint32), off, 0, 7.0, 2048, 0, small=small)
    assert data == po.encode_mphf(np.zeros(0, np.uint32), off, 7.0, 2048, 0,
                                  encoding="rice" if small else "compact")
