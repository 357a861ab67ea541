"""The C++ API's host half (MPHF::setData encoders, serialisation, queries;
phobic-gpu_amd/include) against the oracle -- CPU only.  Pilots and offsets
come from the oracle; the compiled encode_driver runs the MPHF type with the
chosen pilot and offset encoders on them."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "phobic-gpu_amd", "examples", "bin", "encode_driver")


def run_driver(tmp_path, pil, off, n, lam, P, seed, qkeys=None, small=False, pilots=None, offsets=None):
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
    args = [DRIVER] + (["--small"] if small else [])
    args += (["--pilots", pilots] if pilots else []) + (["--offsets", offsets] if offsets else [])
    args += [str(inp), str(out)]
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


@pytest.mark.parametrize("offsets", list(po.OFFSET_ENCODINGS))
@pytest.mark.parametrize("pilots", list(po.ENCODINGS))
@pytest.mark.parametrize("n,lam,P,seed,kseed", [
    (400_000, 7.0, 2048, 0, 11),
    (120_000, 8.0, 2500, 42, 12),
])
def test_every_encoder_layout_matches_oracle(tmp_path, n, lam, P, seed, kseed, pilots, offsets):
    """Every pilot layout (interleaved, one sequence, the dual rice/compact
    split) with every offset encoder (differences to i*P or direct; compact or
    Rice): bytes, queries after loading, and bits per key equal the oracle's."""
    keys = po.splitmix64(kseed, n)
    hk = po.hash_u64(keys, seed)
    st, pil, off = po.build_keys(hk, lam, P)
    assert st == 0
    data, pos, out = run_driver(tmp_path, pil, off, n, lam, P, seed, keys, pilots=pilots, offsets=offsets)
    ref, ref_bpk = po.encode_mphf(pil, off, lam, P, n, seed=seed, encoding=pilots, offset_encoding=offsets,
                                  with_bits_per_key=True)
    assert data == ref
    assert np.array_equal(pos, po.query_keys(hk, pil, off, lam, P))
    bpk = float(out.split("bits_per_key=")[1].split()[0])
    assert abs(bpk - ref_bpk) <= 1e-3 * ref_bpk


def test_dual_split_and_space_order(tmp_path):
    """The dual encoder splits at floor(B * 0.5) (B = 293 -> 146 Rice columns),
    and on PHOBIC pilots its size falls between the all-Rice and the
    all-compact columns' (the reason the reference benchmark defaults to it)."""
    n, lam, P = 1_000_000, 7.0, 2048
    keys = po.splitmix64(1, n)
    hk = po.hash_u64(keys, 0)
    st, pil, off = po.build_keys(hk, lam, P)
    assert st == 0
    np_, B = po.sizes(n, lam, P)
    assert B == 293
    bpk = {}
    for e in ("compact", "rice", "dual_rice_compact"):
        data, _, out = run_driver(tmp_path, pil, off, n, lam, P, 0, pilots=e)
        bpk[e] = float(out.split("bits_per_key=")[1].split()[0])
        if e == "dual_rice_compact":
            header = np.frombuffer(data[:64 + 4 * 2048 + 8], dtype="<u8")
            assert int(header[6]) == 6 and int(header[-1]) == 146
    assert bpk["rice"] < bpk["dual_rice_compact"] < bpk["compact"]
