"""C ABI checks that need no GPU: the library loads, exports every symbol
include/phobic_hip.h declares, and its host-side configuration logic matches
the oracle (no compute calls)."""
import os
import re

import numpy as np
import pytest

import hip_rt
import phobic_hip as ph
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "phobic_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(phobic_\w+)\s*\(", text, re.M)))


def test_header_declares_all_bound_symbols():
    assert declared_symbols() == sorted(ph.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ph.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_status_strings():
    lib = ph.load()
    for code in range(0, 11):
        assert lib.phobic_status_string(code)


@pytest.mark.parametrize("lam,P,n", [(7.0, 2048, 1_000_000), (8.0, 2500, 100_000_000), (7.5, 2048, 10**9),
                                     (6.0, 2048, 200_000_000), (3.0, 64, 1000), (7.0, 2048, 0)])
def test_sizes_and_fulcrums_match_oracle(lam, P, n):
    cfg = ph.make_config(lam, P)
    assert ph.sizes(cfg, n) == po.sizes(n, lam, P)
    assert np.array_equal(ph.fulcrums(cfg), po.fulcrums(lam, P))


@pytest.mark.parametrize("lam,P", [(0.0, 2048), (-1.0, 2048), (float("nan"), 2048), (7.0, 0), (7.0, 1),
                                   (0.001, 2048), (7.0, 1 << 21)])
def test_invalid_configs_rejected(lam, P):
    with pytest.raises(ph.PhobicError) as e:
        ph.sizes(ph.make_config(lam, P), 100)
    assert e.value.status == 1


def test_context_without_device_reports_no_device():
    if hip_rt.device_count() > 0:
        pytest.skip("a device is present")
    with pytest.raises(ph.PhobicError) as e:
        ph.Context(0)
    assert e.value.status == 4


def test_custom_fulcrum_table_bounds():
    """A custom table is accepted with every entry at most B << 16 (then every
    bucket index stays below B) and refused otherwise."""
    import pyoracle as po

    B, f = po.bucketer_fulcrums("uniform", 7.0, 2048)
    cfg = ph.make_config_table(B, f, 7.0, 2048)
    assert ph.sizes(cfg, 1_000_000) == (489, B)
    bad = f.copy()
    bad[2047] = (B << 16) + 1
    with pytest.raises(ph.PhobicError):
        ph.sizes(ph.make_config_table(B, bad, 7.0, 2048), 1_000_000)
