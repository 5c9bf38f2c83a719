"""CPU checks of the C ABI: the built library loads and exports every symbol
that include/ipeps_hip.h declares; the ctypes table covers the header."""
import os
import re

import pytest

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "ipeps_hip.h")).read()
    return sorted(set(re.findall(r"\b(ipeps_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ipeps_rg_step" in names and "ipeps_gs_energy_grad" in names and "ipeps_tcon" in names


def test_library_exports_header_symbols(pkg):
    lib = pkg.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_ctypes_table_matches_header(pkg):
    from ipeps_amd import _lib
    assert sorted(_lib.SIGNATURES) == _declared()


def test_missing_library_fails_loudly(pkg, tmp_path, monkeypatch):
    from ipeps_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(_lib.IpepsError):
        _lib.lib()


def test_models_match_oracle(pkg):
    import numpy as np
    import ctmrg_oracle as O
    M = pkg.models
    for a, b in [(M.ising(2.0), O.ising(2.0)), (M.heisenberg(1.0), O.heisenberg(1.0)),
                 (M.honeycomb(1, 1, dir="XX"), O.honeycomb(1, 1, dir="XX"))]:
        for x, y in zip(a, b):
            np.testing.assert_array_equal(x, y.numpy())
    np.testing.assert_array_equal(M.init_hb_gs(2, "XX"), O.init_hb_gs(2, "XX").numpy())
