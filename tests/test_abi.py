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


def test_compute_es_generalised_eigenproblem(pkg):
    """compute_es on the reference's effH/effN: H2 v = e N2 v, v^H N2 v = 1."""
    import numpy as np
    from conftest import GOLDEN
    es = np.load(os.path.join(GOLDEN, "es_ising_D2_X32.npz"))
    H, Nm = es["effH_0.5_0"], es["effN_0.5_0"]
    e, v, P = pkg.compute_es(H, Nm, 1e-3)
    Hh = (H + H.conj().T) / 2
    Nh = (Nm + Nm.conj().T) / 2
    H2 = P.conj().T @ Hh @ P
    N2 = P.conj().T @ Nh @ P
    assert np.all(np.diff(e) >= -1e-12)
    np.testing.assert_allclose(H2 @ v, (N2 @ v) * e, atol=1e-9)
    np.testing.assert_allclose(v.conj().T @ N2 @ v, np.eye(v.shape[1]), atol=1e-9)
    ev = np.linalg.eigvalsh(Nh)
    assert P.shape[1] == int(np.sum(ev / ev.max() > 1e-3))


def test_make_es_path_matches_files(pkg):
    """The 34 momenta of make_es_path(0.1) are the es_*.jld2 files the reference wrote."""
    import numpy as np
    from conftest import GOLDEN
    es = np.load(os.path.join(GOLDEN, "es_ising_D2_X32.npz"))
    kx, ky = pkg.make_es_path(0.1)
    tags = {k[5:] for k in es.files if k.startswith("effH_")}

    def fmt(x):
        x = round(float(x), 10)
        return str(int(x)) if x == int(x) else repr(x)

    got = {f"{fmt(a)}_{fmt(b)}" for a, b in zip(kx, ky)}
    assert got == tags


def test_tune_table_file_handling(pkg, tmp_path):
    """The GEMM tuning table's text format: an empty table saves and loads,
    a malformed line and a missing file raise (no GPU call involved)."""
    pkg.tune_clear()
    p = tmp_path / "t.txt"
    pkg.tune_save(str(p))
    assert p.read_text().startswith("#")
    pkg.tune_load(str(p))
    assert pkg.tune_stats()["entries"] == 0
    p.write_text("0.3.00.10.10.10.1.0.000 96 1\n")
    with pytest.raises(pkg.IpepsError):
        pkg.tune_load(str(p))
    # column tiles by kernel family (key = dtype.mixed...: dtype 0 FP64, 2 FP32; mixed 6 = FP32 MFMA)
    for bad in ("0.3.00.10.10.2048.1.0.000 128 1", "2.3.00.10.10.2048.1.0.000 96 1"):
        p.write_text(bad + "\n")
        with pytest.raises(pkg.IpepsError):
            pkg.tune_load(str(p))
    p.write_text("# header\n2.3.00.10.10.2048.1.0.000 192 4\n2.6.00.10.10.2048.1.0.000 96 2\n"
                 "0.3.00.10.10.2048.1.0.000 64 4\n")
    pkg.tune_load(str(p))
    assert pkg.tune_stats()["entries"] == 3
    pkg.tune_clear()
    with pytest.raises(pkg.IpepsError):
        pkg.tune_load(str(tmp_path / "missing.txt"))
    with pytest.raises(pkg.IpepsError):
        pkg.tune(5)


def test_committed_tuning_table_loads(pkg):
    """bench.py replays the committed GEMM tuning table by default: it must parse."""
    path = os.path.join(ROOT, "ipeps.jl_amd", "tune_tables", "gfx950_ctmrg_D6_chi144.txt")
    pkg.tune_clear()
    try:
        pkg.tune_load(path)
        n = sum(1 for l in open(path) if l.strip() and not l.startswith("#"))
        assert pkg.tune_stats()["entries"] == n > 0
    finally:
        pkg.tune_clear()


def test_tune_load_rejects_split_k_on_batched_products(pkg, tmp_path):
    """A tuning table that gives split-K slices to a batched product (or to a
    K < 1024 one) is refused at load: split-K there would compute only the
    first batch.  Host-only: the table is parsed without touching a device."""
    from ipeps_amd import _lib
    lib = _lib.lib()
    bad = tmp_path / "bad.txt"
    # key = dtype.mixed.tAtB.M.N.K.batch.beta.vec (batch 4, K 2048), 8 slices
    bad.write_text("# table\n2.3.00.256.256.2048.4.0.111 128 8\n")
    assert lib.ipeps_tune_load(str(bad).encode()) != 0
    assert b"split-K" in lib.ipeps_last_error()
    short = tmp_path / "short.txt"
    short.write_text("2.3.00.256.256.512.1.0.111 128 2\n")
    assert lib.ipeps_tune_load(str(short).encode()) != 0
    ok = tmp_path / "ok.txt"
    ok.write_text("2.3.00.256.256.2048.1.0.111 128 8\n2.3.00.256.256.2048.4.0.111 128 1\n")
    assert lib.ipeps_tune_load(str(ok).encode()) == 0
    lib.ipeps_tune_clear()
