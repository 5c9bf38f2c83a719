"""GEMM autotuner: tuned plans give the same products (to the arithmetic's
tolerance) as the shape rules, a saved table replays bit-identically, and a
CTMRG run with tuning on matches one without."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [("mixed3", (2000, 160, 3000)), ("mixed6", (400, 500, 77)), ("mixed3", (150, 100, 4096)),
          ("f64", (300, 200, 1500)), ("c128", (257, 130, 2051))]


def _operands(kind, shape, rng):
    M, N, K = shape
    cplx = kind == "c128"
    A = rng.standard_normal((K, M)) + (1j * rng.standard_normal((K, M)) if cplx else 0)
    B = rng.standard_normal((K, N)) + (1j * rng.standard_normal((K, N)) if cplx else 0)
    return A, B


def test_tuned_plans_match_and_replay(pkg, tmp_path):
    mixed = {"mixed3": 3, "mixed6": 6}
    tol = {None: 1e-12, 3: 1e-4, 6: 1e-5}
    rng = np.random.default_rng(5)
    ops = [(kind, _operands(kind, shape, rng)) for kind, shape in SHAPES]
    pkg.tune_clear()
    try:
        pkg.tune(1, 0.0)
        first = []
        for kind, (A, B) in ops:
            m = mixed.get(kind)
            ref = A.T @ B
            got = pkg.gemm(A, B, transA=True, mixed=m)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < tol[m], (kind, err)
            first.append(got)
        st = pkg.tune_stats()
        assert st["tuned"] == len(SHAPES) and st["entries"] == len(SHAPES), st
        path = str(tmp_path / "tune.txt")
        pkg.tune_save(path)
        pkg.tune_clear()
        assert pkg.tune_stats()["entries"] == 0
        pkg.tune_load(path)
        pkg.tune(2)
        for (kind, (A, B)), g1 in zip(ops, first):
            g2 = pkg.gemm(A, B, transA=True, mixed=mixed.get(kind))
            assert np.array_equal(g1, g2), kind
        st = pkg.tune_stats()
        assert st["hits"] == len(SHAPES) and st["tuned"] == 0, st
    finally:
        pkg.tune(0)
        pkg.tune_clear()


@pytest.mark.parametrize("dtype", ["f64", "mixed"])
def test_ctmrg_with_tuning_matches_rules(pkg, dtype):
    M = pkg.models
    A = M.seeded_tensor((3, 3, 3, 3, 2), 7)
    H = M.heisenberg(1.0)
    chi = 32
    opts = dict(svd_mode=1)

    def run():
        ts = pkg.CTMTensors(A, {}, dtype=dtype, options=opts)
        for _ in range(3):
            pkg.rg_step(ts, chi)
        return pkg.get_gs_energy(ts, H)[0]

    pkg.tune_clear()
    E0 = run()
    try:
        pkg.tune(1, 1e6)
        E1 = run()
        assert pkg.tune_stats()["tuned"] > 0
    finally:
        pkg.tune(0)
        pkg.tune_clear()
    assert abs(E1 - E0) < (1e-12 if dtype == "f64" else 1e-5), (E0, E1)
