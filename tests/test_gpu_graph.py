"""Captured CTMRG steps (HIP graphs) and their graph-safe form: the
one-workgroup Jacobi small SVD against numpy, graph replays against the same
steps run eagerly in graph-safe form (bit for bit, with eager work and
synchronisations between the replays), and batches in graph mode against
eager batches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 100, 111, 112])
def test_jacobi_small_svd_sizes(pkg, n):
    """svd_thin method 2 (Gram eigenpairs by the Jacobi kernel): graded
    singular values to 1e-13 of S_max, A = U S V^T, orthonormal factors."""
    rng = np.random.default_rng(100 + n)
    m = 1500
    s0 = np.logspace(0, -2, n)
    A = (np.linalg.qr(rng.standard_normal((m, n)))[0] * s0) @ np.linalg.qr(rng.standard_normal((n, n)))[0].T
    U, S, V, used = pkg.svd_thin(A, method=2)
    assert used == 2
    assert np.abs(S - s0).max() < 1e-13
    assert np.abs((U * S) @ V.T - A).max() < 1e-12
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


def test_jacobi_small_svd_degenerate(pkg):
    """Repeated singular values (an exact 3-fold and a 2-fold cluster): the
    values to 1e-13 and the reconstruction to 1e-12 (the vectors of a
    cluster are any orthonormal basis of its subspace)."""
    rng = np.random.default_rng(7)
    s0 = np.array([2.0, 1.0, 1.0, 1.0, 0.5, 0.25, 0.25, 0.1] + list(np.logspace(-1.5, -3, 40)))
    n = len(s0)
    A = (np.linalg.qr(rng.standard_normal((900, n)))[0] * s0) @ np.linalg.qr(rng.standard_normal((n, n)))[0].T
    U, S, V, used = pkg.svd_thin(A, method=2)
    assert used == 2
    assert np.abs(S - s0).max() < 1e-13
    assert np.abs((U * S) @ V.T - A).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


def _saturate(pkg, ts, chi, max_steps=10):
    """Eager steps until a step leaves every shape unchanged."""
    for _ in range(max_steps):
        before = [t.shape for t in ts.Cs + ts.Es]
        pkg.rg_step(ts, chi)
        if [t.shape for t in ts.Cs + ts.Es] == before:
            return
    raise AssertionError("shapes did not settle")


@pytest.mark.parametrize("dtype,D,chi", [("mixed", 3, 24), ("mixed", 4, 40), ("f64", 3, 20)])
def test_graph_replays_match_graph_safe_steps(pkg, dtype, D, chi):
    """An environment advanced by graph replays (run_graph) and its clone
    advanced by eager rg_steps in graph-safe form agree bit for bit after
    every replay; each replay is followed by eager work on both (the energy)
    and a synchronisation, and the second replay onwards runs the same
    instantiated graph.  (FP64 with gram_guard 0: the case is about the
    replay, not about the guard's fallback.)"""
    A = pkg.models.seeded_tensor((D, D, D, D, 2), 11 + D)
    H = pkg.models.heisenberg(1.0)
    opts = dict(svd_mode=1, paired_moves=0)
    if dtype == "f64":
        opts["gram_guard"] = 0.0
    ts = pkg.CTMTensors(A, {}, dtype=dtype, options=opts)
    _saturate(pkg, ts, chi)
    ref = ts.clone()
    ref.set_option("graph_safe", 1)
    total = dict(captured=0, replayed=0, undone=0, eager=0)
    for r in range(4):
        st = pkg.run_graph(ts, chi, 1)
        for k in total:
            total[k] += st[k]
        pkg.rg_step(ref, chi)
        for a, b in zip(ts.Cs + ts.Es, ref.Cs + ref.Es):
            assert a.shape == b.shape
            assert np.array_equal(a, b), (r, float(np.abs(a - b).max()))
        Eg, _ = pkg.get_gs_energy(ts, H)
        Er, _ = pkg.get_gs_energy(ref, H)
        assert Eg == Er
        assert ts.draws == ref.draws
    assert total["captured"] == 1 and total["replayed"] == 4, total


def test_graph_dropped_when_the_environment_changes(pkg):
    """An eager step between replays moves the environment's buffers: the
    next run_graph captures again; setting an option drops the graph."""
    A = pkg.models.seeded_tensor((3, 3, 3, 3, 2), 21)
    ts = pkg.CTMTensors(A, {}, dtype="mixed", options=dict(svd_mode=1, paired_moves=0))
    _saturate(pkg, ts, 24)
    assert pkg.run_graph(ts, 24, 2)["captured"] == 1
    assert pkg.run_graph(ts, 24, 1)["captured"] == 0
    pkg.rg_step(ts, 24)
    assert pkg.run_graph(ts, 24, 1)["captured"] == 1
    ts.set_option("mixed_cutoff", 1e-6)
    st = pkg.run_graph(ts, 24, 1)
    assert st["captured"] == 1 and st["replayed"] == 1


def test_batch_graph_mode_matches_eager_batch(pkg):
    """Six uneven instances in graph mode (each step captured once its shapes
    settle, then replayed) against the same batch run eagerly and against the
    FP64 batch: the graph-safe step differs from the eager one in the small
    SVD's eigensolver and the range basis' second CholeskyQR pass, which the
    mixed engine's FP32 rounding turns into energy differences of the size of
    its own error; so graph mode must stay within 2e-5 of eager mode and as
    close to FP64 as eager mode is (within 2x + 1e-6).  The counters show
    captures and replays, also across a second run call."""
    ins = pkg.c4_instances()
    sel = [ins[i] for i in [0, 4, 7, 16, 32, 52]]
    As, chis, Hs = [s[0] for s in sel], [s[1] for s in sel], [s[2] for s in sel]
    bs = {}
    try:
        bs["f64"] = pkg.BatchedCTM(As, chis, Hs, dtype="f64", workers=3, options=dict(svd_mode=1))
        bs["eager"] = pkg.BatchedCTM(As, chis, Hs, dtype="mixed", workers=3, options=dict(svd_mode=1))
        bs["graph"] = pkg.BatchedCTM(As, chis, Hs, dtype="mixed", workers=3, options=dict(svd_mode=1, graph=1))
        st = None
        for steps in (8, 4):
            E = {k: b.run(steps)[0] for k, b in bs.items()}
            st_now = bs["graph"].stats()
            d_ge = np.abs(E["graph"] - E["eager"]).max()
            d_g64 = np.abs(E["graph"] - E["f64"]).max()
            d_e64 = np.abs(E["eager"] - E["f64"]).max()
            print(st_now, d_ge, d_g64, d_e64)
            assert st_now["captured"] >= 1 and st_now["replayed"] >= 1
            if st is not None:
                assert st_now["replayed"] > st["replayed"]
            st = st_now
            assert d_ge < 2e-5
            assert d_g64 < 2 * d_e64 + 1e-6
    finally:
        for b in bs.values():
            b.close()
