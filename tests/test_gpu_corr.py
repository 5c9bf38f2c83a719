"""GPU parity of the transfer-matrix spectrum and the spin correlations
(src/graph/expectation.jl:193-361) against the CPU oracle
(oracle/corr_oracle.py), on the reference's converged loop-gas environment
(complex) and on a real environment made by the engine."""
import os

import numpy as np
import pytest
import torch

import corr_oracle as Q
import ctmrg_oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _loopgas(pkg):
    e = np.load(os.path.join(GOLDEN, "env_loopgas_D2_X32.npz"))
    ts = pkg.CTMTensors(e["A"], dtype="c128")
    ts.set_env([e[f"C{i}"] for i in range(1, 5)], [e[f"E{i}"] for i in range(1, 5)])
    return ts


def _oracle_env(ts):
    t = torch.from_numpy
    A = ts.A
    return O.CTMTensors(t(A), t(np.conj(A)), [t(c) for c in ts.Cs], [t(x) for x in ts.Es])


def _real_env(pkg):
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    ts = pkg.CTMTensors(A, dtype="f64")
    for _ in range(5):
        pkg.rg_step(ts, 20)
    return ts


@pytest.mark.parametrize("which", ["loopgas_c128", "seeded_f64"])
@pytest.mark.parametrize("direction", ["h", "v"])
def test_correlation_spin_matches_oracle(pkg, which, direction):
    ts = _loopgas(pkg) if which == "loopgas_c128" else _real_env(pkg)
    d = ts.A.shape[4]
    g = np.random.default_rng(9)
    op1 = g.standard_normal((d, d)) + 1j * g.standard_normal((d, d))
    op2 = g.standard_normal((d, d)) + 1j * g.standard_normal((d, d))
    s, m1, m2 = pkg.correlation_spin(ts, op1, op2, max_iter=6, direction=direction)
    so, m1o, m2o = Q.correlation_spin(_oracle_env(ts), torch.from_numpy(op1), torch.from_numpy(op2), 6, direction)
    for r in range(6):
        ref = complex(so[r])
        assert abs(s[r] - ref) < 1e-11 * abs(ref), (r, s[r], ref)
    assert abs(m1 - complex(m1o)) < 1e-11 * abs(complex(m1o))
    assert abs(m2 - complex(m2o)) < 1e-11 * abs(complex(m2o))


@pytest.mark.parametrize("which", ["loopgas_c128", "seeded_f64"])
def test_transfer_matrix_and_spectrum(pkg, which):
    ts = _loopgas(pkg) if which == "loopgas_c128" else _real_env(pkg)
    ots = _oracle_env(ts)
    for direction in ("h", "v"):
        TM = pkg.expect.transfer_matrix(ts, direction)
        TMo = Q.transfer_matrix(ots, direction).numpy()
        np.testing.assert_allclose(TM, TMo, rtol=0, atol=1e-12 * np.abs(TMo).max())
        X = np.random.default_rng(2).standard_normal((TM.shape[1], 3)).astype(TM.dtype)
        Y = pkg.expect.transfer_operator(ts, direction).matmat(X)
        np.testing.assert_allclose(Y, TMo @ X, rtol=0, atol=1e-11 * np.abs(TMo @ X).max())
        es1, es, _ = pkg.correlation_TM(ts, direction, N=6)
        es1o, _, _ = Q.correlation_TM(ots, direction, N=6)
        np.testing.assert_allclose(es1[:6], es1o[:6], rtol=1e-9)
        es1f, _, _ = pkg.correlation_TM(ts, direction, full=True)
        np.testing.assert_allclose(es1f[:6], es1o[:6], rtol=1e-9)


def test_mixed_env_correlations_close_to_fp64(pkg):
    """A mixed (FP32) environment holding the same tensors as an FP64 one:
    the correlations run in complex FP64 on the FP32-rounded tensors."""
    tf = _real_env(pkg)
    tm = pkg.CTMTensors(tf.A, dtype="mixed")
    tm.set_env(tf.Cs, tf.Es)
    op = np.diag([0.5, -0.5])
    s64, _, _ = pkg.correlation_spin(tf, op, op, max_iter=4)
    s32, _, _ = pkg.correlation_spin(tm, op, op, max_iter=4)
    for a, b in zip(s64, s32):
        assert abs(a - b) < 1e-5 * abs(a)


@pytest.mark.parametrize("direction", ["h", "v"])
def test_transfer_matrix_product_edges_closed_form(pkg, direction):
    """The engine's transfer matrix (tm_dense and the matrix-free apply) on
    product-structured edges against the closed form (v . z) kron(w, u)
    (tests/test_oracle_corr.py), complex data."""
    from test_oracle_corr import product_edge_env, product_edge_tm
    A, Cs, Es, u, v, w, z = product_edge_env(direction)
    ts = pkg.CTMTensors(A, dtype="c128")
    ts.set_env(Cs, Es)
    ref = product_edge_tm(u, v, w, z)
    TM = pkg.expect.transfer_matrix(ts, direction)
    np.testing.assert_allclose(TM, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    X = np.random.default_rng(3).standard_normal((TM.shape[1], 4)) + 0j
    Y = pkg.expect.transfer_operator(ts, direction).matmat(X)
    np.testing.assert_allclose(Y, ref @ X, rtol=0, atol=1e-11 * np.abs(ref @ X).max())
