"""CPU checks that pin the correlation oracle (oracle/corr_oracle.py).

Anchors: the single-site norm (get_gs_norm, pinned by the reference's
converged environment G4) must equal the one-column transfer contraction;
the nearest-neighbour correlator must equal the two-site density matrix
(get_dm_h / get_dm_v, pinned by G1-G4) traced against op1 (x) op2; con_LT and
con_TR must be adjoint to each other through con_LR; the transfer-matrix
spectrum must agree between the dense matrix and its matrix-free apply."""
import os

import numpy as np
import pytest
import torch

import corr_oracle as Q
import ctmrg_oracle as O

from conftest import GOLDEN


def _env():
    e = np.load(os.path.join(GOLDEN, "env_loopgas_D2_X32.npz"))
    t = lambda k: torch.from_numpy(e[k])
    return O.CTMTensors(t("A"), t("Ad"), [t(f"C{i}") for i in range(1, 5)], [t(f"E{i}") for i in range(1, 5)])


def _ops(d, seed):
    g = np.random.default_rng(seed)
    mk = lambda: torch.from_numpy(g.standard_normal((d, d)) + 1j * g.standard_normal((d, d)))
    return mk(), mk()


@pytest.mark.parametrize("direction", ["h", "v"])
def test_one_column_equals_single_site_norm(direction):
    ts = _env()
    I = torch.eye(ts.A.shape[4], dtype=torch.complex128)
    envs = (Q.get_envs_TM_h if direction == "h" else Q.get_envs_TM_v)(ts, I, I)
    Et, Ed, T, _, _, L0, R0 = envs
    n1 = Q.con_LR(Q.con_LT(Et, Ed, L0, T, direction), R0)
    n_ref = O.get_gs_norm(ts)
    assert abs(n1 - n_ref) < 1e-12 * abs(n_ref)


@pytest.mark.parametrize("direction", ["h", "v"])
def test_nearest_neighbour_equals_two_site_rdm(direction):
    """s2s[0] / s2s_identity[0] = tr(rho (op1 (x) op2)^T) / tr(rho): op1 acts on
    the first site (left / top), A on its first index and Ad on its second;
    rho's row index is the Ad (bra) side, the tout order puts site 1 fastest."""
    ts = _env()
    d = ts.A.shape[4]
    op1, op2 = _ops(d, 1)
    I = torch.eye(d, dtype=torch.complex128)
    s, _, _ = Q.correlation_spin(ts, op1, op2, max_iter=1, direction=direction)
    n, _, _ = Q.correlation_spin(ts, I, I, max_iter=1, direction=direction)
    roh, rov = O.get_dms(ts)
    rho = roh if direction == "h" else rov
    O12 = torch.from_numpy(np.kron(op2.numpy(), op1.numpy()))  # tout(op1, op2): site-1 index fastest
    ref = torch.trace(rho @ O12.T) / torch.trace(rho)
    got = s[0] / n[0]
    assert abs(got - ref) < 1e-11 * max(1.0, abs(ref))
    # the transposed convention is distinguishable for these operators
    wrong = torch.trace(rho @ O12) / torch.trace(rho)
    assert abs(wrong - ref) > 1e-3


@pytest.mark.parametrize("direction", ["h", "v"])
def test_con_LT_con_TR_adjoint(direction):
    ts = _env()
    g = torch.Generator().manual_seed(3)
    Et, Ed, T, T1, _, L0, R0 = (Q.get_envs_TM_h if direction == "h" else Q.get_envs_TM_v)(ts, *_ops(4, 2))
    L = torch.randn(L0.shape, generator=g, dtype=torch.float64).to(torch.complex128)
    R = torch.randn(R0.shape, generator=g, dtype=torch.float64).to(torch.complex128)
    for X in (T, T1):
        a = Q.con_LR(Q.con_LT(Et, Ed, L, X, direction), R)
        b = Q.con_LR(L, Q.con_TR(Et, Ed, X, R, direction))
        assert abs(a - b) < 1e-11 * abs(a)


@pytest.mark.parametrize("direction", ["h", "v"])
def test_correlation_recursion(direction):
    """s2s[r] = <L_r, R> with L_{r+1} = con_LT(L_r, T): the r-th value equals the
    explicit product of r column maps applied to L0 (dense matrices)."""
    ts = _env()
    op1, op2 = _ops(4, 5)
    s2s, m1, m2 = Q.correlation_spin(ts, op1, op2, max_iter=4, direction=direction)
    Et, Ed, T, T1, T2, L0, R0 = (Q.get_envs_TM_h if direction == "h" else Q.get_envs_TM_v)(ts, op1, op2)
    n = L0.numel()
    # column map as a dense matrix: M[p, m] = d con_LT(L)[p] / d L[m]
    basis = torch.eye(n, dtype=torch.complex128).reshape(n, *L0.shape)
    def dense(X):
        cols = [Q.con_LT(Et, Ed, basis[i], X, direction).reshape(-1) for i in range(n)]
        return torch.stack(cols, dim=1)
    MT, M1 = dense(T), dense(T1)
    R = Q.con_TR(Et, Ed, T2, R0, direction).reshape(-1)
    v = M1 @ L0.reshape(-1)
    for r in range(4):
        ref = (v * R).sum()
        assert abs(s2s[r] - ref) < 1e-10 * abs(ref)
        v = MT @ v
    assert abs(m1 - (M1 @ L0.reshape(-1) * R0.reshape(-1)).sum()) < 1e-10 * abs(m1)


def test_transfer_matrix_spectrum():
    ts = _env()
    es1_full, es_full, TM = Q.correlation_TM(ts, "h", full=True)
    es1, es, _ = Q.correlation_TM(ts, "h", N=6)
    np.testing.assert_allclose(es1[:6], es1_full[:6], rtol=1e-10)
    # TM as the contraction of E4 and E2 over their two D legs
    E4, E2 = ts.Es[3], ts.Es[1]
    x = torch.randn(TM.shape[1], dtype=torch.float64, generator=torch.Generator().manual_seed(1)).to(TM.dtype)
    X = O.reshape_f(x, E4.shape[1], E2.shape[1])
    Y = torch.einsum("acij,bdij,cd->ab", E4, E2, X)
    np.testing.assert_allclose(O.flat_f(Y).numpy(), (TM @ x).numpy(), rtol=1e-12, atol=1e-12)


def product_edge_env(direction, chi=5, D=3, seed=11):
    """An environment whose two transfer-matrix edges are products,
    X[a,c,i,j] = u[a,c] v[i,j] and Y[b,d,i,j] = w[b,d] z[i,j], with u, v, w, z
    complex and non-symmetric; returns (A, Cs, Es, u, v, w, z) as numpy."""
    g = np.random.default_rng(seed)
    cr = lambda *s: g.standard_normal(s) + 1j * g.standard_normal(s)
    u, w, v, z = cr(chi, chi), cr(chi, chi), cr(D, D), cr(D, D)
    X = np.einsum("ac,ij->acij", u, v)
    Y = np.einsum("bd,ij->bdij", w, z)
    Es = [cr(chi, chi, D, D) for _ in range(4)]
    if direction == "h":
        Es[3], Es[1] = X, Y   # TM = tcon([E4, E2], ...)
    else:
        Es[0], Es[2] = X, Y   # TM = tcon([E1, E3], ...)
    Cs = [cr(chi, chi) for _ in range(4)]
    A = cr(D, D, D, D, 2)
    return A, Cs, Es, u, v, w, z


def product_edge_tm(u, v, w, z):
    """The closed form: TM[(a,b),(c,d)] = (sum_ij v_ij z_ij) u[a,c] w[b,d]
    (column-major pair indices a + chi b), i.e. (v . z) kron(w, u), whose
    eigenvalues are (v . z) lambda_i(u) lambda_j(w)."""
    return (v * z).sum() * np.kron(w, u)


@pytest.mark.parametrize("direction", ["h", "v"])
def test_transfer_matrix_product_edges_closed_form(direction):
    """correlation_TM's transfer matrix (expectation.jl:193-220) on product
    edges against the closed form: a ket/bra pairing swapped between the two
    edges gives sum v_ij z_ji instead (v, z are not symmetric), a wrong
    reshape or leg order a different matrix and spectrum."""
    A, Cs, Es, u, v, w, z = product_edge_env(direction)
    t = torch.from_numpy
    ts = O.CTMTensors(t(A), t(A.conj()), [t(c) for c in Cs], [t(e) for e in Es])
    es1, es, TM = Q.correlation_TM(ts, direction, full=True)
    ref = product_edge_tm(u, v, w, z)
    np.testing.assert_allclose(TM.numpy(), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    lam = (v * z).sum() * np.outer(np.linalg.eigvals(u), np.linalg.eigvals(w)).ravel()
    for e in lam:  # every closed-form eigenvalue is in the computed spectrum
        assert np.min(np.abs(es - e)) < 1e-10 * np.abs(lam).max()
    np.testing.assert_allclose(es1, np.sort(np.abs(lam))[::-1], rtol=1e-10)
    # the check catches a contraction that pairs the ket and bra legs
    # crosswise: that TM is (sum v_ij z_ji) kron(w, u), another matrix here
    X, Y = (Es[3], Es[1]) if direction == "h" else (Es[0], Es[2])
    chi = u.shape[0]
    bad = np.einsum("acij,bdji->abcd", X, Y).reshape(chi * chi, chi * chi, order="F")
    assert np.abs(bad - ref).max() > 1e-3 * np.abs(ref).max()
