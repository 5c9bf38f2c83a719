"""CPU checks of the tridiagonal eigensolver steps (ipeps.jl_amd/csrc/trieig_impl.h).

tests/trieig_host.cpp runs the steps of tri_bisect_kernel and
tri_invit_kernel on the host (g++, the header's functions compiled without
HIP); the Newton-Schulz orthonormalisation and its acceptance test follow
gram_svd in ops.cpp.  Reference: LAPACK through scipy.linalg.eigh_tridiagonal.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest
import scipy.linalg as sl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EPS = np.finfo(np.float64).eps
TRIEIG_MAXIT = 20


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("trieig") / "libtrieig_host.so"
    subprocess.run([gxx, "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "ipeps.jl_amd", "csrc"),
                    os.path.join(ROOT, "tests", "trieig_host.cpp"), "-o", str(out)], check=True)
    lib = ctypes.CDLL(str(out))
    P = ctypes.POINTER(ctypes.c_double)
    IP = ctypes.POINTER(ctypes.c_int)
    lib.trieig_host.argtypes = [ctypes.c_int, P, P, P, P, IP]
    lib.trieig_host.restype = ctypes.c_int

    def run(d, e):
        n = len(d)
        d = np.ascontiguousarray(d, dtype=np.float64)
        e = np.ascontiguousarray(np.append(e, 0.0), dtype=np.float64)
        lam = np.zeros(n)
        Zt = np.zeros(n * n)
        steps = np.zeros(n, dtype=np.int32)
        rc = lib.trieig_host(n, d.ctypes.data_as(P), e.ctypes.data_as(P), lam.ctypes.data_as(P),
                             Zt.ctypes.data_as(P), steps.ctypes.data_as(IP))
        assert rc == 0
        return lam, Zt.reshape(n, n, order="F"), steps  # Zt[j, i] = component i of vector j

    return run


def orthonormalise(Zt):
    """gram_svd's two Newton-Schulz steps; returns (Z with eigenvectors as columns, first Gram's distance from I)."""
    n = Zt.shape[0]
    S = Zt @ Zt.T
    err = np.abs(S - np.eye(n)).max()
    Z1 = (1.5 * np.eye(n) - 0.5 * S) @ Zt
    S2 = Z1 @ Z1.T
    Z2 = (1.5 * np.eye(n) - 0.5 * S2) @ Z1
    return Z2.T, err


def check_pairs(d, e, lam, Z, tol_lam=8.0, tol_res=64.0):
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    tn = max(np.abs(np.linalg.eigvalsh(T)).max(), 1e-300)
    ref = sl.eigh_tridiagonal(d, e, eigvals_only=True)
    assert np.all(np.diff(lam) >= 0), "eigenvalues not ascending"
    assert np.abs(lam - ref).max() <= tol_lam * n * EPS * tn
    assert np.abs(Z.T @ Z - np.eye(n)).max() <= 64 * n * EPS
    assert np.abs(T @ Z - Z * lam).max() <= tol_res * n * EPS * tn
    return T, tn


def graded_tridiagonal(n, seed):
    """Tridiagonal form of a Gram matrix with the graded spectrum of the D=6
    chi=144 range finder's B^H B: s_1 = 1, s_2 = 0.04, s_144 = 2.75e-4 with a
    flat tail to s_288 = 0.59 s_144 (lam = s^2)."""
    rng = np.random.default_rng(seed)
    m = n // 2
    s = np.empty(n)
    s[0] = 1.0
    s[1:m] = np.geomspace(0.04, 2.75e-4, m - 1)
    s[m:] = np.geomspace(2.75e-4 * 0.991, 2.75e-4 * 0.59, n - m)
    s *= np.exp(1e-3 * rng.standard_normal(n))
    W, _ = np.linalg.qr(rng.standard_normal((n, n)))
    G = (W * s ** 2) @ W.T
    H = sl.hessenberg(G)
    return np.diag(H).copy(), np.diag(H, -1).copy()


@pytest.mark.parametrize("n", [1, 2, 3, 17, 64, 161, 288, 320])
def test_random_tridiagonal(host, n):
    rng = np.random.default_rng(n)
    d = rng.standard_normal(n)
    e = rng.standard_normal(n - 1)
    lam, Zt, steps = host(d, e)
    Z, err = orthonormalise(Zt)
    assert err <= 1e-3
    check_pairs(d, e, lam, Z)
    assert steps.max() <= TRIEIG_MAXIT


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_graded_gram_spectrum(host, seed):
    """The benchmark's shape: no fallback, eigenvalues to u |T| absolute, and
    eigenvectors as accurate as LAPACK's for the well-separated top triplets."""
    d, e = graded_tridiagonal(288, seed)
    lam, Zt, steps = host(d, e)
    Z, err = orthonormalise(Zt)
    assert err <= 1e-4, err  # inverse iteration alone: ~u |T| / gap between neighbours
    T, tn = check_pairs(d, e, lam, Z)
    lr, Zr = sl.eigh_tridiagonal(d, e)
    # every eigenvalue to high relative accuracy (LAPACK stemr's), the small
    # ones included: bisection runs to the rounding of the interval, not to u |T|
    assert np.max(np.abs(lam - lr) / np.abs(lr)) <= 1e-11
    # the top ten eigenvectors (relative gaps >= 0.5) agree with LAPACK's up to sign
    for j in range(1, 11):
        assert abs(abs(Z[:, -j] @ Zr[:, -j]) - 1.0) <= 1e-12
    # the kept subspace (top 144) is the same subspace
    P = Z[:, -144:] @ Z[:, -144:].T - Zr[:, -144:] @ Zr[:, -144:].T
    assert np.abs(P).max() <= 1e-6


def test_split_and_degenerate_blocks(host):
    """Zero off-diagonals (split blocks) with exactly repeated eigenvalues: the
    inverse-iteration vectors of a repeated eigenvalue are independent, so the
    orthonormalised set is a valid eigenbasis (or the check rejects it)."""
    rng = np.random.default_rng(7)
    b = 12
    db = rng.standard_normal(b)
    eb = rng.standard_normal(b - 1)
    d = np.concatenate([db, db, db])
    e = np.concatenate([eb, [0.0], eb, [0.0], eb])
    lam, Zt, _ = host(d, e)
    Z, err = orthonormalise(Zt)
    if err <= 1e-3:
        check_pairs(d, e, lam, Z, tol_res=256.0)
    else:
        # the rejected case: gram_svd then hands (d, e, Q) to rocSOLVER stedc
        assert np.isfinite(err)


def test_wilkinson_pairs(host):
    """W21+: eigenvalue pairs equal to ~1e-14 relative; the accepted result
    must still be an orthonormal eigenbasis."""
    n = 21
    d = np.abs(np.arange(n) - 10.0)
    e = np.ones(n - 1)
    lam, Zt, _ = host(d, e)
    Z, err = orthonormalise(Zt)
    if err <= 1e-3:
        check_pairs(d, e, lam, Z, tol_res=256.0)


def test_tiny_offdiagonal(host):
    """Off-diagonals near the underflow threshold (the pivmin safeguard)."""
    n = 40
    rng = np.random.default_rng(3)
    d = np.sort(rng.standard_normal(n))
    e = 1e-200 * rng.standard_normal(n - 1)
    lam, Zt, _ = host(d, e)
    Z, err = orthonormalise(Zt)
    assert err <= 1e-3
    check_pairs(d, e, lam, Z)


def test_vectors_continuous_in_the_input(host):
    """Inputs equal to rounding give the same vectors, signs included: the
    sign of a raw inverse-iteration vector follows the side of the computed
    eigenvalue, which flips at random; a flipped vector flips the gauge of
    the renormalised bond, and a fixed random test matrix then sees a
    different M.  LAPACK's own vectors move by ~u |T| / gap under the same
    perturbation, so the bound is that."""
    d, e = graded_tridiagonal(288, 5)
    rng = np.random.default_rng(6)
    lam0, Zt0, _ = host(d, e)
    flips = 0
    for trial in range(4):
        dp = d * (1 + 4 * EPS * rng.standard_normal(d.size))
        ep = e * (1 + 4 * EPS * rng.standard_normal(e.size))
        lam1, Zt1, _ = host(dp, ep)
        Z0, _ = orthonormalise(Zt0)
        Z1, _ = orthonormalise(Zt1)
        dots = np.sum(Z0 * Z1, axis=0)
        flips += int(np.sum(dots < 0))
        # every vector of a relative gap >= 1e-3 stays put (no sign change)
        lr = sl.eigh_tridiagonal(dp, ep, eigvals_only=True)
        gap = np.minimum(np.abs(np.diff(lr, prepend=-np.inf)), np.abs(np.diff(lr, append=np.inf))) / np.abs(lr).max()
        ok = gap > 1e-9
        assert np.all(dots[ok] > 1 - 1e-6), np.min(dots[ok])
    assert flips == 0


def test_planted_mistake_is_caught(host):
    """A sign error in the off-diagonal handed to the solver changes the
    eigenvectors (not the eigenvalues: T and its sign-flipped twin are similar),
    which the residual check must catch."""
    rng = np.random.default_rng(11)
    n = 30
    d = rng.standard_normal(n)
    e = rng.standard_normal(n - 1)
    ebad = e.copy()
    ebad[n // 2] *= -1.0
    lam, Zt, _ = host(d, ebad)
    Z, _ = orthonormalise(Zt)
    with pytest.raises(AssertionError):
        check_pairs(d, e, lam, Z)
