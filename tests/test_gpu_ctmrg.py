"""GPU parity tests: the HIP engine (through the C ABI) against the oracle and
the reference's committed fixtures.  Tolerances are stated per test."""
import os

import numpy as np
import pytest
import torch

import ctmrg_oracle as O

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _npz(name):
    return np.load(os.path.join(GOLDEN, name))


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


# ---------------------------------------------------------------- primitives

@pytest.mark.parametrize("cplx", [False, True])
def test_tcon_matches_einsum(pkg, cplx):
    rng = np.random.default_rng(7)

    def r(*s):
        a = rng.standard_normal(s)
        if cplx:
            a = a + 1j * rng.standard_normal(s)
        return a

    cases = [
        ((5, 7), [-1, 1], (7, 3), [1, -2]),                     # plain matmul
        ((7, 5), [1, -1], (7, 3), [1, -2]),                     # A^T
        ((5, 7), [-2, 1], (3, 7), [-1, 1]),                     # output transposed
        ((3, 4, 5, 2), [-1, 1, -3, 2], (2, 4, 6), [2, 1, -2]),  # needs permutes
        ((2, 3, 4, 5, 2), [1, 2, -1, -3, 3], (2, 3, 4, 5, 2), [1, 2, -2, -4, 3]),  # init_ctm1 C1
        ((65, 33), [-1, 1], (33, 129), [1, -2]),                # ragged tiles
        ((4, 1, 3), [-1, -2, 1], (3, 1), [1, -3]),              # unit dims
        ((6,), [1], (6, 5), [1, -1]),                           # vector
        ((20, 20), [1, -1], (20, 20, 3, 2), [-2, 1, -3, -4]),   # batched over E's last legs (middle leg summed)
        ((20, 20, 3, 2), [-1, 1, -3, -4], (20, 20), [1, -2]),   # batched, the batched operand first
        ((20, 3, 20, 2), [-1, -3, 1, -4], (20, 17), [1, -2]),   # needs a permute: plain plan
    ]
    for sa, la, sb, lb in cases:
        A, B = r(*sa), r(*sb)
        got = pkg.tcon([A, B], [la, lb])
        ref = O.tcon([_t(A), _t(B)], [la, lb]).numpy()
        assert got.shape == ref.shape, (sa, la, sb, lb)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * max(1.0, np.abs(ref).max()))


def test_permutes_bit_exact_fp32(pkg):
    """FP32 leg permutations (the mixed engine's storage) through tcon with a
    one-element partner, bit for bit against numpy: the tiled kernel on 16-byte
    words (extents and strides multiples of four: the chi x chi x D^2 edge
    transposes and the corner layouts of the energy network) and on elements
    (odd extents), and the rows kernel."""
    rng = np.random.default_rng(12)
    cases = [((144, 144, 36), (1, 0, 2)), ((36, 144, 144), (2, 0, 1)), ((64, 12, 8, 4), (2, 3, 0, 1)),
             ((8, 12, 6, 2, 4), (4, 1, 3, 0, 2)), ((68, 4, 132), (2, 1, 0)), ((65, 64, 3), (1, 0, 2)),
             ((36, 36, 2), (1, 0, 2)), ((6, 36, 12), (1, 0, 2)), ((144, 6, 6, 6), (0, 2, 1, 3)),
             ((70, 3, 129), (2, 0, 1)), ((4, 4, 4), (2, 1, 0))]
    for shp, perm in cases:
        X = rng.standard_normal(shp).astype(np.float32)
        la = [-(int(np.where(np.array(perm) == a)[0][0]) + 1) for a in range(len(shp))]
        got = pkg.tcon([X, np.ones(1, dtype=np.float32)], [la, [-(len(shp) + 1)]], fp32=True)
        ref = np.transpose(X, perm)[..., None]
        assert got.dtype == np.float32 and got.shape == ref.shape, (shp, perm)
        assert np.array_equal(got, ref), (shp, perm)


@pytest.mark.parametrize("cplx", [False, True])
def test_permutes_bit_exact(pkg, cplx):
    """Leg permutations (rows and LDS-tiled permute kernels) through tcon with a
    one-element partner: X (x) [1.0] with the output legs reordered must equal
    numpy's transpose bit for bit, including shapes whose merged row length
    exceeds one block (row chunking) and tiles with ragged edges."""
    rng = np.random.default_rng(11)
    shapes = [(144, 6, 6, 6), (6, 6, 2, 144, 3), (70, 3, 129), (2, 5000, 3), (5000, 2, 3), (3, 1, 7, 1, 5),
              (9, 8, 7, 6, 5, 4), (65, 65, 3)]
    for shp in shapes:
        for _ in range(4):
            perm = rng.permutation(len(shp))
            X = rng.standard_normal(shp)
            if cplx:
                X = X + 1j * rng.standard_normal(shp)
            la = [-(int(np.where(perm == a)[0][0]) + 1) for a in range(len(shp))]
            got = pkg.tcon([X, np.ones(1, dtype=X.dtype)], [la, [-(len(shp) + 1)]])
            ref = np.transpose(X, perm)[..., None]
            assert got.shape == ref.shape, (shp, perm)
            assert np.array_equal(got, ref), (shp, perm)
    # the fastest axis kept first with even extent and strides: the rows kernel on 16-byte words (real
    # data), and odd extents that fall back to elements
    for shp, perm in (((144, 6, 6, 6), (0, 2, 1, 3)), ((144, 6, 6, 6), (0, 3, 2, 1)), ((6, 144, 3), (0, 2, 1)),
                      ((2, 5000, 3), (0, 2, 1)), ((7, 6, 4), (0, 2, 1)), ((36, 36, 2, 2), (0, 2, 1, 3))):
        X = rng.standard_normal(shp)
        if cplx:
            X = X + 1j * rng.standard_normal(shp)
        perm = np.array(perm)
        la = [-(int(np.where(perm == a)[0][0]) + 1) for a in range(len(shp))]
        got = pkg.tcon([X, np.ones(1, dtype=X.dtype)], [la, [-(len(shp) + 1)]])
        assert np.array_equal(got, np.transpose(X, perm)[..., None]), (shp, perm)


def test_large_gemm(pkg):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((300, 257))
    B = rng.standard_normal((257, 190))
    got = pkg.tcon([A, B], [[-1, 1], [1, -2]])
    np.testing.assert_allclose(got, A @ B, rtol=1e-12, atol=1e-11)


@pytest.mark.parametrize("cplx", [False, True])
def test_svd_and_adjoint(pkg, cplx):
    rng = np.random.default_rng(2)
    N, n = 24, 9
    M = rng.standard_normal((N, N)) + (1j * rng.standard_normal((N, N)) if cplx else 0)
    U, S, V = pkg.svd1(M)
    np.testing.assert_allclose((U * S) @ V.conj().T, M, atol=1e-12)
    assert np.all(np.diff(S) <= 0)
    dU = rng.standard_normal((N, n)) + (1j * rng.standard_normal((N, n)) if cplx else 0)
    dV = rng.standard_normal((N, n)) + (1j * rng.standard_normal((N, n)) if cplx else 0)
    dS = rng.standard_normal(n)
    got = pkg.svd_back(U, S, V, dU, dS, dV)
    Zp = np.zeros((N, N - n), dtype=dU.dtype)
    ref = O.svd_back(_t(U), _t(S), _t(V), _t(np.hstack([dU, Zp])),
                     _t(np.concatenate([dS, np.zeros(N - n)])), _t(np.hstack([dV, Zp]))).numpy()
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("decay", [1e-1, 1e-3, 1e-6])
def test_svd_thin_methods(pkg, cplx, decay):
    """Tall thin SVD, both methods, against numpy on matrices with a known
    spectrum from 1 down to `decay`: values to 1e-12 relative of S_max (Jacobi)
    or u / (S_i/S_max)^2 (Gram); A = U S V^H and orthonormal U, V."""
    rng = np.random.default_rng(3)
    m, n = 300, 40

    def orthonormal(r, c):
        X = rng.standard_normal((r, c)) + (1j * rng.standard_normal((r, c)) if cplx else 0)
        return np.linalg.qr(X)[0]

    s0 = np.logspace(0, np.log10(decay), n)
    A = (orthonormal(m, n) * s0) @ orthonormal(n, n).conj().T
    for method in (0, 1):
        U, S, V, used = pkg.svd_thin(A, method=method)
        assert used == (method if decay >= 1e-4 else 0)
        tol = 1e-12 if used == 0 else 1e-15 / decay ** 2
        assert np.abs(S - s0).max() < tol, (method, np.abs(S - s0).max())
        assert np.abs((U * S) @ V.conj().T - A).max() < 1e-11
        assert np.abs(U.conj().T @ U - np.eye(n)).max() < 1e-12
        assert np.abs(V.conj().T @ V - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("method", [1, 3])
@pytest.mark.parametrize("n", [1, 2, 3, 33, 64, 127, 128, 129, 130, 160, 161, 224, 288, 320])
def test_svd_thin_gram_sizes(pkg, n, method):
    """Gram route (device Householder tridiagonalisation + Q, then the
    tridiagonal eigenpairs by bisection + inverse iteration (method 1) or by
    rocSOLVER stedc (method 3)) at the sizes it serves (k <= 128 in LDS,
    k <= 320 starting from a global workspace; tile, wave and phase edges):
    singular values to 1e-13 of S_max against numpy, A = U S V^T, orthonormal
    U and V, and the kernels that ran."""
    rng = np.random.default_rng(n)
    m = 2000
    s0 = np.logspace(0, -2, n)
    A = (np.linalg.qr(rng.standard_normal((m, n)))[0] * s0) @ np.linalg.qr(rng.standard_normal((n, n)))[0].T
    ctx = pkg.Context()
    ctx.set_timing(1)
    U, S, V, used = pkg.svd_thin(A, method=method, ctx=ctx)
    st = ctx.kernel_stats(reset=True)
    ctx.close()
    assert used == method
    assert "sytrd_kernel" in st and "rocsolver_syevd" not in st
    if method == 1:
        assert "tri_bisect_kernel" in st and "tri_invit_kernel" in st and "rocsolver_stedc" not in st
    else:
        assert "rocsolver_stedc" in st and "tri_invit_kernel" not in st
    assert np.abs(S - s0).max() < 1e-13
    assert np.abs((U * S) @ V.T - A).max() < 1e-12
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_svd_thin_tridiagonal_solvers_agree_graded(pkg, seed):
    """The projector's shape (k = 288 of a 5184-row B, one dominant value, a
    decade-spanning middle and a flat tail across the cut at 144): the device
    bisection + inverse iteration and rocSOLVER stedc give the same values
    (1e-13 of S_max), the same kept subspace and the same truncated
    V S^-1 U^T (to the rounding of the flat tail's relative gaps)."""
    rng = np.random.default_rng(seed)
    m, n, cut = 5184, 288, 144
    s0 = np.empty(n)
    s0[0] = 1.0
    s0[1:cut] = np.geomspace(0.04, 2.75e-4, cut - 1)
    s0[cut:] = np.geomspace(2.75e-4 * 0.991, 2.75e-4 * 0.59, n - cut)
    A = (np.linalg.qr(rng.standard_normal((m, n)))[0] * s0) @ np.linalg.qr(rng.standard_normal((n, n)))[0].T
    ctx = pkg.Context()
    ctx.set_timing(1)
    U1, S1, V1, u1 = pkg.svd_thin(A, method=1, ctx=ctx)
    st = ctx.kernel_stats(reset=True)
    U3, S3, V3, u3 = pkg.svd_thin(A, method=3, ctx=ctx)
    ctx.close()
    assert (u1, u3) == (1, 3)
    assert "rocsolver_stedc" not in st
    assert np.abs(S1 - S3).max() < 1e-13 and np.abs(S1 - s0).max() < 1e-13
    cos_min = np.linalg.svd(U1[:, :cut].T @ U3[:, :cut], compute_uv=False).min()
    assert 1 - cos_min < 1e-12
    P1 = V1[:, :cut] @ np.diag(1 / S1[:cut]) @ U1[:, :cut].T
    P3 = V3[:, :cut] @ np.diag(1 / S3[:cut]) @ U3[:, :cut].T
    assert np.abs(P1 - P3).max() < 1e-8 * np.abs(P3).max()
    assert np.abs(U1.T @ U1 - np.eye(n)).max() < 1e-12
    assert np.abs(V1.T @ V1 - np.eye(n)).max() < 1e-12


def test_svd_thin_repeated_values_fall_back(pkg):
    """Exactly repeated singular values: the inverse-iteration vectors of a
    repeated Gram eigenvalue need not be independent, so the route checks them
    and hands the tridiagonal form to rocSOLVER stedc when they are not; either
    way the result is an SVD to working precision."""
    rng = np.random.default_rng(9)
    m, n = 600, 48
    s0 = np.repeat(np.logspace(0, -2, 12), 4)
    A = (np.linalg.qr(rng.standard_normal((m, n)))[0] * s0) @ np.linalg.qr(rng.standard_normal((n, n)))[0].T
    U, S, V, used = pkg.svd_thin(A, method=1)
    assert used == 1
    assert np.abs(S - s0).max() < 1e-13
    assert np.abs((U * S) @ V.T - A).max() < 1e-12
    assert np.abs(U.T @ U - np.eye(n)).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12


@pytest.mark.parametrize("spread", [0.3, 3e-3])
def test_svd_thin_clustered_tail(pkg, spread):
    """A projector-like spectrum (10 decades-spaced values, then 30 values within
    `spread` of 1.5e-3): the truncated V S^-1 U^H at a cut inside the cluster and
    the kept subspace, both methods, against LAPACK.  Bounds: u / relative gap."""
    rng = np.random.default_rng(5)
    m, n, cut = 300, 40, 24

    def orth(r, c):
        return np.linalg.qr(rng.standard_normal((r, c)))[0]

    s0 = np.concatenate([np.logspace(0, -2, 10), np.sort(1.5e-3 * (1 + spread * rng.random(30)))[::-1]])
    A = (orth(m, n) * s0) @ orth(n, n).T
    Un, Sn, Vnh = np.linalg.svd(A, full_matrices=False)
    Pn = Vnh[:cut].T @ np.diag(1 / Sn[:cut]) @ Un[:, :cut].T
    bound = 1e-8 if spread > 0.1 else 1e-5
    for method in (0, 1):
        U, S, V, used = pkg.svd_thin(A, method=method)
        assert used == method
        P = V[:, :cut] @ np.diag(1 / S[:cut]) @ U[:, :cut].T
        cos_min = np.linalg.svd(U[:, :cut].T @ Un[:, :cut], compute_uv=False).min()
        print(spread, method, np.abs(P - Pn).max() / np.abs(Pn).max(), 1 - cos_min)
        assert np.abs(S - Sn).max() < 1e-13
        assert 1 - cos_min < 1e-12
        assert np.abs(P - Pn).max() < bound * np.abs(Pn).max()


def _gamma_ref(S, UdU, VdV, dS, n, eta=1e-12):
    N = S.size
    G = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        for j in range(N):
            d = S[j] ** 2 - S[i] ** 2
            f = d / (d * d + eta)
            Jij = f * UdU[i, j] if (UdU is not None and j < n) else 0
            Jji = -f * UdU[j, i] if (UdU is not None and i < n) else 0
            Kij = f * VdV[i, j] if (VdV is not None and j < n) else 0
            Kji = -f * VdV[j, i] if (VdV is not None and i < n) else 0
            v = (Jij + np.conj(Jji)) * S[j] + S[i] * (Kij + np.conj(Kji))
            if i == j and i < n:
                if dS is not None:
                    v += dS[i]
                if UdU is not None:
                    v += 1j * UdU[i, i].imag * S[i] / (S[i] ** 2 + eta)
            G[i, j] = v
    return G


def test_svd_back_gamma_terms_complex(pkg):
    """U = V = I isolates Gamma; each cotangent separately."""
    rng = np.random.default_rng(5)
    N, n = 6, 4
    S = np.sort(rng.uniform(0.5, 2.0, N))[::-1].copy()
    I = np.eye(N, dtype=np.complex128)
    dU = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    dV = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    dS = rng.standard_normal(n)
    errs = {}
    for name, args in [("dU", (dU, None, None)), ("dV", (None, None, dV)), ("dS", (None, dS, None)),
                       ("all", (dU, dS, dV))]:
        got = pkg.svd_back(I, S, I, *args)
        ref = _gamma_ref(S, args[0], args[2], args[1], n)
        errs[name] = float(np.abs(got - ref).max())
        if errs[name] > 1e-12:
            D = got - ref
            print(name, "max|diff|", errs[name], "diag", np.abs(np.diag(D)).max(), "re", np.abs(D.real).max(),
                  "im", np.abs(D.imag).max())
            print(np.round(D, 3))
    assert max(errs.values()) < 1e-12, errs


# ---------------------------------------------------------------- CTMRG forward

@pytest.mark.parametrize("which", ["ising_real", "loopgas_complex"])
def test_init_ctm1_parity(pkg, which):
    if which == "ising_real":
        A = O.seeded_tensor((2, 2, 2, 2, 2), 0).numpy()
    else:
        A = O.init_hb_gs(2, dir="XX").numpy()
    ts = pkg.CTMTensors(A, {})
    Cs, Es = O.init_ctm1(_t(A), _t(A).conj())
    for got, ref in zip(ts.Cs + ts.Es, Cs + Es):
        np.testing.assert_allclose(got, ref.numpy(), rtol=0, atol=1e-14)


def test_rg_step_spectra_and_energy(pkg):
    """Gauge-invariant step-by-step parity: left spectrum and energy, 1e-10."""
    A = O.seeded_tensor((2, 2, 2, 2, 2), 11).numpy()
    H = O.ising(2.0)
    ts = pkg.CTMTensors(A, {})
    ots = O.make_ctm(_t(A), {})
    for step in range(4):
        ts, s = pkg.rg_step(ts, 8)
        ots, os_ = O.rg_step(ots, 8)
        np.testing.assert_allclose(s, os_.numpy(), rtol=1e-10)
        E, _ = pkg.get_gs_energy(ts, [h.numpy() for h in H])
        Eo, _ = O.get_gs_energy(ots, H)
        assert abs(E - complex(Eo)) < 1e-10, (step, E, Eo)


def test_env_outlives_context_handle(pkg):
    """Destroying a context handle before its environments is allowed: the
    environment keeps the device context alive and stays usable."""
    A = O.seeded_tensor((2, 2, 2, 2, 2), 0).numpy()
    ctx = pkg.Context()
    ts = pkg.CTMTensors(A, {}, ctx=ctx)
    ctx.close()
    ts, s = pkg.rg_step(ts, 8)
    E, _ = pkg.get_gs_energy(ts, [h.numpy() for h in O.ising(2.0)])
    assert np.isfinite(E.real) and len(s) > 0
    del ts


def test_g4_fixed_environment(pkg):
    """hb_test_D2_X32/basis.jld2: energy of the converged env with the shifted H is 0."""
    e = _npz("env_loopgas_D2_X32.npz")
    ts = pkg.CTMTensors(e["A"], {})
    ts.set_env([e[f"C{i}"] for i in range(1, 5)], [e[f"E{i}"] for i in range(1, 5)])
    E, _ = pkg.get_gs_energy(ts, [e["H1"], e["H2"]])
    assert abs(E) < 1e-12
    assert abs(pkg.get_gs_norm(ts) - 1.0) < 1e-10
    H = O.honeycomb(1, 1, dir="XX")
    Egs, _ = pkg.get_gs_energy(ts, [h.numpy() for h in H])
    ots = O.CTMTensors(_t(e["A"]), _t(e["Ad"]), [_t(e[f"C{i}"]) for i in range(1, 5)],
                       [_t(e[f"E{i}"]) for i in range(1, 5)])
    Eo, _ = O.get_gs_energy(ots, H)
    assert abs(Egs - complex(Eo)) < 1e-12


def test_config1_ising_energy(pkg):
    """Ising h=2, A (2,2,2,2,2) seed 0, chi=30, rg_tol 1e-6, min 4, max 30."""
    A = O.seeded_tensor((2, 2, 2, 2, 2), 0)
    H = O.ising(2.0)
    cfg = dict(chi=30, min_iter=4, max_iter=30, rg_tol=1e-6)
    E, _, ts = pkg.compute_gs_energy(A.numpy(), [h.numpy() for h in H], cfg)
    Eo, _, _, it = O.compute_gs_energy(A, H, cfg)
    assert ts.last_iters == it
    assert abs(E - complex(Eo)) < 1e-10


# ---------------------------------------------------------------- gradient

def test_g1_g2_loopgas_gradient(pkg):
    """energies[1] / gradnorms[1] of simulation/test_test_D2_X32/gs.jld2."""
    ref = _npz("gs_loopgas_D2_X32.npz")
    A0 = O.init_hb_gs(2, dir="XX").numpy()
    H = [h.numpy() for h in O.honeycomb(1, 1, dir="XX")]
    E, g, it = pkg.gs_energy_grad(A0, H, dict(O.DEFAULT_CFG))
    assert abs(E.real - ref["energies"][0]) < 1e-10
    assert abs(np.linalg.norm(g) - ref["gradnorms"][0]) / ref["gradnorms"][0] < 1e-8


def test_g3_gradient_vector(pkg):
    """Optimised A: xk.value and the Optim gradient vector g(x)."""
    ref = _npz("gs_loopgas_D2_X32.npz")
    H = [h.numpy() for h in O.honeycomb(1, 1, dir="XX")]
    E, g, _ = pkg.gs_energy_grad(ref["A"], H, dict(O.DEFAULT_CFG))
    assert abs(E.real - float(ref["value"])) < 1e-10
    assert np.linalg.norm(g - ref["gx"]) < 1e-10


def test_gradient_matches_oracle_real(pkg):
    """Real Ising D=2 chi=12: gradient vector vs the oracle, 1e-9 relative."""
    A0 = O.seeded_tensor((2, 2, 2, 2, 2), 5, unit_max=True)
    H = O.ising(2.0)
    cfg = dict(chi=12, min_iter=4, max_iter=12, ad_max_iter=3, rg_tol=1e-8)
    E, g, _ = pkg.gs_energy_grad(A0.numpy(), [h.numpy() for h in H], cfg)
    Eo, go, _ = O.gs_energy_grad(A0, H, cfg)
    assert abs(E - complex(Eo)) < 1e-10
    go = go.numpy()
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-9


def test_gradient_central_fd(pkg):
    """Central differences (eps=1e-6) of the engine's own energy on 8 entries."""
    A0 = O.seeded_tensor((2, 2, 2, 2, 2), 9, unit_max=True)
    H = O.ising(2.0)
    Hn = [h.numpy() for h in H]
    cfg = dict(chi=8, min_iter=4, max_iter=10, ad_max_iter=2, rg_tol=1e-9)
    E, g, _ = pkg.gs_energy_grad(A0.numpy(), Hn, cfg)
    # the differentiated function: env converged at x, then 2 steps with A = x + d
    x = A0.numpy() / np.abs(A0.numpy()).max()
    base = pkg.CTMTensors(x, cfg)
    pkg.run_ctm(base, "energy", Hn)
    Cs, Es = base.Cs, base.Es

    def f(xx):
        ts = pkg.CTMTensors(xx, cfg)
        ts.set_env(Cs, Es)
        for _ in range(2):
            pkg.rg_step(ts, cfg["chi"])
        return pkg.get_gs_energy(ts, Hn)[0].real

    rng = np.random.default_rng(0)
    idx = rng.choice(x.size, 8, replace=False)
    eps = 1e-6
    gf = g.reshape(-1, order="F")
    for k in idx:
        d = np.zeros(x.size)
        d[k] = eps
        d = d.reshape(x.shape, order="F")
        fd = (f(x + d) - f(x - d)) / (2 * eps)
        assert abs(fd - gf[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, gf[k])


@pytest.mark.parametrize("checkpoint", [0, 1])
def test_randomized_step_adjoint_central_fd(pkg, checkpoint):
    """The benchmark unit's gradient (step_adjoint: rg_step with the randomized
    projector, then the energy, held-constant environment) against central
    differences (eps=1e-6) of the same function -- the test matrices are fixed
    by the environment's draw counter, so the forward is a smooth function of A.
    The adjoint treats the k computed triplets as the truncated SVD of M (the
    reference's svd_back on them), so it is the derivative of the forward when
    the range finder resolves the kept subspace: D=3, chi=20, N=180, k=80,
    four power iterations.  (With k=chi+16 and one power iteration on this
    state, entries differ from the differences by up to ~1e-3 relative: the
    range finder's own approximation.)"""
    M = pkg.models
    A = M.seeded_tensor((3, 3, 3, 3, 2), 3, unit_max=True)
    H = M.heisenberg(1.0)
    chi = 20
    opts = dict(svd_mode=1, oversample=60, power_iters=4, checkpoint=checkpoint)
    base = pkg.CTMTensors(A, {}, dtype="f64", options=opts)
    for _ in range(3):
        pkg.rg_step(base, chi)
    Cs, Es = base.Cs, base.Es

    def fresh(x):
        ts = pkg.CTMTensors(x, {}, dtype="f64", options=opts)
        ts.set_env(Cs, Es)
        return ts

    E, g = pkg.step_adjoint(fresh(A), chi, H)

    def f(x):
        ts = fresh(x)
        pkg.rg_step(ts, chi)
        return pkg.get_gs_energy(ts, H)[0].real

    assert abs(f(A) - E.real) < 1e-13
    rng = np.random.default_rng(1)
    idx = rng.choice(A.size, 8, replace=False)
    eps = 1e-6
    gf = g.reshape(-1, order="F")
    for k in idx:
        d = np.zeros(A.size)
        d[k] = eps
        d = d.reshape(A.shape, order="F")
        fd = (f(A + d) - f(A - d)) / (2 * eps)
        assert abs(fd - gf[k]) < 1e-6 * max(1.0, abs(fd)), (k, fd, gf[k])


def test_config2_gradient_central_fd(pkg):
    """Config 2 (Heisenberg, A (4,4,4,4,2) from seed 1 scaled to unit max-abs,
    chi=64, N=1024): 40 CTMRG steps, then the fg! gradient through 2
    differentiated steps + the energy, against central differences
    (eps=1e-6) of the same function on 32 entries chosen by seed 1.  The
    randomized projector resolves the kept subspace (k = chi + 448 = N / 2,
    q = 4), so its adjoint is the derivative of the forward; the perturbed runs
    continue the environment's test-matrix stream from the same position.
    (With k = chi + 64, q = 3 entries differ from the differences by up to
    1e-3 at this state: the range finder's own approximation.)"""
    M = pkg.models
    A0 = M.seeded_tensor((4, 4, 4, 4, 2), 1, unit_max=True)
    H = M.heisenberg(1.0)
    cfg = dict(chi=64, min_iter=40, max_iter=40, ad_max_iter=2, rg_tol=0.0)
    opts = dict(svd_mode=1, oversample=448, power_iters=4)
    E, g, iters = pkg.gs_energy_grad(A0, H, cfg, dtype="f64", options=opts)
    x = A0 / np.abs(A0).max()
    base = pkg.CTMTensors(x, cfg, options=opts)
    pkg.run_ctm(base, "energy", H)
    Cs, Es, draws = base.Cs, base.Es, base.draws

    def f(xx):
        ts = pkg.CTMTensors(xx, cfg, options=opts)
        ts.set_env(Cs, Es)
        ts.set_option("draws", draws)
        for _ in range(2):
            pkg.rg_step(ts, cfg["chi"])
        return pkg.get_gs_energy(ts, H)[0].real

    assert abs(f(x) - E.real) < 1e-12, (f(x), E)
    rng = np.random.default_rng(1)
    idx = rng.choice(x.size, 32, replace=False)
    eps = 1e-6
    gf = g.reshape(-1, order="F")
    worst = 0.0
    for k in idx:
        d = np.zeros(x.size)
        d[k] = eps
        d = d.reshape(x.shape, order="F")
        fd = (f(x + d) - f(x - d)) / (2 * eps)
        worst = max(worst, abs(fd - gf[k]) / max(1.0, abs(fd)))
    print("config 2: E", E.real, "iters", iters, "worst |fd - g|", worst)
    assert worst < 1e-6


# ---------------------------------------------------------------- fast modes

def test_randomized_projector_energy(pkg):
    """Randomized range finder (svd_mode=1) vs the full SVD on a D=3 state."""
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    cfg = dict(chi=24, min_iter=10, max_iter=10, rg_tol=0.0)
    E_full, _, _ = pkg.compute_gs_energy(A, H, cfg)
    E_rand, _, _ = pkg.compute_gs_energy(A, H, cfg, options=dict(svd_mode=1, oversample=16, power_iters=2))
    E_rand2, _, _ = pkg.compute_gs_energy(A, H, cfg, options=dict(svd_mode=1, oversample=40, power_iters=4))
    print("full", E_full, "rand k+16 q2", abs(E_full - E_rand), "rand k+40 q4", abs(E_full - E_rand2))
    assert abs(E_full - E_rand) < 1e-6
    assert abs(E_full - E_rand2) < abs(E_full - E_rand)


@pytest.mark.parametrize("dtype", ["f64", "c128"])
def test_fast_solver_routes_match_exact(pkg, dtype):
    """Randomized mode, CholeskyQR2 + Gram-eigen small SVD against Householder QR +
    one-sided Jacobi on the same test matrices.  With k = chi+40 and q = 3 the range
    is resolved and the two routes agree (energy 1e-10, gradient 1e-5);
    the fast routes must run on this D=3 state.  (At the
    default k = chi+16, q = 1 both routes carry the range-finder error, ~1e-7 in E
    here, and differ from each other at that level.)"""
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    if dtype == "c128":
        A = A + 0.3j * O.seeded_tensor((3, 3, 3, 3, 2), 5).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    cfg = dict(chi=24, min_iter=6, max_iter=6, ad_max_iter=2, rg_tol=0.0)
    base = dict(svd_mode=1, oversample=40, power_iters=3)
    runs = []
    for extra in ({}, dict(fast_orth=0, fast_svd=0)):
        ctx = pkg.Context()  # fresh context: both routes draw the same test matrices
        ctx.set_timing(1)
        E, g, _ = pkg.gs_energy_grad(A, H, cfg, dtype=dtype, options=dict(base, **extra), ctx=ctx)
        runs.append((E, g, ctx.kernel_stats(reset=True)))
        ctx.close()
    (Ef, gf, st), (Ee, ge, st_exact) = runs
    print(dtype, "fast routes:", sorted(k for k in st if k.startswith("roc")), "exact:",
          sorted(k for k in st_exact if k.startswith("roc")))
    # the Gram route runs (calls whose tail falls below the guard go to gesvdj);
    # CholeskyQR2 on the blocked one-workgroup Cholesky for FP64, rocSOLVER for C128
    chol = "small_chol_inv_kernel" if dtype == "f64" else "rocsolver_potrf"
    eig = "tri_invit_kernel" if dtype == "f64" else "rocsolver_syevd"
    assert eig in st and chol in st
    assert "rocsolver_stedc" not in st
    assert "rocsolver_geqrf_orgqr" not in st
    assert "rocsolver_gesvdj" in st_exact and "rocsolver_geqrf_orgqr" in st_exact
    print("dE", abs(Ef - Ee), "grad rel", np.abs(gf - ge).max() / np.abs(ge).max())
    assert abs(Ef - Ee) < 1e-10
    # the adjoint couples every computed triplet through F_ij = 1/(s_j^2 - s_i^2),
    # which scales the Gram route's tail-vector error u s_1^2 / (s_i ds)
    assert np.abs(gf - ge).max() < 1e-5 * np.abs(ge).max()


@pytest.mark.parametrize("mixed", [3, 6])
def test_mixed_precision_energy(pkg, mixed):
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    cfg = dict(chi=24, min_iter=10, max_iter=10, rg_tol=0.0)
    E64, _, _ = pkg.compute_gs_energy(A, H, cfg)
    Em, _, _ = pkg.compute_gs_energy(A, H, cfg, dtype="mixed", options=dict(mixed=mixed))
    print("mixed", mixed, "E64", E64, "Emixed", Em, "diff", abs(E64 - Em))
    assert abs(E64 - Em) < (1e-4 if mixed == 3 else 1e-6)


@pytest.mark.parametrize("shape", [(70, 45, 37), (300, 200, 150), (36, 72, 5000), (257, 130, 2051),
                                   (1030, 160, 3001), (100, 192, 40), (400, 500, 77), (200, 600, 2100),
                                   (5000, 72, 36), (4097, 17, 5), (12000, 96, 96)])
@pytest.mark.parametrize("kind", ["f64", "c128", "mixed3", "mixed6", "mixed1"])
def test_gemm_all_ops(pkg, kind, shape):
    """Every transpose / conjugate combination against numpy.  Shapes reach the
    FP32 block tiles (N <= 64, 65..128 and N > 192: 128-row tiles; 129..192:
    64 x 192 tiles), aligned (multiple-of-4) and
    unaligned leading dimensions, ragged edges and split-K (long K, few tiles),
    and the FP32 operands pre-split into bf16 planes (B alone, A alone with
    split-K, both; K not a multiple of 32)."""
    rng = np.random.default_rng(11)
    M, N, K = shape
    cplx = kind == "c128"
    mixed = {"mixed1": 1, "mixed3": 3, "mixed6": 6}.get(kind)
    tol = {None: 1e-12, 1: 2e-2, 3: 1e-4, 6: 1e-5}[mixed]
    for tA in (False, True):
        for tB in (False, True):
            for cA in ((False, True) if cplx else (False,)):
                for cB in ((False, True) if cplx else (False,)):
                    sa = (K, M) if tA else (M, K)
                    sb = (N, K) if tB else (K, N)
                    A = rng.standard_normal(sa) + (1j * rng.standard_normal(sa) if cplx else 0)
                    B = rng.standard_normal(sb) + (1j * rng.standard_normal(sb) if cplx else 0)
                    C0 = rng.standard_normal((M, N)) + (1j * rng.standard_normal((M, N)) if cplx else 0)
                    oA = A.T if tA else A
                    oA = oA.conj() if cA else oA
                    oB = B.T if tB else B
                    oB = oB.conj() if cB else oB
                    ref = 0.7 * oA @ oB + 0.3 * C0
                    got = pkg.gemm(A, B, tA, cA, tB, cB, mixed=mixed, alpha=0.7, beta=0.3, C=C0)
                    err = np.abs(got - ref).max() / np.abs(ref).max()
                    assert err < tol, (kind, tA, cA, tB, cB, err)


@pytest.mark.parametrize("emode", ["fp64", "bf16x6", "bf16x3"])
def test_mixed_energy_contractions_fixed_env(pkg, emode):
    """Energy of one fixed FP64 environment evaluated by a mixed engine (FP32
    storage): contractions in FP64, or split-bf16 with 6 / 3 products; the
    density-matrix traces and ratios are FP64 in every mode.  Bound on
    |E_mixed - E_fp64| per site: 1e-7 for FP64 and bf16x6 contractions (the
    stated mixed-path tolerance), 1e-5 for bf16x3."""
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    ts = pkg.CTMTensors(A, dict(chi=24, min_iter=8, max_iter=8, rg_tol=0.0))
    ts, _ = pkg.run_ctm(ts, "energy", H)
    E64, _ = pkg.get_gs_energy(ts, H)
    opts = {"fp64": dict(energy_fp64=1), "bf16x6": dict(energy_fp64=0, energy_mixed=6),
            "bf16x3": dict(energy_fp64=0, energy_mixed=3)}[emode]
    tm = pkg.CTMTensors(A, dtype="mixed", options=opts)
    tm.set_env(ts.Cs, ts.Es)
    Em, _ = pkg.get_gs_energy(tm, H)
    print(emode, "E64", E64.real, "dE", abs(Em.real - E64.real))
    assert abs(Em.real - E64.real) < (1e-5 if emode == "bf16x3" else 1e-7)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 31, 33, 96, 144, 160, 161, 288, 320])
def test_orth_methods(pkg, n):
    """Range bases of a tall Y (5184 x n, cond(Y) ~ 1e4): Householder (0),
    CholeskyQR2 on rocSOLVER (1) and on the blocked one-workgroup Cholesky
    (2: 16 x 16 tile edges, the n = 160 limit of one kernel and the 2 x 2
    block split up to 320 covered).  Q^T Q = I to 1e-13, Q
    spans Y, and 1 and 2 agree to 1e-10 (both are the Q of the QR with a
    positive-diagonal R; 0 agrees up to column signs)."""
    rng = np.random.default_rng(n)
    m = 5184
    Y = rng.standard_normal((m, n)) @ np.diag(np.logspace(0, -4, n)) @ np.linalg.qr(rng.standard_normal((n, n)))[0]
    Qs = [pkg.orth(Y, method=k) for k in (0, 1, 2)]
    for Q in Qs:
        assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
        assert np.abs(Q @ (Q.T @ Y) - Y).max() < 1e-12 * np.abs(Y).max()
    assert np.abs(Qs[1] - Qs[2]).max() < 1e-10
    sg = np.sign(np.sum(Qs[0] * Qs[1], axis=0))
    assert np.abs(Qs[0] * sg - Qs[1]).max() < 1e-10


def test_orth_blocked_cholesky_kernel_runs(pkg):
    """method 2 goes through small_chol_inv_kernel (not rocSOLVER potrf)."""
    ctx = pkg.Context()
    ctx.set_timing(1)
    pkg.orth(np.random.default_rng(0).standard_normal((2000, 160)), method=2, ctx=ctx)
    st = ctx.kernel_stats(reset=True)
    ctx.close()
    assert "small_chol_inv_kernel" in st and "rocsolver_potrf" not in st


@pytest.mark.parametrize("n", [96, 200])
def test_orth_cholesky_failure_falls_back(pkg, n):
    """A rank-deficient Y (two equal columns): the one-workgroup Cholesky of
    its Gram reports a non-positive pivot (or the second pass is far from
    I), and orth falls back to Householder QR: Q orthonormal, Q Q^T Y = Y.
    n = 200 runs the 2 x 2 block split of the Cholesky."""
    rng = np.random.default_rng(3)
    Y = rng.standard_normal((1500, n))
    Y[:, 5] = Y[:, 2]
    ctx = pkg.Context()
    ctx.set_timing(1)
    Q = pkg.orth(Y, method=2, ctx=ctx)
    st = ctx.kernel_stats(reset=True)
    ctx.close()
    assert "small_chol_inv_kernel" in st and "rocsolver_geqrf_orgqr" in st
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-12
    assert np.abs(Q @ (Q.T @ Y) - Y).max() < 1e-12


def test_small_cholesky_route_matches_rocsolver(pkg):
    """fast_orth=2 (CholeskyQR2 on the one-workgroup FP64 Cholesky + triangular
    inverse) against fast_orth=1 (rocSOLVER potrf + trsm) on the same test
    matrices: energy 1e-10, gradient 1e-6 relative."""
    A = O.seeded_tensor((3, 3, 3, 3, 2), 4).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    cfg = dict(chi=24, min_iter=6, max_iter=6, ad_max_iter=2, rg_tol=0.0)
    runs = []
    for fo in (1, 2):
        ctx = pkg.Context()
        ctx.set_timing(1)
        E, g, _ = pkg.gs_energy_grad(A, H, cfg, dtype="f64", options=dict(svd_mode=1, fast_orth=fo), ctx=ctx)
        runs.append((E, g, ctx.kernel_stats(reset=True)))
        ctx.close()
    (E1, g1, st1), (E2, g2, st2) = runs
    assert "small_chol_inv_kernel" in st2 and "small_chol_inv_kernel" not in st1
    print("dE", abs(E1 - E2), "grad rel", np.abs(g1 - g2).max() / np.abs(g1).max())
    assert abs(E1 - E2) < 1e-10
    assert np.abs(g1 - g2).max() < 1e-6 * np.abs(g1).max()


@pytest.mark.parametrize("shape", [(5000, 72, 36), (4096, 96, 96), (70001, 36, 72), (4100, 1, 1)])
def test_gemm_lowk_fp32_products(pkg, shape):
    """FP32 products with K <= 96, N <= 96 over >= 4096 rows run on the low-K
    kernel with FP32 MFMA products whatever the bf16 split asked for: plain
    bf16 (mixed=1, ~1e-2 error through the split kernel) must come out at FP32
    accuracy, for every transpose and with beta."""
    rng = np.random.default_rng(13)
    M, N, K = shape
    for tA in (False, True):
        for tB in (False, True):
            A = rng.standard_normal((K, M) if tA else (M, K)).astype(np.float32).astype(np.float64)
            B = rng.standard_normal((N, K) if tB else (K, N)).astype(np.float32).astype(np.float64)
            C0 = rng.standard_normal((M, N)).astype(np.float32).astype(np.float64)
            ref = 0.7 * (A.T if tA else A) @ (B.T if tB else B) + 0.3 * C0
            got = pkg.gemm(A, B, tA, False, tB, False, mixed=1, alpha=0.7, beta=0.3, C=C0)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < 2e-6, (shape, tA, tB, err)


@pytest.mark.parametrize("dtype", ["f64", "mixed"])
def test_paired_moves_match_sequential(pkg, dtype):
    """rg_step with left || right and top || bottom on two streams / threads
    (paired_moves=1, the default), and the reverse sweep over each pair on two
    lanes, give bit-identical environments, energy and step gradient to the
    sequential schedule, randomized projector included (draw index = position
    of the move in the step)."""
    A = O.seeded_tensor((4, 4, 4, 4, 2), 3).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    out = []
    for pm in (0, 1):
        ts = pkg.CTMTensors(A, dtype=dtype, options=dict(svd_mode=1, paired_moves=pm))
        for _ in range(3):
            pkg.rg_step(ts, 24)
        E, g = pkg.step_adjoint(ts, 24, H)
        out.append((E, g, ts.Cs, ts.Es))
    (E0, g0, C0, E0s), (E1, g1, C1, E1s) = out
    assert E0 == E1
    assert np.array_equal(g0, g1)
    for a, b in zip(C0 + E0s, C1 + E1s):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype,checkpoint", [("f64", 0), ("f64", 1), ("c128", 0)])
def test_run_gs_vjp_is_the_fg_pullback(pkg, dtype, checkpoint):
    """ipeps_run_gs_vjp (the rrule of run_gs, src/optimize/optim_gs.jl:133-143)
    from a converged environment equals the gradient of fg!
    (ipeps_gs_energy_grad), scales with the cotangent dE, and leaves the
    environment it starts from unchanged."""
    M = pkg.models
    A0 = M.seeded_tensor((2, 2, 2, 2, 2), 6, unit_max=False)
    if dtype == "c128":
        A0 = A0 + 0.3j * M.seeded_tensor((2, 2, 2, 2, 2), 7)
    H = M.ising(2.0)
    cfg = dict(chi=12, min_iter=4, max_iter=12, ad_max_iter=3, rg_tol=1e-9)
    opts = dict(svd_mode=0, checkpoint=checkpoint)
    E, g, it = pkg.gs_energy_grad(A0, H, cfg, dtype=dtype, options=opts)
    x = A0 / np.abs(A0).max()
    ts = pkg.CTMTensors(x, cfg, dtype=dtype, options=opts)
    pkg.run_ctm(ts, "energy", H)
    Cs0 = [c.copy() for c in ts.Cs]
    E1, g1, n = pkg.run_gs_vjp(ts, x, H, cfg)
    assert n == it[1]
    # the same arithmetic up to the rounding of rocSOLVER on two concurrent streams
    assert abs(E1 - E) < 1e-13
    assert np.linalg.norm(g1 - g) < 1e-10 * np.linalg.norm(g)
    for a, b in zip(ts.Cs, Cs0):
        np.testing.assert_array_equal(a, b)
    E2, g2, _ = pkg.run_gs_vjp(ts, x, H, cfg, dE=2.5)
    assert np.linalg.norm(g2 - 2.5 * g1) < 1e-10 * np.linalg.norm(g1)
