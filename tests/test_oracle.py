"""CPU checks that pin the oracle (oracle/ctmrg_oracle.py).

Anchors are the reference's committed simulation outputs (tests/golden/*.npz,
extracted by tests/golden/make_fixtures.py), its known-answer test
(test/test_model.jl:10-15), closed forms and finite differences.
"""
import math
import os

import numpy as np
import pytest
import torch

import ctmrg_oracle as O

from conftest import GOLDEN


def _npz(name):
    return np.load(os.path.join(GOLDEN, name))


def test_tcon_bruteforce():
    g = torch.Generator().manual_seed(0)
    A = torch.randn(2, 3, 4, generator=g, dtype=torch.float64)
    B = torch.randn(4, 2, 5, generator=g, dtype=torch.float64)
    # C[o1, o3] = sum_{c1, c2} A[c2, o3, c1] B[c1, c2, o1], written as loops
    C = O.tcon([A, B], [[2, -2, 1], [1, 2, -1]])
    An, Bn = A.numpy(), B.numpy()
    ref = np.zeros((5, 3))
    for o1 in range(5):
        for o3 in range(3):
            s = 0.0
            for c1 in range(4):
                for c2 in range(2):
                    s += An[c2, o3, c1] * Bn[c1, c2, o1]
            ref[o1, o3] = s
    assert C.shape == (5, 3)
    np.testing.assert_allclose(C.numpy(), ref, rtol=1e-13, atol=1e-13)


def test_reshape_f_is_column_major():
    x = torch.arange(24, dtype=torch.float64).reshape(2, 3, 4)
    y = O.reshape_f(x, 6, 4)
    xn = x.numpy()
    np.testing.assert_array_equal(y.numpy(), xn.reshape(6, 4, order="F"))


def test_cutoff_rule():
    s = torch.tensor([3.0, 2.0, 1.0, 1.0 + 5e-11, 0.5, 1e-13])
    s = torch.sort(s, descending=True).values
    # n=3 cuts inside the near-degenerate pair -> one extra
    assert O.cutoff_count(s, 3) == 4
    assert O.cutoff_count(s, 2) == 2
    # relative 1e-12 cutoff bounds n
    assert O.cutoff_count(s, 10) == 5


def test_g1_g2_loopgas_energy_and_gradnorm():
    """G1/G2: energies[1], gradnorms[1] of simulation/test_test_D2_X32/gs.jld2."""
    ref = _npz("gs_loopgas_D2_X32.npz")
    A0 = O.init_hb_gs(2, dir="XX")
    H = O.honeycomb(1, 1, dir="XX")
    E, g, _ = O.gs_energy_grad(A0, H, dict(O.DEFAULT_CFG))
    assert abs(E.real.item() - ref["energies"][0]) < 1e-12
    gn = torch.linalg.vector_norm(g).item()
    assert abs(gn - ref["gradnorms"][0]) / ref["gradnorms"][0] < 1e-9


def test_g3_gradient_vector():
    """G3: energy and Optim gradient vector g(x) stored with the optimised A."""
    ref = _npz("gs_loopgas_D2_X32.npz")
    A = torch.from_numpy(ref["A"])
    H = O.honeycomb(1, 1, dir="XX")
    E, g, _ = O.gs_energy_grad(A, H, dict(O.DEFAULT_CFG))
    assert abs(E.real.item() - float(ref["value"])) < 1e-12
    gx = torch.from_numpy(ref["gx"])
    assert torch.linalg.vector_norm(g - gx).item() < 1e-12
    assert abs(torch.linalg.vector_norm(g).item() - ref["gradnorms"][-1]) < 1e-12


def test_g4_fixed_environment_energy():
    """G4: converged env of hb_test_D2_X32/basis.jld2; H was shifted by E_gs/2."""
    e = _npz("env_loopgas_D2_X32.npz")
    ts = O.CTMTensors(torch.from_numpy(e["A"]), torch.from_numpy(e["Ad"]),
                      [torch.from_numpy(e[f"C{i}"]) for i in range(1, 5)],
                      [torch.from_numpy(e[f"E{i}"]) for i in range(1, 5)])
    Hs = [torch.from_numpy(e["H1"]), torch.from_numpy(e["H2"])]
    H = O.honeycomb(1, 1, dir="XX")
    E_shift, _ = O.get_gs_energy(ts, Hs)
    E_gs, _ = O.get_gs_energy(ts, H)
    offset = float((H[0] - Hs[0]).diagonal()[0])
    assert abs(offset - E_gs.real.item() / 2) < 1e-12
    assert abs(E_shift) < 1e-12
    assert abs(O.get_gs_norm(ts) - 1.0) < 1e-10   # normalize_gs (optim_es.jl:212-222)


def test_kat_loopgas_energy():
    """test/test_model.jl:10-15: real(e0)/8 ~ -0.16349 (atol 1e-5)."""
    E, _, _, _ = O.compute_gs_energy(O.init_hb_gs(2, dir="XX"), O.honeycomb(1, 1, dir="XX"),
                                     dict(O.DEFAULT_CFG))
    assert abs(E.real.item() / 8 + 0.16349) < 1e-5


def test_product_state_closed_form():
    """D=1: CTMRG is exact, E = 2 <pp|H|pp>/<p|p>^2."""
    phi = torch.tensor([0.3, -1.1], dtype=torch.float64)
    A = phi.reshape(1, 1, 1, 1, 2).clone()
    H = O.ising(2.0)
    cfg = dict(chi=4, min_iter=2, max_iter=3, rg_tol=1e-12)
    E, _, _, _ = O.compute_gs_energy(A, H, cfg)
    pp = torch.kron(phi, phi)
    ref = 2 * (pp @ H[0] @ pp) / (phi @ phi) ** 2
    assert abs(E.real.item() - ref.item()) < 1e-13


def test_energy_scale_and_gauge_invariance():
    A = O.seeded_tensor((2, 2, 2, 2, 2), 11)
    H = O.ising(2.0)
    cfg = dict(chi=8, min_iter=6, max_iter=6, rg_tol=0.0)
    E0, _, _, _ = O.compute_gs_energy(A, H, cfg)
    E1, _, _, _ = O.compute_gs_energy(3.7 * A, H, cfg)
    # orthogonal virtual gauge on the up/down and left/right bonds (SVD-based
    # projectors are covariant under unitary gauge transformations)
    c, s_ = math.cos(0.7), math.sin(0.7)
    g = torch.tensor([[c, -s_], [s_, c]], dtype=torch.float64)
    gi = g.T
    Ag = torch.einsum("uldrp,Uu,dD->UlDrp", A, g, gi)
    Ag = torch.einsum("uldrp,Ll,rR->uLdRp", Ag, g, gi)
    E2, _, _, _ = O.compute_gs_energy(Ag, H, cfg)
    assert abs(E0 - E1) < 1e-12
    assert abs(E0 - E2) < 1e-8


def test_gradient_central_fd():
    """Autograd through rg_step + svd_back vs central differences (eps=1e-6)."""
    A0 = O.seeded_tensor((2, 2, 2, 2, 2), 5, unit_max=True)
    H = O.ising(2.0)
    cfg = dict(chi=6, min_iter=4, max_iter=8, ad_max_iter=3, rg_tol=1e-8)
    E, g, ts = O.gs_energy_grad(A0, H, cfg)
    x = O.renormalize(A0)
    ts.params = dict(ts.params, max_iter=3)

    def f(xx):
        with torch.no_grad():
            return O.run_gs(ts, H, xx)[0].real.item()

    rng = np.random.default_rng(0)
    idx = rng.choice(x.numel(), 6, replace=False)
    eps = 1e-6
    for k in idx:
        d = torch.zeros(x.numel(), dtype=torch.float64)
        d[k] = eps
        d = d.reshape(x.shape)
        fd = (f(x + d) - f(x - d)) / (2 * eps)
        assert abs(fd - g.reshape(-1)[k].item()) < 1e-7 * max(1.0, abs(fd))


def test_svd_back_matches_fd():
    g = torch.Generator().manual_seed(3)
    M = torch.randn(6, 6, generator=g, dtype=torch.float64)
    W = torch.randn(6, 3, generator=g, dtype=torch.float64)

    def loss(m):
        u, s, v = O.wrap_svd(m, 3)
        return ((u @ torch.diag(s) @ W.T) ** 2).sum() + (v * W).sum() ** 2

    Mg = M.clone().requires_grad_(True)
    loss(Mg).backward()
    eps = 1e-6
    for (i, j) in [(0, 0), (2, 5), (5, 1)]:
        d = torch.zeros_like(M)
        d[i, j] = eps
        with torch.no_grad():
            fd = (loss(M + d) - loss(M - d)) / (2 * eps)
        assert abs(fd.item() - Mg.grad[i, j].item()) < 1e-6


@pytest.mark.parametrize("ia,iad", [
    ([-2, 1, -4, -5, -7], [-3, 1, -6, -8, 2]),   # proj_left
    ([-2, -5, -4, 1, -7], [-3, -8, -6, 1, 2]),   # proj_right
    ([1, -2, -5, -4, -7], [1, -3, -8, -6, 2]),   # proj_top
    ([-5, -2, 1, -4, -7], [-8, -3, 1, -6, 2]),   # proj_bottom
])
def test_edge_canonical_order_matches_reference_order(ia, iad):
    """The canonical-order edge update (CPU-baseline timing variant) equals the
    reference-order one, complex data, every direction's label lists."""
    g = torch.Generator().manual_seed(3)
    D, chi, chi1, d = 3, 5, 4, 2
    cr = lambda *s: torch.complex(torch.randn(*s, generator=g, dtype=torch.float64),
                                  torch.randn(*s, generator=g, dtype=torch.float64))
    E, A, Ad = cr(chi, chi, D, D), cr(D, D, D, D, d), cr(D, D, D, D, d)
    P, Pd = cr(chi * D * D, chi1), cr(chi1, chi * D * D)
    O.EDGE_ORDER = "reference"
    r = O._edge(E, A, Ad, ia, iad, P, Pd)
    try:
        O.EDGE_ORDER = "canonical"
        cno = O._edge(E, A, Ad, ia, iad, P, Pd)
    finally:
        O.EDGE_ORDER = "reference"
    assert cno.shape == r.shape
    assert torch.allclose(cno, r, rtol=1e-12, atol=1e-12 * r.abs().max())
