"""CPU checks that pin the nested (excited-state) oracle, oracle/nested_oracle.py.

Anchors: the reference's committed simulation/test_test_D2_X32/es_*.jld2
(effN and envB columns: they depend on the converged nested environment and
get_all_norm only), the product rule as a directional derivative, and central
differences of the excited-state energy w.r.t. B.  effH of those files was
produced with a four-site (plaquette, MPO-form) Hamiltonian -- basis.jld2
stores H as a 4-tensor MPO -- so it is not a fixture for the two-site
get_es_energy and is not used here.
"""
import math
import os

import numpy as np
import pytest
import torch

import ctmrg_oracle as O
import nested_oracle as N

from conftest import GOLDEN


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def _ising_env():
    e = np.load(os.path.join(GOLDEN, "env_ising_D2_X32.npz"))
    cfg = dict(O.DEFAULT_CFG)
    plain = O.CTMTensors(_t(e["A"]), _t(e["Ad"]), [_t(e[f"C{i}"]) for i in range(1, 5)],
                         [_t(e[f"E{i}"]) for i in range(1, 5)], cfg)
    return plain, _t(e["basis"]), cfg


@pytest.mark.parametrize("px,py,tag,cols", [(0.5, 0.0, "0.5_0", [0, 7]), (1.0, 1.0, "1_1", [3])])
def test_g5_effN_envB_columns(px, py, tag, cols):
    es = np.load(os.path.join(GOLDEN, "es_ising_D2_X32.npz"))
    plain, basis, cfg = _ising_env()
    H = N.substract_gs_energy(plain, O.ising(2.0))
    ts = N.NestedCTM(plain.A, plain.Ad, plain.Cs, plain.Es, params=dict(cfg, px=px * math.pi, py=py * math.pi))
    for i in cols:
        gH, gN = N.get_es_grad(ts, H, basis[:, i])
        effN = (basis.mH @ gN).numpy()
        assert np.abs(effN - es[f"effN_{tag}"][:, i]).max() < 1e-9
        assert np.abs(gN.numpy() - es[f"envB_{tag}"][:, i]).max() < 1e-6


def test_nested_product_rule_is_a_derivative():
    """Component 2 of a nested contraction chain = d/de of component 1 with A -> A + e B."""
    g = torch.Generator().manual_seed(1)
    A = torch.randn(2, 3, 4, generator=g, dtype=torch.float64)
    B = torch.randn(2, 3, 4, generator=g, dtype=torch.float64)
    X = torch.randn(4, 5, generator=g, dtype=torch.float64)
    Y = torch.randn(3, 5, generator=g, dtype=torch.float64)
    nA = N.Nested([A, B, None, None])
    r = N.ntcon([N.ntcon([nA, X], [[-1, -2, 1], [1, -3]]), Y], [[-1, 1, 2], [1, 2]])
    eps = 1e-6
    f = lambda a: O.tcon([O.tcon([a, X], [[-1, -2, 1], [1, -3]]), Y], [[-1, 1, 2], [1, 2]])
    fd = (f(A + eps * B) - f(A - eps * B)) / (2 * eps)
    assert r[2] is None and r[3] is None
    assert torch.allclose(r[1], fd, atol=1e-8)


def test_nested_energy_gradient_fd():
    """d Re(E4) / dB by autograd vs central differences (eps=1e-6), complex B."""
    A = O.seeded_tensor((2, 2, 2, 2, 2), 7).to(torch.complex128)
    H = O.ising(2.0)
    cfg = dict(chi=6, min_iter=3, max_iter=8, ad_max_iter=2, rg_tol=1e-9)
    with torch.no_grad():
        _, _, plain, _ = O.compute_gs_energy(A, H, cfg)
    Hs = N.substract_gs_energy(plain, H)
    ts = N.NestedCTM(plain.A, plain.Ad, plain.Cs, plain.Es, params=dict(cfg, px=0.3 * math.pi, py=0.7 * math.pi))
    gen = torch.Generator().manual_seed(3)
    Bv = torch.randn(32, generator=gen, dtype=torch.float64) + 1j * torch.randn(32, generator=gen, dtype=torch.float64)
    gH, _ = N.get_es_grad(ts, Hs, Bv)
    # the differentiated function: converged nested env (held fixed), 2 steps, E4
    B = O.reshape_f(Bv, *A.shape)
    with torch.no_grad():
        ts1 = N.run_ctm_nested(ts.copy(B=B, Bd=B.conj()), Hs)

    def f(bv):
        b = O.reshape_f(bv, *A.shape)
        with torch.no_grad():
            t2 = N.run_ctm_nested(ts1.copy(B=b, Bd=b.conj()), Hs, max_iter=2)
            return N.get_es_energy(t2, Hs).item()

    eps = 1e-6
    for k in [0, 5, 17, 30]:
        for part, unit in (("re", 1.0), ("im", 1j)):
            d = torch.zeros(32, dtype=torch.complex128)
            d[k] = eps * unit
            fd = (f(Bv + d) - f(Bv - d)) / (2 * eps)
            an = gH[k].real.item() if part == "re" else gH[k].imag.item()
            assert abs(fd - an) < 1e-6 * max(1.0, abs(fd)), (k, part, fd, an)
