"""CPU checks that pin the plaquette oracle (oracle/plaquette_oracle.py)."""
import itertools

import numpy as np
import pytest
import torch

import ctmrg_oracle as O
import plaquette_oracle as P


def _contract_local(sl):
    s1, s2, s3, s4 = sl
    return torch.einsum("abi,icdj,jefk,kgh->abcdefgh", s1, s2, s3, s4)


@pytest.mark.parametrize("model", ["ising_h4", "hb_h4_ZZ"])
def test_get_local_h_reconstructs_h(model):
    """get_local_h's four factors contract back to h (local_h.jl:1-31)."""
    h = getattr(P, model)() if model == "hb_h4_ZZ" else P.ising_h4(2.5)
    sl = P.get_local_h(h)
    assert [s.ndim for s in sl] == [3, 4, 4, 3]
    rec = _contract_local(sl)          # (a,b) site 1 ... (g,h) site 4
    d = int(round(h.shape[0] ** 0.5))  # 2 (Ising) or 4 (two spins per honeycomb site)
    ref = O.reshape_f(h, d, d, d, d, d, d, d, d)
    assert torch.allclose(rec.to(ref.dtype), ref, atol=1e-12)


def test_product_state_closed_form():
    """D = 1 (a product state |a>): every environment tensor is a scalar and
    the plaquette energy is <a a a a| h |a a a a> / <a|a>^4 for any four-site
    h (random Hermitian here), with site i's ket/bra pair as h's i-th index."""
    g = torch.Generator().manual_seed(2)
    d = 2
    a = torch.complex(torch.randn(d, generator=g, dtype=torch.float64), torch.randn(d, generator=g,
                                                                                     dtype=torch.float64))
    X = torch.complex(torch.randn(16, 16, generator=g, dtype=torch.float64),
                      torch.randn(16, 16, generator=g, dtype=torch.float64))
    H16 = X + X.mH
    # h index i = (ket_i, bra_i) of site i; H16 rows = kets (site 1 fastest), cols = bras
    h8 = O.reshape_f(H16, d, d, d, d, d, d, d, d)          # (k1, k2, k3, k4, b1, b2, b3, b4)
    h = h8.permute(0, 4, 1, 5, 2, 6, 3, 7)                 # (k1, b1, k2, b2, ...)
    h = O.reshape_f(h, d * d, d * d, d * d, d * d)
    A = a.reshape(1, 1, 1, 1, d)
    ts = O.make_ctm(A)
    E, e, n = P.get_gs_energy_4x4(ts, P.get_local_h(h))
    # <a|^4 h |a>^4 with the reference's op_Ad: op acts on Ad's physical leg
    aa = a.conj()
    ref = torch.einsum("abcdefgh,a,b,c,d,e,f,g,h->", h8, a, a, a, a, aa, aa, aa, aa) / (aa @ a) ** 4
    assert abs(complex(E) - complex(ref)) < 1e-12 * abs(complex(ref))


def test_kat_loop_gas_zz():
    """test/test_local_h4.jl:5-27: the loop-gas state (ZZ geometry), default
    config, run_ctm with the spectrum as convergence measure, hb_h4_ZZ:
    e0 / 8 = -0.16349 within 1e-5."""
    A = O.init_hb_gs(2, dir="ZZ")
    ts = O.make_ctm(A, dict(O.DEFAULT_CFG))
    with torch.no_grad():
        ts, _, _, _ = O.run_ctm(ts)
        E, _, _ = P.get_gs_energy_4x4(ts, P.get_local_h(P.hb_h4_ZZ()))
    print(E / 8)
    assert abs(complex(E).real / 8 - (-0.16349)) < 1e-5
