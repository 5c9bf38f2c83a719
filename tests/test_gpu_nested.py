"""GPU parity of the excited-state (nested-tensor) path against the nested
oracle and the reference's committed es_*.jld2 (effN / envB columns)."""
import math
import os

import numpy as np
import pytest
import torch

import ctmrg_oracle as O
import nested_oracle as N

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def test_es_grad_matches_g5_and_oracle(pkg):
    e = np.load(os.path.join(GOLDEN, "env_ising_D2_X32.npz"))
    es = np.load(os.path.join(GOLDEN, "es_ising_D2_X32.npz"))
    cfg = dict(O.DEFAULT_CFG)
    ts = pkg.CTMTensors(e["A"], cfg)
    ts.set_env([e[f"C{i}"] for i in range(1, 5)], [e[f"E{i}"] for i in range(1, 5)])
    H = pkg.excited.substract_gs_energy(ts, pkg.models.ising(2.0))
    plain = O.CTMTensors(_t(e["A"]), _t(e["Ad"]), [_t(e[f"C{i}"]) for i in range(1, 5)],
                         [_t(e[f"E{i}"]) for i in range(1, 5)], cfg)
    Ho = N.substract_gs_energy(plain, O.ising(2.0))
    for h, ho in zip(H, Ho):
        np.testing.assert_allclose(h, ho.numpy(), atol=1e-12)
    basis = e["basis"]
    ots = N.NestedCTM(plain.A, plain.Ad, plain.Cs, plain.Es, params=dict(cfg, px=0.5 * math.pi, py=0.0))
    for i in (0, 7):
        gH, gN, (E4, Nb) = pkg.get_es_grad(ts, H, basis[:, i], 0.5, 0.0, cfg)
        effN = basis.conj().T @ gN
        effH = basis.conj().T @ gH / 2
        print("column", i, "E4", E4, "Nb", Nb)
        assert np.abs(effN - es["effN_0.5_0"][:, i]).max() < 1e-9
        assert np.abs(gN - es["envB_0.5_0"][:, i]).max() < 1e-6
        ogH, ogN = N.get_es_grad(ots, Ho, _t(basis[:, i]))
        oeffH = (_t(basis).mH @ ogH / 2).numpy()
        assert np.abs(effH - oeffH).max() < 1e-9 * max(1.0, np.abs(oeffH).max())
        assert np.abs(gN - ogN.numpy()).max() < 1e-9


def test_optim_es_columns_and_bands(pkg):
    """effH/effN for a few columns at (pi, pi) and the compute_es pipeline on them."""
    e = np.load(os.path.join(GOLDEN, "env_ising_D2_X32.npz"))
    es = np.load(os.path.join(GOLDEN, "es_ising_D2_X32.npz"))
    cfg = dict(O.DEFAULT_CFG)
    ts = pkg.CTMTensors(e["A"], cfg)
    ts.set_env([e[f"C{i}"] for i in range(1, 5)], [e[f"E{i}"] for i in range(1, 5)])
    H = pkg.excited.substract_gs_energy(ts, pkg.models.ising(2.0))
    cols = [2, 11]
    effH, effN, envB = pkg.optim_es(ts, H, e["basis"], 1.0, 1.0, cfg, columns=cols)
    for i in cols:
        assert np.abs(effN[:, i] - es["effN_1_1"][:, i]).max() < 1e-9
