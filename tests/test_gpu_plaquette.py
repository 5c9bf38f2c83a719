"""The 2 x 2 plaquette energy path (get_local_h, get_op_Ad4,
energy_norm_4x4, get_gs_energy_4x4; src/model/local_h.jl,
src/graph/evaluation.jl:340-519) on the GPU against the reference's KAT and
the oracle (oracle/plaquette_oracle.py)."""
import numpy as np
import pytest
import torch

import ctmrg_oracle as O
import plaquette_oracle as P

pytestmark = pytest.mark.gpu


def _oracle_env(ts):
    t = torch.from_numpy
    A = ts.A
    return O.CTMTensors(t(A), t(np.conj(A)), [t(c) for c in ts.Cs], [t(x) for x in ts.Es])


def test_kat_loop_gas_plaquette(pkg):
    """test/test_local_h4.jl:5-27: loop-gas state in the ZZ geometry, default
    config, run_ctm on the spectrum, hb_h4_ZZ: e0 / 8 = -0.16349 +- 1e-5; and
    the oracle on the same environment to 1e-12."""
    M = pkg.models
    cfg = dict(chi=32, min_iter=4, max_iter=20, ad_max_iter=4, rg_tol=1e-6)
    ts = pkg.CTMTensors(M.init_hb_gs(2, "ZZ"), cfg, dtype="c128", options=dict(svd_mode=0))
    pkg.run_ctm(ts, None)
    sl = M.get_local_h(M.hb_h4_ZZ())
    E, N = pkg.expect.get_gs_energy_4x4(ts, sl)
    print("e0/8", E / 8)
    assert abs(E.real / 8 - (-0.16349)) < 1e-5
    Eo, _, _ = P.get_gs_energy_4x4(_oracle_env(ts), [torch.from_numpy(s) for s in sl])
    assert abs(E - complex(Eo)) < 1e-12 * abs(complex(Eo))


@pytest.mark.parametrize("dtype", ["f64", "mixed"])
def test_ising_plaquette_vs_oracle(pkg, dtype):
    """ising_h4 on a random real D=3 environment (the mixed engine's FP32
    environment is evaluated in FP64, against the oracle on the same FP32
    values) at 1e-12."""
    M = pkg.models
    A = M.seeded_tensor((3, 3, 3, 3, 2), 8)
    ts = pkg.CTMTensors(A, {}, dtype=dtype, options=dict(svd_mode=1, oversample=20, power_iters=2))
    for _ in range(4):
        pkg.rg_step(ts, 16)
    sl = M.get_local_h(M.ising_h4(2.5))
    E, N = pkg.expect.get_gs_energy_4x4(ts, sl)
    Eo, _, no = P.get_gs_energy_4x4(_oracle_env(ts), [torch.from_numpy(s) for s in sl])
    print(dtype, E, complex(Eo))
    assert abs(E - complex(Eo)) < 1e-12 * abs(complex(Eo))
    assert abs(N - complex(no)) < 1e-12 * abs(complex(no))
