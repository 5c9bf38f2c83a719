"""CPU tests of host-side logic: workload generators and option tables."""
import numpy as np


def test_c4_instances_deterministic(pkg):
    a = pkg.c4_instances()
    b = pkg.c4_instances()
    assert len(a) == 64
    Ds = {x[0].shape[0] for x in a}
    chis = {x[1] for x in a}
    assert Ds == {2, 3, 4} and chis == {16, 32, 48, 64, 96}
    assert {x[3] for x in a} == {"ising", "heisenberg"}
    for (A1, c1, H1, k1), (A2, c2, H2, k2) in zip(a, b):
        assert c1 == c2 and k1 == k2
        np.testing.assert_array_equal(A1, A2)
        # float32 values stored as float64
        np.testing.assert_array_equal(A1, A1.astype(np.float32).astype(np.float64))
    # instance i uses seed 3 + i (Fortran order)
    A5 = a[5][0]
    D = A5.shape[0]
    v = np.random.default_rng(8).standard_normal(D ** 4 * 2).astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(A5.reshape(-1, order="F"), v)


def test_env_options_documented(pkg):
    import os
    from conftest import ROOT
    hdr = open(os.path.join(ROOT, "include", "ipeps_hip.h")).read()
    for k in pkg.ENV_OPTIONS:
        assert f'"{k}"' in hdr, k


def test_act_Q_op_reproduces_loop_gas_state(pkg):
    """act_Q_op (src/model/lee_gs.jl:54-73) on the unit site tensor carrying
    s111 (x) s111 is the ZZ loop-gas state of init_hb_gs (lee_gs.jl:2-28), so
    that state lies in the span of the no-flux columns of the gauge basis;
    the sigma^z-flux image differs from it."""
    import math
    M = pkg.models
    u = 1 / math.sqrt(3)
    s = np.array([1 + u, u + 1j * u]) / math.sqrt(2 + 2 * u)
    B0 = np.kron(s, s).reshape(1, 1, 1, 1, 4, order="F")  # (q1, q2) column-major
    A = M.act_Q_op(B0, add=0)
    ref = M.init_hb_gs(2, "ZZ")
    assert A.shape == ref.shape == (2, 2, 2, 2, 4)
    np.testing.assert_allclose(A, ref, atol=1e-14)
    cols = np.stack([M.act_Q_op(np.eye(4)[i].reshape(1, 1, 1, 1, 4, order="F"), 0).reshape(-1, order="F")
                     for i in range(4)], axis=1)
    x = np.linalg.lstsq(cols, ref.reshape(-1, order="F"), rcond=None)[0]
    assert np.abs(cols @ x - ref.reshape(-1, order="F")).max() < 1e-13
    Az = M.act_Q_op(B0, add=3)
    assert np.abs(Az - ref).max() > 1e-3


def test_plaquette_models_match_oracle(pkg):
    """The plaquette Hamiltonians and get_local_h of models.py against the
    oracle's restatement: equal tensors, and factors whose product is h."""
    import torch
    import plaquette_oracle as P
    M = pkg.models
    np.testing.assert_array_equal(M.ising_h4(2.5), P.ising_h4(2.5).numpy())
    np.testing.assert_allclose(M.hb_h4_ZZ(), P.hb_h4_ZZ().numpy(), atol=0)
    for h in (M.ising_h4(2.5), M.hb_h4_ZZ(), M.honeycomb_h4().reshape(16, 16, 16, 16, order="F")):
        s1, s2, s3, s4 = M.get_local_h(h)
        d = int(round(h.shape[0] ** 0.5))
        rec = np.einsum("abi,icdj,jefk,kgh->abcdefgh", s1, s2, s3, s4)
        np.testing.assert_allclose(rec, h.reshape((d,) * 8, order="F"), atol=1e-12)
