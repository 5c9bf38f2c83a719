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
