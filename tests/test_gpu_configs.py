"""The large configurations against the oracle's reference semantics (full
LAPACK SVD of M = R1 R2 and the reference's svd_back), through committed
oracle outputs (tests/golden/make_oracle_fixtures.py: the CPU runs take
minutes), and the benchmark unit's precision.

Config 3 -- Heisenberg, A (6,6,6,6,2) from default_rng(2), chi = 144
(N = 5184).  The half-system spectrum of this random state is flat at the
cut (s_145 / s_144 = 0.99 in the benchmark's state), so a projector matches
the reference only once its range finder resolves the kept subspace: the
FP64 runs here use k = chi + 256 and 8 power iterations, and the adjoint's
discarded-subspace term by conjugate gradients.
Config 2 -- Heisenberg, A (4,4,4,4,2) from default_rng(1) at unit max-abs,
chi = 64 (N = 1024): the fg! contract with 40 steps and 2 differentiated ones.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

RESOLVED = dict(svd_mode=1, oversample=256, power_iters=8, complement_iters=60, complement_tol=1e-10)
BENCH = dict(svd_mode=1, oversample=144, power_iters=4)  # bench.py's default projector


def _c3(pkg):
    M = pkg.models
    return M.seeded_tensor((6, 6, 6, 6, 2), 2), M.heisenberg(1.0)


def _c3_state(pkg, steps=27):
    A, H = _c3(pkg)
    ts = pkg.CTMTensors(A, {}, dtype="f64", options=BENCH)
    for _ in range(steps):
        pkg.rg_step(ts, 144)
    return A, H, ts.Cs, ts.Es


def _unit(pkg, A, H, Cs, Es, dtype="f64", **opts):
    ts = pkg.CTMTensors(A, {}, dtype=dtype, options=opts)
    ts.set_env(Cs, Es)
    E, g = pkg.step_adjoint(ts, 144, H)
    return E.real, g


def test_c3_two_steps_match_oracle(pkg):
    """Two FP64 rg_steps from init_ctm1 (chi 36 -> 144, then N = 5184) with
    the resolved projector against the oracle: the left move's kept spectra
    and the energy after the second step at 1e-10."""
    ref = np.load(os.path.join(GOLDEN, "oracle_c3_D6_X144.npz"))
    A, H = _c3(pkg)
    ts = pkg.CTMTensors(A, {}, dtype="f64", options=RESOLVED)
    _, s1 = pkg.rg_step(ts, 144)
    _, s2 = pkg.rg_step(ts, 144)
    E, _ = pkg.get_gs_energy(ts, H)
    for s, so in ((s1, ref["s1"]), (s2, ref["s2"])):
        assert s.shape == so.shape
        err = np.abs(s - so).max() / so.max()
        print("spectrum", err)
        assert err < 1e-10, err
    print("E2", E.real, complex(ref["E2"]).real)
    assert abs(E.real - complex(ref["E2"]).real) < 1e-10


def test_c3_resolved_gradient_central_fd(pkg):
    """The resolved FP64 step + adjoint against central differences
    (eps = 1e-5) of the same forward on 8 entries: |fd - g| <= 1e-5 max|g|.
    (The benchmark's k = 2 chi, q = 4 projector is not smooth enough in A at
    this state for differences to resolve its gradient; its gradient is
    compared with this one below.)"""
    A, H, Cs, Es = _c3_state(pkg)
    E, g = _unit(pkg, A, H, Cs, Es, **RESOLVED)

    def f(x):
        ts = pkg.CTMTensors(x, {}, dtype="f64", options=RESOLVED)
        ts.set_env(Cs, Es)
        pkg.rg_step(ts, 144)
        return pkg.get_gs_energy(ts, H)[0].real

    assert abs(f(A) - E) < 1e-12
    gf = g.reshape(-1, order="F")
    scale = np.abs(gf).max()
    idx = np.random.default_rng(7).choice(A.size, 8, replace=False)
    eps = 1e-5
    worst = 0.0
    for k in idx:
        d = np.zeros(A.size)
        d[k] = eps
        d = d.reshape(A.shape, order="F")
        fd = (f(A + d) - f(A - d)) / (2 * eps)
        worst = max(worst, abs(fd - gf[k]) / scale)
    print("fd worst / max|g|", worst)
    assert worst < 1e-5, worst


def test_c3_benchmark_unit_precision(pkg):
    """The benchmark unit (bench.py's projector, mixed engine with FP32
    projector products) against the FP64 engine from the same state with the
    same test matrices: |dE/site| within the metric's 1e-7; the gradients of
    both engines within 3e-4 of the resolved FP64 gradient (the 1e-5 target is
    not met at this state: FP32 arithmetic moves the near-degenerate triplets
    at the cut)."""
    A, H, Cs, Es = _c3_state(pkg)
    E64, g64 = _unit(pkg, A, H, Cs, Es, **BENCH)
    Em, gm = _unit(pkg, A, H, Cs, Es, dtype="mixed", mixed_proj=6, **BENCH)
    Er, gr = _unit(pkg, A, H, Cs, Es, **RESOLVED)
    n = np.linalg.norm(gr)
    dE, dEr = abs(Em - E64), abs(E64 - Er)
    e64, em = np.linalg.norm(g64 - gr) / n, np.linalg.norm(gm - gr) / n
    print(f"dE mixed-fp64 {dE:.2e}, fp64-resolved {dEr:.2e}; grad fp64 {e64:.2e}, mixed {em:.2e}")
    assert dE < 1e-7 and abs(Em - Er) < 1e-7 and dEr < 1e-8
    assert e64 < 3e-4 and em < 3e-4


def test_c2_energy_and_gradient_match_oracle(pkg):
    """Config 2 through the fg! contract (gs_energy_grad: 40 steps from
    init_ctm1, the environment held constant, 2 differentiated steps + the
    energy) against the oracle's energy after the 40 steps, its energy after
    the differentiated steps and its autograd gradient (the reference's
    svd_back as a custom function)."""
    ref = np.load(os.path.join(GOLDEN, "oracle_c2_D4_X64.npz"))
    M = pkg.models
    A0 = M.seeded_tensor((4, 4, 4, 4, 2), 1, unit_max=True)
    H = M.heisenberg(1.0)
    cfg = dict(chi=64, min_iter=40, max_iter=40, ad_max_iter=2, rg_tol=0.0)
    opts = dict(svd_mode=1, oversample=448, power_iters=6, complement_iters=80, complement_tol=1e-12)
    x = A0 / np.abs(A0).max()
    ts = pkg.CTMTensors(x, cfg, dtype="f64", options=opts)
    pkg.run_ctm(ts, "energy", H)
    E40, _ = pkg.get_gs_energy(ts, H)
    E, g, iters = pkg.gs_energy_grad(A0, H, cfg, dtype="f64", options=opts)
    gr = ref["g"]
    dE40 = abs(E40 - complex(ref["E40"]))
    dE = abs(E - complex(ref["E"]))
    dg = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print(f"c2: iters {iters} dE40 {dE40:.2e} dE {dE:.2e} grad {dg:.2e}")
    assert dE40 < 1e-10 and dE < 1e-10
    assert dg < 1e-8
