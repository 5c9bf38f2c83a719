"""optim_gs (src/optimize/optim_gs.jl:14-99) end to end on the GPU: the
LBFGS loop over the engine's fg!, its per-iteration checkpoint and resume,
on the reference's own loop-gas run (test/optim_gs.jl:27-31), whose first
energy and gradient norm the committed gs.jld2 holds (G1, G2)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_optim_gs_loopgas(pkg, tmp_path):
    ref = np.load(os.path.join(GOLDEN, "gs_loopgas_D2_X32.npz"))
    M = pkg.models
    A0 = M.init_hb_gs(2, "XX")
    H = M.honeycomb(1, 1, dir="XX")
    cfg = dict(chi=32, min_iter=4, max_iter=20, ad_max_iter=4, rg_tol=1e-6)
    path = str(tmp_path / "gs.npz")
    r = pkg.optim_gs(H, A0, cfg, iterations=2, save_path=path, verbose=True)
    E = r["energies"]
    # the first iterate is A0 itself: the reference's energies[1], gradnorms[1]
    assert abs(E[0] - ref["energies"][0]) < 1e-10
    assert abs(r["gradnorms"][0] - ref["gradnorms"][0]) < 1e-8 * ref["gradnorms"][0]
    assert len(E) == 3 and np.all(np.diff(E) < 0)
    z = np.load(path)
    np.testing.assert_array_equal(z["energies"], E)
    np.testing.assert_array_equal(z["A"], r["A"])
    # resume from the checkpoint (the reference's cfg["resume"]) and go on
    r2 = pkg.optim_gs(H, A0, cfg, iterations=4, save_path=path, resume=True, verbose=True)
    E2 = r2["energies"]
    print("energies", E2, "reference", ref["energies"])
    assert len(E2) >= len(E) + 2 and np.array_equal(E2[:len(E)], E)
    assert np.all(np.diff(E2) < 0)
    # the reference's LBFGS reached -1.39506849 (G3) at its seventh iterate
    assert E2[-1] < -1.3950
