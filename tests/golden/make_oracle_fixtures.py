"""Oracle outputs for the configurations whose CPU runs take minutes, kept as
fixtures so the GPU tests can compare against them without running the
oracle there.  The oracle (oracle/ctmrg_oracle.py) follows the reference's
op order (M = R1 R2, full LAPACK SVD, the reference's svd_back), so these are
the reference semantics on synthetic inputs.

    python tests/golden/make_oracle_fixtures.py [c2] [c3]

* ``oracle_c2_D4_X64.npz``: Heisenberg (Jz=1), A (4,4,4,4,2) from
  numpy default_rng(1).standard_normal (Fortran order) scaled to unit max-abs,
  chi=64: the fg! contract of src/optimize/optim_gs.jl:53-91 with 40
  non-differentiated steps from init_ctm1 (min_iter = max_iter = 40) and 2
  differentiated steps -- E40 (energy after the 40 steps), E and the gradient g.
* ``oracle_c3_D6_X144.npz``: Heisenberg, A (6,6,6,6,2) from default_rng(2),
  chi=144 (N = 5184 in the second step): two rg_steps from init_ctm1 -- the
  left move's kept spectra s1, s2 and the energy E2 after the second step.
"""
import os
import sys
import time

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import ctmrg_oracle as O  # noqa: E402


def make_c2():
    A0 = O.seeded_tensor((4, 4, 4, 4, 2), 1, unit_max=True)
    H = O.heisenberg(1.0)
    cfg = dict(chi=64, min_iter=40, max_iter=40, ad_max_iter=2, rg_tol=0.0)
    x = O.renormalize(A0)
    ts, n_iter = O.get_conv_boundary(x, H, cfg)
    with torch.no_grad():
        E40, _ = O.get_gs_energy(ts, H)
    E, g, _ = O.gs_energy_grad(A0, H, cfg)
    np.savez(os.path.join(HERE, "oracle_c2_D4_X64.npz"), E40=np.complex128(complex(E40)), E=np.complex128(complex(E)),
             g=g.numpy(), n_iter=n_iter)
    print("c2", n_iter, complex(E40), complex(E), float(g.norm()))


def make_c3():
    A = O.seeded_tensor((6, 6, 6, 6, 2), 2)
    H = O.heisenberg(1.0)
    ts = O.make_ctm(A, dict(chi=144))
    with torch.no_grad():
        ts, s1 = O.rg_step(ts, 144)
        ts, s2 = O.rg_step(ts, 144)
        E2, _ = O.get_gs_energy(ts, H)
    np.savez(os.path.join(HERE, "oracle_c3_D6_X144.npz"), s1=s1.numpy(), s2=s2.numpy(),
             E2=np.complex128(complex(E2)), chis=np.array([c.shape for c in ts.Cs]))
    print("c3", complex(E2), s1.shape, s2.shape)


if __name__ == "__main__":
    torch.set_num_threads(len(os.sched_getaffinity(0)))
    which = sys.argv[1:] or ["c2", "c3"]
    for w in which:
        t0 = time.time()
        {"c2": make_c2, "c3": make_c3}[w]()
        print(w, f"{time.time() - t0:.0f} s", flush=True)
