"""Fused site absorption (op::absorb, absorb.hip): the mixed engine with the
FP32-MFMA kernel against the same engine with two split-bf16 GEMMs and
against FP64, from one FP64 environment with the same random test matrices
(the fused forward and both VJP kernels must run), and the op itself against
a float64 einsum and its autograd VJP."""
import numpy as np
import pytest

import ctmrg_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("D,chi", [(4, 16), (5, 20), (6, 24)])
def test_fused_absorb_step_adjoint(pkg, D, chi):
    A = O.seeded_tensor((D, D, D, D, 2), 2).numpy()
    H = [h.numpy() for h in O.heisenberg(1.0)]
    base = dict(svd_mode=1)
    t64 = pkg.CTMTensors(A, dtype="f64", options=base)
    for _ in range(3):
        pkg.rg_step(t64, chi)
    Cs, Es = t64.Cs, t64.Es
    res = {}
    for name, dt, ex in (("f64", "f64", {}), ("fused", "mixed", dict(fused_absorb=1, energy_mixed=6)),
                         ("gemm", "mixed", dict(fused_absorb=0, energy_mixed=6))):
        ctx = pkg.Context()
        ts = pkg.CTMTensors(A, dtype=dt, options=dict(base, **ex), ctx=ctx)
        ts.set_env(Cs, Es)
        ctx.set_timing(1)
        E, g = pkg.step_adjoint(ts, chi, H)
        st = ctx.kernel_stats(reset=True)
        res[name] = (E.real, g, st)
        del ts
        ctx.close()
    assert "absorb_fwd_kernel" in res["fused"][2] and "absorb_wgrad_kernel" in res["fused"][2]
    assert "absorb_fwd_kernel" not in res["gemm"][2]
    E64, g64 = res["f64"][0], res["f64"][1]
    n64 = np.linalg.norm(g64)
    errs = {k: (abs(v[0] - E64), np.linalg.norm(v[1] - g64) / n64) for k, v in res.items() if k != "f64"}
    print(D, chi, errs)
    # the step + adjoint of these random states amplifies rounding through the
    # projector (flat spectra at the cut), so both FP32 routes sit at a
    # similar level; the op itself is checked element-wise below
    for name in ("fused", "gemm"):
        assert errs[name][0] < 1e-6 and errs[name][1] < 1e-3, (name, errs[name])


PATTERNS = [  # (lx, la, lad) of the eight enlarged corners (left LU / LB, right RU / RB, top, bottom)
    ([-1, -4, 11, 12, 21, 22], [21, 22, -2, -5, 3], [11, 12, -3, -6, 3]),
    ([-1, -4, 11, 12, 21, 22], [-2, 21, 22, -5, 3], [-3, 11, 12, -6, 3]),
    ([-1, -4, 11, 12, 21, 22], [21, -2, -5, 22, 3], [11, -3, -6, 12, 3]),
    ([-1, -4, 11, 12, 21, 22], [-2, -5, 21, 22, 3], [-3, -6, 11, 12, 3]),
    ([-3, 11, -1, 21, 12, 22], [21, 22, -2, -5, 3], [11, 12, -4, -6, 3]),  # permuted X legs / outputs
]


def _einsum_ref(X, lx, A, la, Ad, lad):
    letters = {}
    def s(ls):
        out = ""
        for l in ls:
            if l not in letters:
                letters[l] = chr(ord("a") + len(letters))
            out += letters[l]
        return out
    sx, sa, sd = s(lx), s(la), s(lad)
    nout = sum(1 for l in lx + la + lad if l < 0)
    so = "".join(letters[-k] for k in range(1, nout + 1))
    return np.einsum(f"{sx},{sa},{sd}->{so}", X, A, Ad)


@pytest.mark.parametrize("D", [4, 5, 6])
@pytest.mark.parametrize("pat", range(len(PATTERNS)))
def test_absorb_op_forward_and_vjp(pkg, D, pat):
    """op::absorb (fused FP32-MFMA kernel and the two-contraction route) on
    float32 data against a float64 einsum: Y and the VJP of X, A and Ad at
    relative 2e-6 (fused) / 5e-5 (split-bf16 GEMMs), ragged row extents."""
    lx, la, lad = PATTERNS[pat]
    rng = np.random.default_rng(D * 10 + pat)
    chi1, chi2 = 17, 5
    # X legs: open legs get chi1 / chi2, shared legs D
    shX = []
    opens = [chi1, chi2]
    for l in lx:
        shX.append(opens.pop(0) if l < 0 else D)
    X = rng.standard_normal(shX).astype(np.float32)
    A = rng.standard_normal((D, D, D, D, 2)).astype(np.float32)
    Ad = rng.standard_normal((D, D, D, D, 2)).astype(np.float32)
    Y64 = _einsum_ref(X.astype(np.float64), lx, A.astype(np.float64), la, Ad.astype(np.float64), lad)
    G = rng.standard_normal(Y64.shape).astype(np.float32)
    # reference VJP by einsum: gX = G A Ad with X's legs open, etc.
    import torch
    Xt, At, Adt = (torch.from_numpy(x.astype(np.float64)).requires_grad_() for x in (X, A, Ad))
    letters = {}
    def s(ls):
        return "".join(letters.setdefault(l, chr(ord("a") + len(letters))) for l in ls)
    sx, sa, sd = s(lx), s(la), s(lad)
    so = "".join(letters[-k] for k in range(1, 7))
    Yt = torch.einsum(f"{sx},{sa},{sd}->{so}", Xt, At, Adt)
    Yt.backward(torch.from_numpy(G.astype(np.float64)))
    refs = [Y64, Xt.grad.numpy(), At.grad.numpy(), Adt.grad.numpy()]
    for fused, tol in ((True, 2e-6), (False, 5e-5)):
        Y, gX, gA, gAd = pkg.absorb(X, lx, A, la, Ad, lad, G=G, fused=fused)
        for got, ref, name in zip((Y, gX, gA, gAd), refs, ("Y", "gX", "gA", "gAd")):
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < tol, (fused, name, err)
