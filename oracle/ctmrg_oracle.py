"""CPU FP64 restatement of the reference CTMRG hot path -- TEST INFRASTRUCTURE.

This module is the oracle that the HIP engine is checked against.  It is
imported only by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py``; the product path never calls it.

It restates, in the reference's own index lists and op order, the Julia
functions of yitan1/iPEPS.jl (file:line citations are into that repository):

* ``tcon``                     src/utility/wrapper.jl:15-23
* ``renormalize``              src/utility/wrapper.jl:57
* ``wrap_svd``/``cutoff_matrix`` src/utility/wrapper.jl:25-55
* ``svd1`` + ``svd_back``      src/utility/autodiff.jl:22-75
* ``init_ctm1``                src/ctm/ctm_tensor.jl:97-130
* ``run_ctm``/``rg_step``      src/ctm/ctmrg.jl:1-64
* ``get_projector_*``          src/ctm/ctmrg.jl:122-265
* ``proj_*``                   src/ctm/ctmrg.jl:267-375
* ``get_gs_energy``/``get_dms``/``get_dm_h``/``get_dm_v``
                               src/graph/evaluation.jl:15-29,124-248
* ``get_single_dm``/``get_gs_norm`` src/graph/evaluation.jl:47-59,110-122
* gradient contract of ``optim_gs``'s ``fg!`` / ``run_gs``
                               src/optimize/optim_gs.jl:53-91,120-143

Arrays are torch tensors on the CPU in float64/complex128 and are indexed like
the Julia arrays (X[i1, i2, ...] == Julia X[i1+1, i2+1, ...]).  Julia's
column-major ``reshape`` is ``reshape_f``.  Reverse mode uses torch autograd
with the reference's own SVD adjoint (Lorentzian F, eta = 1e-12) installed as
a custom function, so the gradient follows the same graph Zygote walks.

Pinning (see tests/test_oracle.py): energies G1/G3 and the gradient norm G2 of
the reference's committed gs.jld2, the fixed-environment energy G4, the
KAT -0.16349 per site, the D=1 product-state closed form, the scale/gauge
invariances, and central finite differences of the gradient.
"""
import math

import numpy as np
import torch

ETA = 1e-12          # src/utility/autodiff.jl:41
CUTOFF = 1e-12       # src/utility/wrapper.jl:32
DEGEN_TOL = 1e-10    # src/utility/wrapper.jl:46

# --------------------------------------------------------------------------
# primitives
# --------------------------------------------------------------------------


def reshape_f(x, *shape):
    """Julia (column-major) reshape."""
    if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
        shape = tuple(shape[0])
    shape = list(shape)
    if -1 in shape:
        k = shape.index(-1)
        rest = 1
        for i, s in enumerate(shape):
            if i != k:
                rest *= s
        shape[k] = x.numel() // rest
    xr = x.permute(*reversed(range(x.dim()))) if x.dim() > 1 else x
    y = xr.reshape(tuple(reversed(shape)))
    return y.permute(*reversed(range(len(shape)))) if len(shape) > 1 else y


def tcon(xs, inds):
    """ncon-style contraction (src/utility/wrapper.jl:15-23).

    Positive labels are summed, negative labels are the output legs in the
    order -1, -2, ...
    """
    letters = {}

    def let(l):
        if l not in letters:
            n = len(letters)
            letters[l] = chr(ord("a") + n) if n < 26 else chr(ord("A") + n - 26)
        return letters[l]

    subs = ["".join(let(l) for l in ind) for ind in inds]
    nout = sum(1 for ind in inds for l in ind if l < 0)
    out = "".join(let(-k) for k in range(1, nout + 1))
    return torch.einsum(",".join(subs) + "->" + out, *xs)


def flat_f(x):
    """Column-major flattening (Julia linear index order)."""
    return x.permute(*reversed(range(x.dim()))).reshape(-1) if x.dim() > 1 else x.reshape(-1)


def renormalize(x):
    """x ./ maximum(abs.(x)) (src/utility/wrapper.jl:57).

    The gradient of ``maximum`` goes to the first maximal element in Julia's
    column-major order, as ChainRules' findmax-based rule does.
    """
    a = flat_f(x).abs()
    idx = int(torch.argmax(a.detach()))
    return x / a[idx]


class _SVD1(torch.autograd.Function):
    """svd1 with the reference adjoint svd_back (src/utility/autodiff.jl:22-75)."""

    @staticmethod
    def forward(ctx, A):
        U, S, Vh = torch.linalg.svd(A, full_matrices=True)
        V = Vh.mH.contiguous()
        ctx.save_for_backward(U, S, V)
        return U, S, V

    @staticmethod
    def backward(ctx, dU, dS, dV):
        U, S, V = ctx.saved_tensors
        return svd_back(U, S, V, dU, dS, dV)


def svd_back(U, S, V, dU, dS, dV, eta=ETA):
    """Reference SVD adjoint (src/utility/autodiff.jl:41-75)."""
    S2 = S * S
    Sinv = S / (S2 + eta)
    F = S2[None, :] - S2[:, None]
    F = F / (F * F + eta)
    cdt = U.dtype
    res = torch.zeros(U.shape[1], V.shape[1], dtype=cdt)
    if dU is not None:
        UdU = U.mH @ dU
        J = F.to(cdt) * UdU
        res = res + (J + J.mH) @ torch.diag(S).to(cdt)
        if cdt.is_complex:
            res = res + torch.diag(1j * torch.diagonal(UdU).imag * Sinv)
    if dV is not None:
        VdV = V.mH @ dV
        K = F.to(cdt) * VdV
        res = res + torch.diag(S).to(cdt) @ (K + K.mH)
    if dS is not None:
        res = res + torch.diag(dS).to(cdt)
    res = U @ res @ V.mH
    if dU is not None and U.shape[0] != U.shape[1]:
        res = res + (dU - U @ (U.mH @ dU)) @ torch.diag(Sinv).to(cdt) @ V.mH
    if dV is not None and V.shape[0] != V.shape[1]:
        res = res + U @ torch.diag(Sinv).to(cdt) @ (dV.mH - (dV.mH @ V) @ V.mH)
    return res


def svd1(A):
    return _SVD1.apply(A)


def cutoff_count(s, n):
    """The truncation rank rule of cutoff_matrix (src/utility/wrapper.jl:37-55)."""
    smax = float(s.max())
    n_above = int((s / smax > CUTOFF).sum())
    n = min(n, n_above)
    if s.shape[0] > n and abs(float(s[n]) - float(s[n - 1])) < DEGEN_TOL:
        n = n + 1
    return n


def wrap_svd(A, n=math.inf):
    """Sorted, truncated SVD (src/utility/wrapper.jl:25-35)."""
    u, s, v = svd1(A)
    ix = np.argsort(s.detach().numpy(), kind="stable")[::-1].copy()
    ix = torch.from_numpy(ix)
    s = s[ix]
    u = u[:, ix]
    v = v[:, ix]
    if n == math.inf:
        n = s.shape[0]
    k = cutoff_count(s.detach(), int(n))
    return u[:, :k], s[:k], v[:, :k]


# --------------------------------------------------------------------------
# environment
# --------------------------------------------------------------------------


class CTMTensors:
    """Plain (ground-state) environment (src/ctm/ctm_tensor.jl:3-18)."""

    def __init__(self, A, Ad, Cs, Es, params=None):
        self.A, self.Ad = A, Ad
        self.Cs, self.Es = list(Cs), list(Es)
        self.params = dict(params or {})

    def replace(self, **kw):
        t = CTMTensors(self.A, self.Ad, self.Cs, self.Es, self.params)
        for k, v in kw.items():
            setattr(t, k, v)
        return t


def init_ctm1(A, Ad):
    """Boundary initialisation from A and conj(A) (src/ctm/ctm_tensor.jl:97-130)."""
    D = A.shape[0]
    C1 = reshape_f(tcon([A, Ad], [[1, 2, -1, -3, 3], [1, 2, -2, -4, 3]]), D * D, D * D)
    C2 = reshape_f(tcon([A, Ad], [[1, -1, -3, 2, 3], [1, -2, -4, 2, 3]]), D * D, D * D)
    C3 = reshape_f(tcon([A, Ad], [[-1, -3, 1, 2, 3], [-2, -4, 1, 2, 3]]), D * D, D * D)
    C4 = reshape_f(tcon([A, Ad], [[-1, 1, 2, -3, 3], [-2, 1, 2, -4, 3]]), D * D, D * D)
    Cs = [renormalize(c) for c in (C1, C2, C3, C4)]
    E1 = reshape_f(tcon([A, Ad], [[1, -1, -5, -3, 2], [1, -2, -6, -4, 2]]), D * D, D * D, D, D)
    E2 = reshape_f(tcon([A, Ad], [[-1, -5, -3, 1, 2], [-2, -6, -4, 1, 2]]), D * D, D * D, D, D)
    E3 = reshape_f(tcon([A, Ad], [[-5, -1, 1, -3, 2], [-6, -2, 1, -4, 2]]), D * D, D * D, D, D)
    E4 = reshape_f(tcon([A, Ad], [[-1, 1, -3, -5, 2], [-2, 1, -4, -6, 2]]), D * D, D * D, D, D)
    Es = [renormalize(e) for e in (E1, E2, E3, E4)]
    return Cs, Es


def make_ctm(A, params=None):
    """CTMTensors(A, cfg) (src/ctm/ctm_tensor.jl:63-69)."""
    Ad = A.conj()
    Cs, Es = init_ctm1(A, Ad)
    return CTMTensors(A, Ad, Cs, Es, params)


# --------------------------------------------------------------------------
# projectors (src/ctm/ctmrg.jl:122-265)
# --------------------------------------------------------------------------


def get_projector(R1, R2, chi):
    """src/ctm/ctmrg.jl:254-265."""
    new_chi = min(chi, R1.shape[1])
    U, S, V = wrap_svd(R1 @ R2, new_chi)
    isq = torch.diag(1.0 / torch.sqrt(S)).to(R1.dtype)
    P1 = R2 @ V @ isq
    P2 = isq @ U.mH @ R1
    return P1, P2, S


def _mat(x, rows_ndim):
    sz = list(x.shape)
    r = 1
    for s in sz[:rows_ndim]:
        r *= s
    return reshape_f(x, r, -1)


def get_projector_left(ts, chi):
    C1, C4 = ts.Cs[0], ts.Cs[3]
    E1, E3, E4 = ts.Es[0], ts.Es[2], ts.Es[3]
    A, Ad = ts.A, ts.Ad
    C1E1 = tcon([C1, E1], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C1E1, E4], [[1, -2, -5, -3], [1, -1, -6, -4]])
    CEEA = tcon([CEE4, A], [[-1, -2, -3, -4, 1, 2], [1, 2, -5, -6, -7]])
    LU = tcon([CEEA, Ad], [[-1, -4, 1, 2, -2, -5, 3], [1, 2, -3, -6, 3]])
    R1 = _mat(LU, 3).transpose(0, 1)
    C4E3 = tcon([C4, E3], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C4E3, E4], [[1, -2, -6, -4], [-1, 1, -5, -3]])
    CEEA = tcon([CEE4, A], [[-1, -2, -3, -4, 1, 2], [-5, 1, 2, -6, -7]])
    LB = tcon([CEEA, Ad], [[-1, -4, 1, 2, -2, -5, 3], [-3, 1, 2, -6, 3]])
    R2 = _mat(LB, 3)
    return get_projector(R1, R2, chi)


def get_projector_right(ts, chi):
    C2, C3 = ts.Cs[1], ts.Cs[2]
    E1, E2, E3 = ts.Es[0], ts.Es[1], ts.Es[2]
    A, Ad = ts.A, ts.Ad
    C2E1 = tcon([C2, E1], [[1, -2], [-1, 1, -3, -4]])
    CEE2 = tcon([C2E1, E2], [[-1, 1, -5, -3], [1, -2, -6, -4]])
    CEEA = tcon([CEE2, A], [[-1, -2, -3, -4, 1, 2], [1, -5, -6, 2, -7]])
    RU = tcon([CEEA, Ad], [[-1, -4, 1, 2, -2, -5, 3], [1, -3, -6, 2, 3]])
    R1 = _mat(RU, 3)
    C3E3 = tcon([C3, E3], [[-1, 1], [-2, 1, -3, -4]])
    CEE2 = tcon([C3E3, E2], [[1, -2, -5, -3], [-1, 1, -6, -4]])
    CEEA = tcon([CEE2, A], [[-1, -2, -5, -6, 1, 2], [-3, -4, 1, 2, -7]])
    RB = tcon([CEEA, Ad], [[-1, -4, -2, -5, 1, 2, 3], [-3, -6, 1, 2, 3]])
    R2 = _mat(RB, 3)
    return get_projector(R1, R2, chi)


def get_projector_top(ts, chi):
    C1, C2 = ts.Cs[0], ts.Cs[1]
    E1, E2, E4 = ts.Es[0], ts.Es[1], ts.Es[3]
    A, Ad = ts.A, ts.Ad
    C1E4 = tcon([C1, E4], [[1, -2], [1, -1, -3, -4]])
    CEE1 = tcon([C1E4, E1], [[-1, 1, -6, -4], [1, -2, -5, -3]])
    CEEA = tcon([CEE1, A], [[-1, -2, -5, -6, 1, 2], [1, 2, -3, -4, -7]])
    UL = tcon([CEEA, Ad], [[-1, -4, -2, -5, 1, 2, 3], [1, 2, -3, -6, 3]])
    R1 = _mat(UL, 3)
    C2E2 = tcon([C2, E2], [[-1, 1], [1, -2, -3, -4]])
    CEE1 = tcon([C2E2, E1], [[1, -2, -6, -4], [-1, 1, -5, -3]])
    CEEA = tcon([CEE1, A], [[-1, -2, -5, -6, 1, 2], [1, -3, -4, 2, -7]])
    UR = tcon([CEEA, Ad], [[-1, -4, -2, -5, 1, 2, 3], [1, -3, -6, 2, 3]])
    R2 = _mat(UR, 3)
    return get_projector(R1, R2, chi)


def get_projector_bottom(ts, chi):
    C3, C4 = ts.Cs[2], ts.Cs[3]
    E2, E3, E4 = ts.Es[1], ts.Es[2], ts.Es[3]
    A, Ad = ts.A, ts.Ad
    C4E4 = tcon([C4, E4], [[1, -2], [-1, 1, -3, -4]])
    CEE3 = tcon([C4E4, E3], [[-1, 1, -5, -3], [1, -2, -6, -4]])
    CEEA = tcon([CEE3, A], [[-1, -2, -5, -6, 1, 2], [-3, 1, 2, -4, -7]])
    BL = tcon([CEEA, Ad], [[-1, -4, -2, -5, 1, 2, 3], [-3, 1, 2, -6, 3]])
    R1 = _mat(BL, 3)
    C3E2 = tcon([C3, E2], [[1, -2], [-1, 1, -3, -4]])
    CEE3 = tcon([C3E2, E3], [[-1, 1, -6, -4], [-2, 1, -5, -3]])
    CEEA = tcon([CEE3, A], [[-1, -2, -5, -6, 1, 2], [-3, -4, 1, 2, -7]])
    BR = tcon([CEEA, Ad], [[-1, -4, -2, -5, 1, 2, 3], [-3, -6, 1, 2, 3]])
    R2 = _mat(BR, 3).transpose(0, 1)
    return get_projector(R1, R2, chi)


# --------------------------------------------------------------------------
# renormalisation (src/ctm/ctmrg.jl:267-375)
# --------------------------------------------------------------------------


EDGE_ORDER = "reference"   # or "canonical": _edge_canonical (CPU-baseline timing variant)


def _edge_canonical(E, A, Ad, ia, iad, P, Pd):
    """The edge update of _edge in the canonical pairwise order: Pd with E
    (chi^3 D^4), then A and Ad (2 chi^2 D^6 d), then P (chi^3 D^4), instead of
    the reference's chi^2 D^6 intermediate E A Ad.  Same result to rounding
    (tests/test_oracle.py pins it against _edge)."""
    # legs of the reference's EAAd (src/ctm/ctmrg.jl:285-289): o1 = E leg 1,
    # o2 / o5 / o7 = A's labels -2 / -4 / -5, o3 / o6 / o8 = Ad's -3 / -6 / -8,
    # o4 = E leg 2; E leg 3 meets A's label 1, E leg 4 meets Ad's label 1, and
    # A's -7 meets Ad's label 2 (the physical leg)
    la = {-2: "b", -4: "d", -5: "g", -7: "p", 1: "x"}
    lad = {-3: "e", -6: "h", -8: "k", 1: "y", 2: "p"}
    sa = "".join(la[l] for l in ia)
    sd = "".join(lad[l] for l in iad)
    chi1, D = Pd.shape[0], A.shape[0]
    Pd4 = reshape_f(Pd, chi1, E.shape[0], D, D)            # (i, a, b, e)
    P4 = reshape_f(P, E.shape[1], D, D, P.shape[1])         # (c, d, h, j)
    T1 = torch.einsum("iabe,acxy->ibecxy", Pd4, E)
    T2 = torch.einsum(f"ibecxy,{sa}->iecydgp", T1, A)
    T3 = torch.einsum(f"iecydgp,{sd}->icdghk", T2, Ad)
    return torch.einsum("icdghk,cdhj->ijgk", T3, P4)


def _edge(E, A, Ad, ia, iad, P, Pd):
    if EDGE_ORDER == "canonical":
        return _edge_canonical(E, A, Ad, ia, iad, P, Pd)
    EA = tcon([E, A], [[-1, -3, 1, -6], ia])
    EAAd = tcon([EA, Ad], [[-1, -2, -4, -5, -7, 1, 2], iad])
    s = EAAd.shape
    EAAd = reshape_f(EAAd, s[0] * s[1] * s[2], s[3] * s[4] * s[5], s[6], s[7])
    EPd = tcon([Pd, EAAd], [[-1, 1], [1, -2, -3, -4]])
    return tcon([EPd, P], [[-1, 1, -3, -4], [1, -2]])


def proj_left(ts, P, Pd):
    C1, C4 = ts.Cs[0], ts.Cs[3]
    E1, E3, E4 = ts.Es[0], ts.Es[2], ts.Es[3]
    C1E1 = tcon([C1, E1], [[-2, 1], [1, -1, -3, -4]])
    newC1 = (reshape_f(C1E1, C1E1.shape[0], -1) @ P).transpose(0, 1)
    C4E3 = tcon([C4, E3], [[-1, 1], [1, -4, -2, -3]])
    newC4 = Pd @ reshape_f(C4E3, -1, C4E3.shape[3])
    newE4 = _edge(E4, ts.A, ts.Ad, [-2, 1, -4, -5, -7], [-3, 1, -6, -8, 2], P, Pd)
    return newC1, newE4, newC4


def proj_right(ts, P, Pd):
    C2, C3 = ts.Cs[1], ts.Cs[2]
    E1, E2, E3 = ts.Es[0], ts.Es[1], ts.Es[2]
    C2E1 = tcon([C2, E1], [[1, -2], [-1, 1, -3, -4]])
    newC2 = reshape_f(C2E1, C2E1.shape[0], -1) @ P
    C3E3 = tcon([C3, E3], [[-1, 1], [-4, 1, -2, -3]])
    newC3 = Pd @ reshape_f(C3E3, -1, C3E3.shape[3])
    newE2 = _edge(E2, ts.A, ts.Ad, [-2, -5, -4, 1, -7], [-3, -8, -6, 1, 2], P, Pd)
    return newC2, newE2, newC3


def proj_top(ts, P, Pd):
    C1, C2 = ts.Cs[0], ts.Cs[1]
    E1, E2, E4 = ts.Es[0], ts.Es[1], ts.Es[3]
    C1E4 = tcon([C1, E4], [[1, -2], [1, -1, -3, -4]])
    newC1 = reshape_f(C1E4, C1E4.shape[0], -1) @ P
    C2E2 = tcon([C2, E2], [[-1, 1], [1, -4, -2, -3]])
    newC2 = Pd @ reshape_f(C2E2, -1, C2E2.shape[3])
    newE1 = _edge(E1, ts.A, ts.Ad, [1, -2, -5, -4, -7], [1, -3, -8, -6, 2], P, Pd)
    return newC1, newE1, newC2


def proj_bottom(ts, P, Pd):
    C3, C4 = ts.Cs[2], ts.Cs[3]
    E2, E3, E4 = ts.Es[1], ts.Es[2], ts.Es[3]
    C4E4 = tcon([C4, E4], [[1, -2], [-1, 1, -3, -4]])
    newC4 = reshape_f(C4E4, C4E4.shape[0], -1) @ P
    C3E2 = tcon([C3, E2], [[1, -1], [-4, 1, -2, -3]])
    newC3 = (Pd @ reshape_f(C3E2, -1, C3E2.shape[3])).transpose(0, 1)
    newE3 = _edge(E3, ts.A, ts.Ad, [-5, -2, 1, -4, -7], [-8, -3, 1, -6, 2], P, Pd)
    return newC4, newE3, newC3


def left_rg(ts, chi):
    P, Pd, s = get_projector_left(ts, chi)
    C1, E4, C4 = (renormalize(x) for x in proj_left(ts, P, Pd))
    Cs = [C1, ts.Cs[1], ts.Cs[2], C4]
    Es = [ts.Es[0], ts.Es[1], ts.Es[2], E4]
    return ts.replace(Cs=Cs, Es=Es), s


def right_rg(ts, chi):
    P, Pd, s = get_projector_right(ts, chi)
    C2, E2, C3 = (renormalize(x) for x in proj_right(ts, P, Pd))
    Cs = [ts.Cs[0], C2, C3, ts.Cs[3]]
    Es = [ts.Es[0], E2, ts.Es[2], ts.Es[3]]
    return ts.replace(Cs=Cs, Es=Es), s


def top_rg(ts, chi):
    P, Pd, s = get_projector_top(ts, chi)
    C1, E1, C2 = (renormalize(x) for x in proj_top(ts, P, Pd))
    Cs = [C1, C2, ts.Cs[2], ts.Cs[3]]
    Es = [E1, ts.Es[1], ts.Es[2], ts.Es[3]]
    return ts.replace(Cs=Cs, Es=Es), s


def bottom_rg(ts, chi):
    P, Pd, s = get_projector_bottom(ts, chi)
    C4, E3, C3 = (renormalize(x) for x in proj_bottom(ts, P, Pd))
    Cs = [ts.Cs[0], ts.Cs[1], C3, C4]
    Es = [ts.Es[0], ts.Es[1], E3, ts.Es[3]]
    return ts.replace(Cs=Cs, Es=Es), s


def rg_step(ts, chi):
    """left -> right -> top -> bottom; returns left's spectrum (src/ctm/ctmrg.jl:57-64)."""
    ts, s = left_rg(ts, chi)
    ts, _ = right_rg(ts, chi)
    ts, _ = top_rg(ts, chi)
    ts, _ = bottom_rg(ts, chi)
    return ts, s


def run_ctm(ts, conv_fun=None, verbose=False):
    """src/ctm/ctmrg.jl:1-55 with its exact stopping rule.

    Returns (ts, conv, n_iter, diffs).
    """
    p = ts.params
    chi = p.get("chi", 30)
    min_iter = p.get("min_iter", 4)
    max_iter = p.get("max_iter", 20)
    tol = p.get("rg_tol", 1e-6)
    diffs = [1.0]
    conv = torch.tensor(1.0, dtype=torch.float64)
    it = 0
    for i in range(1, max_iter + 1):
        it = i
        ts, s = rg_step(ts, chi)
        old_conv = conv
        with torch.no_grad():
            if conv_fun is not None:
                conv = conv_fun(ts).detach().reshape(-1)
            else:
                conv = s.detach().reshape(-1)
            oc = old_conv.reshape(-1)
            if conv.numel() == oc.numel():
                diffs.append(float(torch.linalg.vector_norm(conv.to(torch.complex128) - oc.to(torch.complex128))))
        if verbose:
            print(f"CTM step {i:3d}, diff = {diffs[-1]:.4e}")
        if i >= min_iter and diffs[-1] < tol:
            break
    return ts, conv, it, diffs


# --------------------------------------------------------------------------
# observables (src/graph/evaluation.jl)
# --------------------------------------------------------------------------


def get_dm_h(C1, C2, C3, C4, E1l, E1r, E2, E3l, E3r, E4, Al, Ar, Adl, Adr):
    """src/graph/evaluation.jl:169-200."""
    C1E1 = tcon([C1, E1l], [[-1, 1], [1, -2, -3, -4]])
    LU = tcon([C1E1, Adl], [[-1, -2, -3, 1], [1, -4, -5, -6, -7]])
    C4E3 = tcon([C4, E3l], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([E4, C4E3], [[-1, 1, -3, -5], [1, -2, -4, -6]])
    LB = tcon([CEE4, Al], [[-1, -2, 1, 2, -4, -5], [-3, 1, 2, -6, -7]])
    L = tcon([LB, LU], [[1, -2, 2, 3, 4, -3, -5], [1, -1, 2, 3, 4, -4, -6]])
    C2E1 = tcon([C2, E1r], [[1, -2], [-1, 1, -3, -4]])
    RU = tcon([C2E1, Ar], [[-1, -2, 1, -3], [1, -4, -5, -6, -7]])
    C3E3 = tcon([C3, E3r], [[-1, 1], [-2, 1, -3, -4]])
    CEE2 = tcon([C3E3, E2], [[1, -2, -3, -5], [-1, 1, -4, -6]])
    RB = tcon([CEE2, Adr], [[-1, -2, -3, -4, 1, 2], [-5, -6, 1, 2, -7]])
    R = tcon([RU, RB], [[-1, 1, 2, -3, 3, 4, -5], [1, -2, 3, 4, 2, -4, -6]])
    roh = tcon([L, R], [[1, 2, 3, 4, -1, -3], [1, 2, 3, 4, -2, -4]])
    d = Al.shape[4]
    return reshape_f(roh, d * d, d * d)


def get_dm_v(C1, C2, C3, C4, E1, E2u, E2d, E3, E4u, E4d, Au, Ad, Adu, Add):
    """src/graph/evaluation.jl:217-248."""
    C1E1 = tcon([C1, E1], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C1E1, E4u], [[1, -2, -3, -5], [1, -1, -4, -6]])
    UL = tcon([CEE4, Au], [[-1, -2, 1, 2, -3, -4], [1, 2, -5, -6, -7]])
    C2E2 = tcon([C2, E2u], [[-1, 1], [1, -2, -3, -4]])
    UR = tcon([C2E2, Adu], [[-1, -2, -3, 1], [-4, -5, -6, 1, -7]])
    U = tcon([UL, UR], [[-1, 1, 3, 4, -3, 2, -5], [1, -2, 2, 3, 4, -4, -6]])
    C4E3 = tcon([C4, E3], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = tcon([C4E3, E4d], [[1, -2, -4, -6], [-1, 1, -3, -5]])
    BL = tcon([CEE4, Add], [[-1, -2, -3, -4, 1, 2], [-5, 1, 2, -6, -7]])
    C3E2 = tcon([C3, E2d], [[1, -2], [-1, 1, -3, -4]])
    BR = tcon([C3E2, Ad], [[-1, -2, 1, -3], [-4, -5, -6, 1, -7]])
    B = tcon([BL, BR], [[-1, 1, 2, 3, -4, 4, -6], [-2, 1, 4, -3, 2, 3, -5]])
    rov = tcon([U, B], [[1, 2, 3, 4, -1, -3], [1, 2, 3, 4, -2, -4]])
    d = Au.shape[4]
    return reshape_f(rov, d * d, d * d)


def get_dms(ts):
    """Ground-state branch of src/graph/evaluation.jl:124-156."""
    C1, C2, C3, C4 = ts.Cs
    E1, E2, E3, E4 = ts.Es
    A, Ad = ts.A, ts.Ad
    roh = get_dm_h(C1, C2, C3, C4, E1, E1, E2, E3, E3, E4, A, A, Ad, Ad)
    rov = get_dm_v(C1, C2, C3, C4, E1, E2, E2, E3, E4, E4, A, A, Ad, Ad)
    return roh, rov


def get_gs_energy(ts, H):
    """E = tr(H1 roh)/tr(roh) + tr(H2 rov)/tr(rov) (src/graph/evaluation.jl:15-29).

    Returns (E, Na) as complex scalars.
    """
    roh, rov = get_dms(ts)
    Nh = torch.trace(roh)
    Nv = torch.trace(rov)
    Na = (Nh + Nv) / 2
    roh = roh / Nh
    rov = rov / Nv
    cdt = torch.promote_types(roh.dtype, H[0].dtype)
    H1, H2 = H[0].to(cdt), H[1].to(cdt)
    roh, rov = roh.to(cdt), rov.to(cdt)
    Eh = torch.trace(H1 @ roh)
    Ev = torch.trace(H2 @ rov)
    return Eh + Ev, Na


def get_single_dm(C1, C2, C3, C4, E1, E2, E3, E4):
    """src/graph/evaluation.jl:110-122."""
    C1E1 = tcon([C1, E1], [[-1, 1], [1, -2, -3, -4]])
    CEC2 = tcon([C1E1, C2], [[-1, 1, -3, -4], [1, -2]])
    CECE4 = tcon([CEC2, E4], [[1, -2, -3, -5], [1, -1, -4, -6]])
    C4E3 = tcon([C4, E3], [[-1, 1], [1, -2, -3, -4]])
    CEC3 = tcon([C4E3, C3], [[-1, 1, -3, -4], [-2, 1]])
    CECE2 = tcon([CEC3, E2], [[-1, 1, -3, -5], [-2, 1, -4, -6]])
    return tcon([CECE4, CECE2], [[1, 2, -1, -2, -5, -6], [1, 2, -3, -4, -7, -8]])


def get_gs_norm(ts):
    """src/graph/evaluation.jl:47-59."""
    n_dm = get_single_dm(*ts.Cs, *ts.Es)
    ndm_Ad = tcon([n_dm, ts.Ad], [[-1, -2, -3, -4, 1, 2, 3, 4], [1, 2, 3, 4, -5]])
    return (ndm_Ad * ts.A).sum()


# --------------------------------------------------------------------------
# drivers
# --------------------------------------------------------------------------

DEFAULT_CFG = dict(chi=32, min_iter=4, max_iter=20, ad_max_iter=4, rg_tol=1e-6)


def compute_gs_energy(A, H, cfg):
    """src/graph/evaluation.jl:1-13 (forward only)."""
    ts = make_ctm(A, cfg)
    ts, _, n_iter, _ = run_ctm(ts, conv_fun=lambda t: get_gs_energy(t, H)[0])
    E, N = get_gs_energy(ts, H)
    return E, N, ts, n_iter


def get_conv_boundary(x, H, cfg):
    """src/optimize/optim_gs.jl:120-131."""
    ts = make_ctm(x, cfg)
    with torch.no_grad():
        ts, _, n_iter, _ = run_ctm(ts, conv_fun=lambda t: get_gs_energy(t, H)[0])
    return ts, n_iter


def run_gs(ts, H, A):
    """src/optimize/optim_gs.jl:133-143: restart from ts with A, conj(A)."""
    ts1 = ts.replace(A=A, Ad=A.conj())
    ts1, _, _, _ = run_ctm(ts1, conv_fun=lambda t: get_gs_energy(t, H)[0])
    return get_gs_energy(ts1, H)[0], ts1


def gs_energy_grad(A0, H, cfg):
    """Energy and gradient as fg! computes them (src/optimize/optim_gs.jl:53-91).

    x = renormalize(A0) outside the AD; the converged environment is held
    constant; ``ad_max_iter`` steps are differentiated.  Returns (E, g, ts)
    with g = dRe(E)/dRe(x) + i dRe(E)/dIm(x) for complex x.
    """
    x = renormalize(A0.detach())
    ts, _ = get_conv_boundary(x, H, cfg)
    ts.params = dict(ts.params)
    ts.params["max_iter"] = cfg.get("ad_max_iter", 4)
    xg = x.clone().requires_grad_(True)
    E, ts1 = run_gs(ts, H, xg)
    E.real.backward()
    return E.detach(), xg.grad.detach(), ts


# --------------------------------------------------------------------------
# models (src/model/basis.jl, honeycomb.jl, lee_gs.jl)
# --------------------------------------------------------------------------

C128 = torch.complex128
F64 = torch.float64
sigmax = torch.tensor([[0, 1], [1, 0]], dtype=F64)
sigmay = torch.tensor([[0, -1j], [1j, 0]], dtype=C128)
sigmaz = torch.tensor([[1, 0], [0, -1]], dtype=F64)
sI = torch.eye(2, dtype=F64)
Sx, Sy, Sz, SI = sigmax / 2, sigmay / 2, sigmaz / 2, sI


def tout(a, b):
    """src/model/basis.jl:16-21 (site-1 index fastest)."""
    if a.dtype != b.dtype:
        a, b = a.to(C128), b.to(C128)
    c = tcon([a, b], [[-1, -3], [-2, -4]])
    s = c.shape
    return reshape_f(c, s[0] * s[1], s[2] * s[3])


def ising(h=2.0):
    """src/model/basis.jl:23-29."""
    H = -tout(sigmax, sigmax) - h * tout(sigmaz, sI) / 4 - h * tout(sI, sigmaz) / 4
    return [H, H]


def heisenberg(Jz=1.0):
    """src/model/basis.jl:31-35 (ComplexF64-typed, real-valued)."""
    H = Jz * tout(Sz, Sz).to(C128) - tout(Sx, Sx).to(C128) - tout(Sy, Sy)
    return [H, H]


def honeycomb(Jx=1.0, Jy=1.0, Jz=1.0, dir="ZZ"):
    """src/model/honeycomb.jl:1-14."""
    def t4(a, b, c, d):
        return tout(tout(a, b), tout(c, d))
    if dir == "ZZ":
        hv = Jx * t4(sI, sigmax, sigmax, sI) + (t4(sigmaz, sigmaz, sI, sI) + t4(sI, sI, sigmaz, sigmaz)) / 4
        hh = Jy * t4(sI, sigmay, sigmay, sI) + (t4(sigmaz, sigmaz, sI, sI) + t4(sI, sI, sigmaz, sigmaz)) / 4
    elif dir == "XX":
        hv = Jy * t4(sI, sigmay, sigmay, sI) + (t4(sigmax, sigmax, sI, sI) + t4(sI, sI, sigmax, sigmax)) / 4
        hh = Jz * t4(sI, sigmaz, sigmaz, sI) + (t4(sigmax, sigmax, sI, sI) + t4(sI, sI, sigmax, sigmax)) / 4
    elif dir == "YY":
        hv = Jz * t4(sI, sigmaz, sigmaz, sI) + (t4(sigmay, sigmay, sI, sI) + t4(sI, sI, sigmay, sigmay)) / 4
        hh = Jx * t4(sI, sigmax, sigmax, sI) + (t4(sigmay, sigmay, sI, sI) + t4(sI, sI, sigmay, sigmay)) / 4
    else:
        raise ValueError(dir)
    hv = hv.real.to(F64) if hv.dtype.is_complex else hv
    hh = hh.real.to(F64) if hh.dtype.is_complex else hh
    return [-hh, -hv]


def get_Q_op():
    """src/model/lee_gs.jl:88-96."""
    Q = torch.zeros(2, 2, 2, 2, 2, dtype=C128)
    Q[0, 0, 0] = sI.to(C128)
    Q[0, 1, 1] = sigmax.to(C128)
    Q[1, 0, 1] = sigmay
    Q[1, 1, 0] = sigmaz.to(C128)
    return Q


def init_hb_gs(D=2, dir="XX"):
    """Loop-gas state (src/model/lee_gs.jl:2-28), D=2 branch."""
    if D != 2:
        raise NotImplementedError("only the D=2 loop-gas state is restated")
    Q = get_Q_op()
    u = 1 / math.sqrt(3)
    s111 = torch.tensor([1 + u, u + 1j * u], dtype=C128) / math.sqrt(2 + 2 * u)
    T = torch.einsum("abcdp,p->abcd", Q, s111)
    if dir == "ZZ":
        A = tcon([T, T], [[-1, -2, 1, -5], [-3, -4, 1, -6]])
    elif dir == "XX":
        A = tcon([T, T], [[1, -1, -2, -5], [1, -3, -4, -6]])
    elif dir == "YY":
        A = tcon([T, T], [[-2, 1, -1, -5], [-4, 1, -3, -6]])
    else:
        raise ValueError(dir)
    return reshape_f(A, 2, 2, 2, 2, 4)


def seeded_tensor(shape, seed, dtype=F64, unit_max=False):
    """Synthetic A: numpy default_rng(seed).standard_normal, Fortran order."""
    n = int(np.prod(shape))
    v = np.random.default_rng(seed).standard_normal(n)
    a = v.reshape(shape, order="F")
    if unit_max:
        a = a / np.abs(a).max()
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype)
