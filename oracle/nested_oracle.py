"""Nested-tensor (excited-state) CTMRG restatement -- TEST INFRASTRUCTURE.

CPU FP64/complex128 restatement of the reference's excited-state path, used
only as the checker of the HIP engine (imported by tests/ only):

* ``Nested`` = [X, X_B, X_Bd, X_BB] with ``None`` for EmptyT
                         src/nested_tensor/nested_tensor.jl:1-66, emptyT.jl:1-29
* nested ``tcon`` product rule            nested_tensor.jl:48-63
* ``renormalize`` (scaled by max|X_1|)     nested_tensor.jl:41-46
* ``shift`` (momentum phase)               nested_tensor.jl:65
* nested environment accessors            src/ctm/ctm_tensor.jl:20-55
* ``proj_*`` / ``up_*`` on nested tensors src/ctm/ctmrg.jl:267-491
* ``get_dms(only_gs=false)``, ``get_es_energy``, ``get_all_norm``
                                          src/graph/evaluation.jl:31-77,124-156
* ``get_es_grad`` / ``run_es`` and the effH/effN columns
                                          src/optimize/optim_es.jl:127-194
* ``substract_gs_energy``                 src/optimize/optim_es.jl:224-234

Projectors come from the plain ground-state parts (ctmrg.jl:122-252), as in
the reference.  Pinned by the reference's committed es_*.jld2 (effH, effN,
envB columns), see tests/test_oracle_nested.py.
"""
import cmath

import torch

import ctmrg_oracle as O


class Nested(list):
    """[X, X_B, X_Bd, X_BB]; None stands for EmptyT."""


def isN(x):
    return isinstance(x, Nested)


def _add(a, b):
    if a is None:
        return b
    if b is None:
        return a
    return a + b


def _tc(a, la, b, lb):
    if a is None or b is None:
        return None
    if a.dtype != b.dtype:
        a, b = a.to(torch.complex128), b.to(torch.complex128)
    return O.tcon([a, b], [la, lb])


def ntcon(xs, inds):
    """tcon with the nested product rule (nested_tensor.jl:48-63)."""
    A, B = xs
    la, lb = inds
    if isN(A) and isN(B):
        r1 = _tc(A[0], la, B[0], lb)
        r2 = _add(_tc(A[0], la, B[1], lb), _tc(A[1], la, B[0], lb))
        r3 = _add(_tc(A[0], la, B[2], lb), _tc(A[2], la, B[0], lb))
        r4 = _add(_add(_tc(A[0], la, B[3], lb), _tc(A[3], la, B[0], lb)),
                  _add(_tc(A[1], la, B[2], lb), _tc(A[2], la, B[1], lb)))
        return Nested([r1, r2, r3, r4])
    if isN(A):
        return Nested([_tc(a, la, B, lb) for a in A])
    if isN(B):
        return Nested([_tc(A, la, b, lb) for b in B])
    return O.tcon([A, B], [la, lb])


def nmap(f, x):
    if isN(x):
        return Nested([None if c is None else f(c) for c in x])
    return f(x)


def rs(x, *shape):
    return nmap(lambda t: O.reshape_f(t, *shape), x)


def tp(x):
    return nmap(lambda t: t.transpose(0, 1), x)


def mm(a, b):
    """a * b for matrices, one side possibly nested."""
    if isN(a):
        return Nested([None if c is None else c @ b for c in a])
    if isN(b):
        return Nested([None if c is None else a @ c for c in b])
    return a @ b


def nrenormalize(x):
    """nested_tensor.jl:41-46: all components divided by max|X_1|."""
    if not isN(x):
        return O.renormalize(x)
    a = O.flat_f(x[0]).abs()
    idx = int(torch.argmax(a.detach()))
    f = a[idx]
    return Nested([None if c is None else c / f for c in x])


def shift(x, phi):
    """nested_tensor.jl:65: component 2 times e^{i phi}, component 3 times e^{-i phi}."""
    if not isN(x):
        return x
    e = cmath.exp(1j * phi)
    return Nested([x[0], None if x[1] is None else x[1] * e,
                   None if x[2] is None else x[2] * e.conjugate(), x[3]])


class NestedCTM:
    """CTMTensors with B (ctm_tensor.jl:3-18): plain parts + B components."""

    def __init__(self, A, Ad, Cs, Es, B=None, Bd=None, BC=None, BdC=None, BBC=None, BE=None, BdE=None, BBE=None,
                 params=None):
        self.A, self.Ad, self.Cs, self.Es = A, Ad, list(Cs), list(Es)
        self.B, self.Bd = B, Bd
        self.BC = list(BC or [None] * 4)
        self.BdC = list(BdC or [None] * 4)
        self.BBC = list(BBC or [None] * 4)
        self.BE = list(BE or [None] * 4)
        self.BdE = list(BdE or [None] * 4)
        self.BBE = list(BBE or [None] * 4)
        self.params = dict(params or {})

    def copy(self, **kw):
        t = NestedCTM(self.A, self.Ad, self.Cs, self.Es, self.B, self.Bd, self.BC, self.BdC, self.BBC, self.BE,
                      self.BdE, self.BBE, self.params)
        for k, v in kw.items():
            setattr(t, k, v)
        return t

    # ctm_tensor.jl:20-55
    def all_A(self):
        return Nested([self.A, self.B, None, None])

    def all_Ad(self):
        return Nested([self.Ad, None, self.Bd, None])

    def all_Cs(self):
        return [Nested([self.Cs[i], self.BC[i], self.BdC[i], self.BBC[i]]) for i in range(4)]

    def all_Es(self):
        return [Nested([self.Es[i], self.BE[i], self.BdE[i], self.BBE[i]]) for i in range(4)]

    def plain(self):
        return O.CTMTensors(self.A, self.Ad, self.Cs, self.Es, self.params)

    def store(self, which_C, which_E, Cn, En):
        """Write nested results back (up_* of ctmrg.jl:377-491)."""
        Cs, BC, BdC, BBC = list(self.Cs), list(self.BC), list(self.BdC), list(self.BBC)
        Es, BE, BdE, BBE = list(self.Es), list(self.BE), list(self.BdE), list(self.BBE)
        for i, X in zip(which_C, Cn):
            Cs[i], BC[i], BdC[i], BBC[i] = X
        for i, X in zip(which_E, En):
            Es[i], BE[i], BdE[i], BBE[i] = X
        return self.copy(Cs=Cs, BC=BC, BdC=BdC, BBC=BBC, Es=Es, BE=BE, BdE=BdE, BBE=BBE)


def _edge(E, A, Ad, ia, iad, P, Pd):
    EA = ntcon([E, A], [[-1, -3, 1, -6], ia])
    EAAd = ntcon([EA, Ad], [[-1, -2, -4, -5, -7, 1, 2], iad])
    s = EAAd[0].shape
    EAAd = rs(EAAd, s[0] * s[1] * s[2], s[3] * s[4] * s[5], s[6], s[7])
    EPd = ntcon([Pd, EAAd], [[-1, 1], [1, -2, -3, -4]])
    return ntcon([EPd, P], [[-1, 1, -3, -4], [1, -2]])


def _rows(x):
    return x[0].shape[0] if isN(x) else x.shape[0]


def _last(x):
    return x[0].shape[3] if isN(x) else x.shape[3]


def left_rg(ts, chi):
    P, Pd, s = O.get_projector_left(ts.plain(), chi)
    A, Ad = ts.all_A(), ts.all_Ad()
    C1, _, _, C4 = ts.all_Cs()
    E1, _, E3, E4 = ts.all_Es()
    C1E1 = ntcon([C1, E1], [[-2, 1], [1, -1, -3, -4]])
    nC1 = tp(mm(rs(C1E1, _rows(C1E1), -1), P))
    C4E3 = ntcon([C4, E3], [[-1, 1], [1, -4, -2, -3]])
    nC4 = mm(Pd, rs(C4E3, -1, _last(C4E3)))
    nE4 = _edge(E4, A, Ad, [-2, 1, -4, -5, -7], [-3, 1, -6, -8, 2], P, Pd)
    px = ts.params.get("px", 0.0)
    nC1, nE4, nC4 = (shift(nrenormalize(x), -px) for x in (nC1, nE4, nC4))
    return ts.store([0, 3], [3], [nC1, nC4], [nE4]), s


def right_rg(ts, chi):
    P, Pd, s = O.get_projector_right(ts.plain(), chi)
    A, Ad = ts.all_A(), ts.all_Ad()
    _, C2, C3, _ = ts.all_Cs()
    E1, E2, E3, _ = ts.all_Es()
    C2E1 = ntcon([C2, E1], [[1, -2], [-1, 1, -3, -4]])
    nC2 = mm(rs(C2E1, _rows(C2E1), -1), P)
    C3E3 = ntcon([C3, E3], [[-1, 1], [-4, 1, -2, -3]])
    nC3 = mm(Pd, rs(C3E3, -1, _last(C3E3)))
    nE2 = _edge(E2, A, Ad, [-2, -5, -4, 1, -7], [-3, -8, -6, 1, 2], P, Pd)
    px = ts.params.get("px", 0.0)
    nC2, nE2, nC3 = (shift(nrenormalize(x), px) for x in (nC2, nE2, nC3))
    return ts.store([1, 2], [1], [nC2, nC3], [nE2]), s


def top_rg(ts, chi):
    P, Pd, s = O.get_projector_top(ts.plain(), chi)
    A, Ad = ts.all_A(), ts.all_Ad()
    C1, C2, _, _ = ts.all_Cs()
    E1, E2, _, E4 = ts.all_Es()
    C1E4 = ntcon([C1, E4], [[1, -2], [1, -1, -3, -4]])
    nC1 = mm(rs(C1E4, _rows(C1E4), -1), P)
    C2E2 = ntcon([C2, E2], [[-1, 1], [1, -4, -2, -3]])
    nC2 = mm(Pd, rs(C2E2, -1, _last(C2E2)))
    nE1 = _edge(E1, A, Ad, [1, -2, -5, -4, -7], [1, -3, -8, -6, 2], P, Pd)
    py = ts.params.get("py", 0.0)
    nC1, nE1, nC2 = (shift(nrenormalize(x), -py) for x in (nC1, nE1, nC2))
    return ts.store([0, 1], [0], [nC1, nC2], [nE1]), s


def bottom_rg(ts, chi):
    P, Pd, s = O.get_projector_bottom(ts.plain(), chi)
    A, Ad = ts.all_A(), ts.all_Ad()
    _, _, C3, C4 = ts.all_Cs()
    _, E2, E3, E4 = ts.all_Es()
    C4E4 = ntcon([C4, E4], [[1, -2], [-1, 1, -3, -4]])
    nC4 = mm(rs(C4E4, _rows(C4E4), -1), P)
    C3E2 = ntcon([C3, E2], [[1, -1], [-4, 1, -2, -3]])
    nC3 = tp(mm(Pd, rs(C3E2, -1, _last(C3E2))))
    nE3 = _edge(E3, A, Ad, [-5, -2, 1, -4, -7], [-8, -3, 1, -6, 2], P, Pd)
    py = ts.params.get("py", 0.0)
    nC4, nE3, nC3 = (shift(nrenormalize(x), py) for x in (nC4, nE3, nC3))
    return ts.store([3, 2], [2], [nC4, nC3], [nE3]), s


def rg_step(ts, chi):
    ts, s = left_rg(ts, chi)
    ts, _ = right_rg(ts, chi)
    ts, _ = top_rg(ts, chi)
    ts, _ = bottom_rg(ts, chi)
    return ts, s


# ---- observables (evaluation.jl) -------------------------------------------

def _dm_h(C1, C2, C3, C4, E1l, E1r, E2, E3l, E3r, E4, Al, Ar, Adl, Adr):
    C1E1 = ntcon([C1, E1l], [[-1, 1], [1, -2, -3, -4]])
    LU = ntcon([C1E1, Adl], [[-1, -2, -3, 1], [1, -4, -5, -6, -7]])
    C4E3 = ntcon([C4, E3l], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = ntcon([E4, C4E3], [[-1, 1, -3, -5], [1, -2, -4, -6]])
    LB = ntcon([CEE4, Al], [[-1, -2, 1, 2, -4, -5], [-3, 1, 2, -6, -7]])
    L = ntcon([LB, LU], [[1, -2, 2, 3, 4, -3, -5], [1, -1, 2, 3, 4, -4, -6]])
    C2E1 = ntcon([C2, E1r], [[1, -2], [-1, 1, -3, -4]])
    RU = ntcon([C2E1, Ar], [[-1, -2, 1, -3], [1, -4, -5, -6, -7]])
    C3E3 = ntcon([C3, E3r], [[-1, 1], [-2, 1, -3, -4]])
    CEE2 = ntcon([C3E3, E2], [[1, -2, -3, -5], [-1, 1, -4, -6]])
    RB = ntcon([CEE2, Adr], [[-1, -2, -3, -4, 1, 2], [-5, -6, 1, 2, -7]])
    R = ntcon([RU, RB], [[-1, 1, 2, -3, 3, 4, -5], [1, -2, 3, 4, 2, -4, -6]])
    roh = ntcon([L, R], [[1, 2, 3, 4, -1, -3], [1, 2, 3, 4, -2, -4]])
    d = (Al[0] if isN(Al) else Al).shape[4]
    return rs(roh, d * d, d * d)


def _dm_v(C1, C2, C3, C4, E1, E2u, E2d, E3, E4u, E4d, Au, Ad_, Adu, Add):
    C1E1 = ntcon([C1, E1], [[-1, 1], [1, -2, -3, -4]])
    CEE4 = ntcon([C1E1, E4u], [[1, -2, -3, -5], [1, -1, -4, -6]])
    UL = ntcon([CEE4, Au], [[-1, -2, 1, 2, -3, -4], [1, 2, -5, -6, -7]])
    C2E2 = ntcon([C2, E2u], [[-1, 1], [1, -2, -3, -4]])
    UR = ntcon([C2E2, Adu], [[-1, -2, -3, 1], [-4, -5, -6, 1, -7]])
    U = ntcon([UL, UR], [[-1, 1, 3, 4, -3, 2, -5], [1, -2, 2, 3, 4, -4, -6]])
    C4E3 = ntcon([C4, E3], [[-1, 1], [1, -2, -3, -4]])
    CEE4b = ntcon([C4E3, E4d], [[1, -2, -4, -6], [-1, 1, -3, -5]])
    BL = ntcon([CEE4b, Add], [[-1, -2, -3, -4, 1, 2], [-5, 1, 2, -6, -7]])
    C3E2 = ntcon([C3, E2d], [[1, -2], [-1, 1, -3, -4]])
    BR = ntcon([C3E2, Ad_], [[-1, -2, 1, -3], [-4, -5, -6, 1, -7]])
    Bm = ntcon([BL, BR], [[-1, 1, 2, 3, -4, 4, -6], [-2, 1, 4, -3, 2, 3, -5]])
    rov = ntcon([U, Bm], [[1, 2, 3, 4, -1, -3], [1, 2, 3, 4, -2, -4]])
    d = (Au[0] if isN(Au) else Au).shape[4]
    return rs(rov, d * d, d * d)


def get_dms_nested(ts):
    """get_dms(ts; only_gs=false) (evaluation.jl:138-150) with the phase shifts."""
    A, Ad = ts.all_A(), ts.all_Ad()
    C1, C2, C3, C4 = ts.all_Cs()
    E1, E2, E3, E4 = ts.all_Es()
    px = ts.params.get("px", 0.0)
    py = ts.params.get("py", 0.0)
    roh = _dm_h(C1, shift(C2, px), shift(C3, px), C4, E1, shift(E1, px), shift(E2, px), E3, shift(E3, px), E4, A,
                shift(A, px), Ad, shift(Ad, px))
    rov = _dm_v(C1, C2, shift(C3, py), shift(C4, py), E1, E2, shift(E2, py), shift(E3, py), E4, shift(E4, py), A,
                shift(A, py), Ad, shift(Ad, py))
    return roh, rov


def get_es_energy(ts, H):
    """evaluation.jl:31-45: Re(E[4]) after normalising by the component-1 trace."""
    roh, rov = get_dms_nested(ts)
    Nh = torch.trace(roh[0]).real
    Nv = torch.trace(rov[0]).real
    H1 = H[0].to(torch.complex128)
    H2 = H[1].to(torch.complex128)
    comps = []
    for i in range(4):
        a = roh[i]
        b = rov[i]
        eh = None if a is None else torch.trace(H1 @ (a.to(torch.complex128) / Nh))
        ev = None if b is None else torch.trace(H2 @ (b.to(torch.complex128) / Nv))
        comps.append(_add(eh, ev))
    return comps[3].real


def get_single_dm(Cs, Es):
    C1, C2, C3, C4 = Cs
    E1, E2, E3, E4 = Es
    C1E1 = ntcon([C1, E1], [[-1, 1], [1, -2, -3, -4]])
    CEC2 = ntcon([C1E1, C2], [[-1, 1, -3, -4], [1, -2]])
    CECE4 = ntcon([CEC2, E4], [[1, -2, -3, -5], [1, -1, -4, -6]])
    C4E3 = ntcon([C4, E3], [[-1, 1], [1, -2, -3, -4]])
    CEC3 = ntcon([C4E3, C3], [[-1, 1, -3, -4], [-2, 1]])
    CECE2 = ntcon([CEC3, E2], [[-1, 1, -3, -5], [-2, 1, -4, -6]])
    return ntcon([CECE4, CECE2], [[1, 2, -1, -2, -5, -6], [1, 2, -3, -4, -7, -8]])


def get_all_norm(ts):
    """evaluation.jl:61-77: (Nb, envB)."""
    n_dm = get_single_dm(ts.all_Cs(), ts.all_Es())
    ndm_A = ntcon([n_dm, ts.all_A()], [[1, 2, 3, 4, -1, -2, -3, -4], [1, 2, 3, 4, -5]])
    all_norm = ntcon([ndm_A, ts.all_Ad()], [[1, 2, 3, 4, 5], [1, 2, 3, 4, 5]])
    envB = ndm_A[1] / all_norm[0]
    Nb = (ndm_A[1] * ts.Bd).sum() / all_norm[0]
    return Nb, envB


# ---- driver (optim_es.jl) ---------------------------------------------------

def substract_gs_energy(ts, H):
    """optim_es.jl:224-234."""
    E, _ = O.get_gs_energy(ts, H)
    e = E.real.item() / 2
    return [h - e * torch.eye(h.shape[0], dtype=h.dtype) for h in H]


def run_ctm_nested(ts, H, max_iter=None, verbose=False):
    """run_ctm with conv_fun = get_es_energy / get_all_norm[1] (optim_es.jl:175-176)."""
    p = ts.params
    chi = p.get("chi", 30)
    min_iter = p.get("min_iter", 4)
    max_iter = p.get("max_iter", 20) if max_iter is None else max_iter
    tol = p.get("rg_tol", 1e-6)
    diffs = [1.0]
    conv = torch.tensor(1.0 + 0j, dtype=torch.complex128)
    for i in range(1, max_iter + 1):
        ts, s = rg_step(ts, chi)
        old = conv
        with torch.no_grad():
            c = get_es_energy(ts, H) / get_all_norm(ts)[0]
            conv = c.detach().to(torch.complex128).reshape(())
            diffs.append(abs(complex(conv) - complex(old)))
        if verbose:
            print(f"nested CTM step {i}, diff {diffs[-1]:.3e}")
        if i >= min_iter and diffs[-1] < tol:
            break
    return ts


def get_es_grad(ts, H, Bvec):
    """optim_es.jl:167-194: returns (gradH, gradN) as flat column-major vectors."""
    B = O.reshape_f(Bvec, *ts.A.shape)
    ts1 = ts.copy(B=B, Bd=B.conj())
    with torch.no_grad():
        ts1 = run_ctm_nested(ts1, H)
    Bg = B.detach().clone().requires_grad_(True)
    ts2 = ts1.copy(B=Bg, Bd=Bg.conj())
    ts2 = run_ctm_nested(ts2, H, max_iter=ts.params.get("ad_max_iter", 4))
    E = get_es_energy(ts2, H)
    E.backward()
    gH = Bg.grad.detach()
    with torch.no_grad():
        ts3 = ts2.copy(B=B, Bd=B.conj(), BC=[None if x is None else x.detach() for x in ts2.BC],
                       BdC=[None if x is None else x.detach() for x in ts2.BdC],
                       BBC=[None if x is None else x.detach() for x in ts2.BBC],
                       BE=[None if x is None else x.detach() for x in ts2.BE],
                       BdE=[None if x is None else x.detach() for x in ts2.BdE],
                       BBE=[None if x is None else x.detach() for x in ts2.BBE],
                       Cs=[x.detach() for x in ts2.Cs], Es=[x.detach() for x in ts2.Es])
        _, envB = get_all_norm(ts3)
    return O.flat_f(gH), O.flat_f(envB)
