"""CPU FP64 restatement of the reference's transfer-matrix spectrum and spin
correlations -- TEST INFRASTRUCTURE.

Imported only by ``tests/``; the product path never calls it.  It restates
src/graph/expectation.jl of yitan1/iPEPS.jl in the reference's own index
lists (each ``@ein`` written as one einsum):

* ``get_envs_TM_h`` / ``get_envs_TM_v``   expectation.jl:299-361
* ``con_LT`` / ``con_TR`` / ``con_LR``     expectation.jl:273-297
* ``correlation_spin``                     expectation.jl:246-271
* ``correlation_TM``                       expectation.jl:193-220 (``full``:
  dense eigenvalues; otherwise the N largest-magnitude eigenvalues, which
  the reference takes from KrylovKit's ``eigsolve(TM, N, :LM)``)

Tensors are torch CPU tensors indexed like the Julia arrays.

Pinning (tests/test_oracle_corr.py): ``correlation_spin`` against an
independent dense formulation (explicit column transfer matrices and matrix
powers), the D = 1 product-state closed form, and the spectrum of
``correlation_TM`` against its own dense matrix and the matrix-free apply.
"""
import numpy as np
import torch

from ctmrg_oracle import reshape_f, tcon


def _ein(spec, *xs):
    return torch.einsum(spec, *xs)


def double_layer(A, Ad, op=None):
    """T[m1..m8] = A[m1,m3,m5,m7,p1] op[p1,p2] Ad[m2,m4,m6,m8,p2]
    (expectation.jl:310-312, 340-342); op = None is the identity."""
    if op is None:
        return _ein("acegp,bdfhp->abcdefgh", A, Ad)
    return _ein("acegp,pq,bdfhq->abcdefgh", A, op, Ad)


def get_envs_TM_h(ts, op1, op2):
    """expectation.jl:299-327: (Et, Ed, T, T_op1, T_op2, L0, R0)."""
    C1, C2, C3, C4 = ts.Cs
    E1, E2, E3, E4 = ts.Es
    L0 = _ein("pa,pqbc,qd->abcd", C1, E4, C4)
    R0 = _ein("ap,pqbc,qd->abcd", C2, E2, C3)
    T = double_layer(ts.A, ts.Ad)
    return E1, E3, T, double_layer(ts.A, ts.Ad, op1), double_layer(ts.A, ts.Ad, op2), L0, R0


def get_envs_TM_v(ts, op1, op2):
    """expectation.jl:329-361: (El, Er, T, T_op1, T_op2, U0, D0)."""
    C1, C2, C3, C4 = ts.Cs
    E1, E2, E3, E4 = ts.Es
    U0 = _ein("ap,pqbc,qd->abcd", C1, E1, C2)
    D0 = _ein("ap,pqbc,dq->abcd", C4, E3, C3)
    T = double_layer(ts.A, ts.Ad)
    return E4, E2, T, double_layer(ts.A, ts.Ad, op1), double_layer(ts.A, ts.Ad, op2), U0, D0


def con_LT(Et, Ed, L0, T, direction="h"):
    """expectation.jl:273-281."""
    # L[p1,p2,p3,p4] := L0[m1,m4,m5,m8] Et[m1,p1,m2,m3] T[...] Ed[m8,p4,m6,m7]
    if direction == "h":
        return _ein("adeh,awbc,bcdefgxy,hzfg->wxyz", L0, Et, T, Ed)
    return _ein("adeh,awbc,debcxyfg,hzfg->wxyz", L0, Et, T, Ed)


def con_TR(Et, Ed, T, R0, direction="h"):
    """expectation.jl:283-291."""
    # R[p1,p2,p3,p4] := R0[m1,m4,m5,m8] Et[p1,m1,m2,m3] T[...] Ed[p4,m8,m6,m7]
    if direction == "h":
        return _ein("adeh,wabc,bcxyfgde,zhfg->wxyz", R0, Et, T, Ed)
    return _ein("adeh,wabc,xybcdefg,zhfg->wxyz", R0, Et, T, Ed)


def con_LR(L, R):
    """expectation.jl:293-297."""
    return _ein("abcd,abcd->", L, R)


def _promoted(ts, op1, op2):
    """Environment and operators in one dtype (Julia promotes real tensors
    when an operator is complex)."""
    dt = ts.A.dtype
    for x in (op1, op2, ts.Cs[0]):
        dt = torch.promote_types(dt, x.dtype)
    cv = lambda x: x.to(dt)
    return ts.replace(A=cv(ts.A), Ad=cv(ts.Ad), Cs=[cv(c) for c in ts.Cs], Es=[cv(e) for e in ts.Es]), cv(op1), cv(op2)


def correlation_spin(ts, op1, op2, max_iter=10, direction="h"):
    """expectation.jl:246-271: (s2s list, m1, m2)."""
    ts, op1, op2 = _promoted(ts, op1, op2)
    if direction == "h":
        Et, Ed, T, T1, T2, L0, R0 = get_envs_TM_h(ts, op1, op2)
    else:
        Et, Ed, T, T1, T2, L0, R0 = get_envs_TM_v(ts, op1, op2)
    L = con_LT(Et, Ed, L0, T1, direction)
    R = con_TR(Et, Ed, T2, R0, direction)
    s2s = [con_LR(L, R)]
    m1 = con_LR(L, R0)
    m2 = con_LR(L0, R)
    for _ in range(max_iter - 1):
        L = con_LT(Et, Ed, L, T, direction)
        s2s.append(con_LR(L, R))
    return s2s, m1, m2


def transfer_matrix(ts, direction="h"):
    """TM of expectation.jl:199-211 reshaped to (chi^2, chi^2) column-major."""
    E1, E2, E3, E4 = ts.Es
    if direction == "h":
        TM = tcon([E4, E2], [[-1, -3, 1, 2], [-2, -4, 1, 2]])
    else:
        TM = tcon([E1, E3], [[-1, -3, 1, 2], [-2, -4, 1, 2]])
    s = TM.shape
    return reshape_f(TM, s[0] * s[1], s[2] * s[3])


def correlation_TM(ts, direction="h", N=10, full=False):
    """expectation.jl:193-220: (|es| descending and > 0, es, TM)."""
    TM = transfer_matrix(ts, direction)
    es = np.linalg.eigvals(TM.numpy())
    if not full:
        es = es[np.argsort(-np.abs(es))][:N]
    es1 = np.sort(np.abs(es))[::-1]
    return es1[es1 > 0], es, TM
