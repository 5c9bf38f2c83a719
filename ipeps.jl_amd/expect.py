"""Transfer-matrix spectrum and spin correlations over the HIP engine.

Mirrors src/graph/expectation.jl:

* ``correlation_spin(ts, op1, op2, max_iter, direction)``  expectation.jl:246-271
  (get_envs_TM_h / get_envs_TM_v, con_LT, con_TR, con_LR: :273-361), all
  contractions on the GPU
* ``correlation_TM(ts, direction, N, full)``                expectation.jl:193-220:
  the transfer matrix of two edge tensors; ``full`` brings the dense matrix
  to the host for LAPACK's eigenvalues, otherwise the N largest-magnitude
  eigenvalues come from ARPACK (implicitly restarted Arnoldi, the role of
  KrylovKit's ``eigsolve(TM, N, :LM)`` in the reference) driving the
  matrix-free product on the GPU.
"""
import ctypes as C

import numpy as np
import scipy.sparse.linalg as spla

from ._lib import C128, check, lib
from .ctm import _dp, _ptr

_DIR = {"h": 0, "v": 1}


def _op(o, d):
    a = np.asfortranarray(np.asarray(o, dtype=np.complex128))
    if a.shape != (d, d):
        raise ValueError(f"operator must be {d} x {d}")
    return a


def correlation_spin(ts, op1, op2=None, max_iter=10, direction="h"):
    """(s2s, m1, m2): s2s[r] = <L_r, R>, r = 0..max_iter-1, and the
    one-operator contractions m1 = <L, R0>, m2 = <L0, R> (unnormalised, as
    the reference returns them)."""
    d = ts.A.shape[4]
    o1 = _op(op1, d)
    o2 = _op(op1 if op2 is None else op2, d)
    s = np.zeros(2 * int(max_iter))
    m1, m2 = np.zeros(2), np.zeros(2)
    check(lib().ipeps_correlation_spin(ts.h, _dp(o1), _dp(o2), d, int(max_iter),
                                       _DIR[direction], _dp(s), _dp(m1), _dp(m2)))
    return list(s[0::2] + 1j * s[1::2]), complex(m1[0], m1[1]), complex(m2[0], m2[1])


def transfer_matrix_dims(ts, direction="h"):
    r, c = C.c_int64(), C.c_int64()
    check(lib().ipeps_tm_dims(ts.h, _DIR[direction], C.byref(r), C.byref(c)))
    return r.value, c.value


def transfer_matrix(ts, direction="h"):
    """TM (rows x cols), dense, on the host."""
    rows, cols = transfer_matrix_dims(ts, direction)
    out = np.empty((rows, cols), dtype=ts._np, order="F")
    check(lib().ipeps_tm_dense(ts.h, _DIR[direction], _ptr(out)))
    return out


def transfer_operator(ts, direction="h"):
    """TM as a scipy LinearOperator whose products run on the GPU."""
    rows, cols = transfer_matrix_dims(ts, direction)
    dt = ts._np

    def mm(X):
        X = np.asarray(X)
        vec = X.ndim == 1
        Xf = np.asfortranarray(X.reshape(cols, -1), dtype=dt)
        if not np.iscomplexobj(X) or ts.code == C128:
            Y = np.empty((rows, Xf.shape[1]), dtype=dt, order="F")
            check(lib().ipeps_tm_apply(ts.h, _DIR[direction], Xf.shape[1], _ptr(Xf), _ptr(Y)))
        else:  # complex vectors on a real operator: real and imaginary parts
            Xr = np.asfortranarray(X.reshape(cols, -1).real)
            Xi = np.asfortranarray(X.reshape(cols, -1).imag)
            Yr = np.empty((rows, Xr.shape[1]), order="F")
            Yi = np.empty_like(Yr)
            check(lib().ipeps_tm_apply(ts.h, _DIR[direction], Xr.shape[1], _ptr(Xr), _ptr(Yr)))
            check(lib().ipeps_tm_apply(ts.h, _DIR[direction], Xi.shape[1], _ptr(Xi), _ptr(Yi)))
            Y = Yr + 1j * Yi
        return Y.reshape(-1) if vec else Y

    return spla.LinearOperator((rows, cols), matvec=mm, matmat=mm, dtype=dt)


def correlation_TM(ts, direction="h", N=10, full=False):
    """(es1, es, TM): es1 = |eigenvalues| in descending order (positive ones),
    es the eigenvalues; TM is the dense matrix when ``full``, else the GPU
    LinearOperator."""
    if full:
        TM = transfer_matrix(ts, direction)
        es = np.linalg.eigvals(TM)
    else:
        TM = transfer_operator(ts, direction)
        es = spla.eigs(TM, k=int(N), which="LM", return_eigenvectors=False)
    es1 = np.sort(np.abs(es))[::-1]
    return es1[es1 > 0], es, TM


def get_gs_energy_4x4(ts, sl):
    """get_gs_energy_4x4 (src/graph/evaluation.jl:340-355) on the GPU for the
    factors sl = get_local_h(h) of a plaquette Hamiltonian: (energy / norm,
    norm) as complex numbers."""
    import numpy as np
    fs = [np.asfortranarray(np.asarray(x), dtype=np.complex128) for x in sl]
    dims = np.ones((4, 4), dtype=np.int64)
    for i, f in enumerate(fs):
        dims[i, : f.ndim] = f.shape
    E, N = np.zeros(2), np.zeros(2)
    dflat = np.ascontiguousarray(dims)
    check(lib().ipeps_gs_energy_4x4(ts.h, dflat.ctypes.data_as(C.POINTER(C.c_int64)), *[_ptr(f) for f in fs],
                                    E.ctypes.data_as(C.POINTER(C.c_double)), N.ctypes.data_as(C.POINTER(C.c_double))))
    return complex(E[0], E[1]), complex(N[0], N[1])
