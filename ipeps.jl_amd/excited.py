"""Excited states (nested-tensor CTMRG) over the HIP engine.

Mirrors src/optimize/optim_es.jl and src/graph/expectation.jl:

* ``get_es_grad(ts, H, Bi, px, py, cfg)``   optim_es.jl:167-194 (GPU)
* ``optim_es(ts, H, basis, px, py, cfg)``   optim_es.jl:91-165: effH, effN,
  envB columns for every tangent basis vector (GPU per column)
* ``normalize_gs``, ``substract_gs_energy``, ``get_tangent_basis``
                                            optim_es.jl:212-258 (+ get_gauge_basis :294-344)
* ``prepare_basis(A, H, cfg)``              optim_es.jl:34-73 (no file I/O)
* ``compute_es(effH, effN, nrmB_cut)``      expectation.jl:13-56: the small
  generalised eigenproblem, host FP64 (LAPACK via scipy), as in the reference
* ``make_es_path(step)``                    optim_es.jl:1-20
* ``band_structure(ts, H, basis, momenta, cfg)`` the bands over a momentum path
                                            (optim_es + compute_es per point)
"""
import ctypes as C

import numpy as np
import scipy.linalg as sla

from ._lib import C128, check, lib
from .ctm import CTMTensors, _dp, _hpair, _ptr, get_gs_energy, get_gs_norm


class ProjectorPath:
    """The projectors of the first ``steps`` moves x 4 of every nested run
    from ``ts`` at bond dimension cfg["chi"] (ipeps_es_path_create).  The plain
    parts of a nested environment evolve alike whatever B and the momentum
    (nested_tensor.jl:48-63), so all columns and momenta started from ``ts``
    share them (ctmrg.jl:122-265); ``steps`` defaults to max_iter +
    ad_max_iter, the most a column runs.  Valid for ``ts`` and its clones
    while they stay in the state it was made from."""

    def __init__(self, ts, cfg, steps=None):
        if ts.code != C128:
            raise ValueError("ProjectorPath needs a complex (c128) environment")
        if steps is None:
            steps = int(cfg.get("max_iter", 20)) + int(cfg.get("ad_max_iter", 4))
        self.chi = int(cfg.get("chi", 30))
        h = C.c_void_p()
        check(lib().ipeps_es_path_create(ts.h, self.chi, int(steps), C.byref(h)))
        self.h = h

    @property
    def moves(self):
        v = C.c_int64()
        check(lib().ipeps_es_path_moves(self.h, C.byref(v)))
        return v.value

    def close(self):
        if getattr(self, "h", None):
            lib().ipeps_es_path_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_es_grad(ts, H, Bi, px, py, cfg, path=None):
    """(gradH, gradN) flattened in Julia (column-major) order, plus (E4, Nb).

    ``ts``: converged, normalised complex ground-state ``CTMTensors``;
    ``H``: the E_gs-shifted two-site terms; (px, py) in units of pi;
    ``path``: an optional ``ProjectorPath`` made from ``ts`` (or an
    environment in the same state).
    """
    if ts.code != C128:
        raise ValueError("get_es_grad needs a complex (c128) environment")
    shape = ts.A.shape
    B = np.asfortranarray(np.asarray(Bi, dtype=np.complex128).reshape(shape, order="F"))
    (h1, h2), m = _hpair(H)
    gH = np.empty(shape, dtype=np.complex128, order="F")
    envB = np.empty(shape, dtype=np.complex128, order="F")
    E4 = np.zeros(1)
    Nb = np.zeros(2)
    args = (_ptr(h1), _ptr(h2), m, float(px), float(py), _ptr(B), int(cfg.get("chi", 30)),
            int(cfg.get("min_iter", 4)), int(cfg.get("max_iter", 20)), int(cfg.get("ad_max_iter", 4)),
            float(cfg.get("rg_tol", 1e-6)), _ptr(gH), _ptr(envB), _dp(E4), _dp(Nb))
    if path is None:
        check(lib().ipeps_es_grad(ts.h, *args))
    else:
        check(lib().ipeps_es_grad_path(ts.h, path.h, *args))
    return gH.reshape(-1, order="F"), envB.reshape(-1, order="F"), (float(E4[0]), complex(Nb[0], Nb[1]))


def optim_es(ts, H, basis, px, py, cfg, columns=None, workers=1, progress=None, path=True):
    """effH, effN, envB for the tangent basis at momentum (px, py) pi (optim_es.jl:127-153).

    The columns are independent nested runs (the reference splits them over
    jobs with es_resume / es_num, optim_es.jl:127-135).  ``workers`` > 1 runs
    them concurrently: each worker owns a context (stream, allocator, solver
    handle) with a copy of the environment and a host thread, and takes every
    ``workers``-th column.  Run alone, a column gives the same bits on any
    context; run concurrently, rare differences of ~5e-7 relative in a column
    have been seen, below the nested run's stopping tolerance -- as in the
    batched environments, from the rocSOLVER factorizations (zpotrf / zheevd)
    on this path: this library's own kernels use no atomics and read no
    unwritten memory (the GPU tests pass with every fresh allocation filled
    with NaN, IPEPS_POOL_POISON=1).

    ``path``: True (default) computes the columns' shared projectors once
    (``ProjectorPath``), False recomputes them in every column, or a
    ``ProjectorPath`` made from ``ts``.
    """
    own_path = path is True
    if own_path:
        path = ProjectorPath(ts, cfg)
    elif path is False:
        path = None
    try:
        return _optim_es(ts, H, basis, px, py, cfg, columns, workers, progress, path)
    finally:
        if own_path:
            path.close()


def _optim_es(ts, H, basis, px, py, cfg, columns, workers, progress, path):
    basis = np.asarray(basis, dtype=np.complex128)
    n = basis.shape[1]
    effH = np.zeros((n, n), dtype=np.complex128)
    effN = np.zeros((n, n), dtype=np.complex128)
    envB = np.zeros(basis.shape, dtype=np.complex128)
    cols = list(range(n) if columns is None else columns)

    done = [0]

    def store(i, gH, gN):
        envB[:, i] = gN
        effH[:, i] = basis.conj().T @ gH / 2
        effN[:, i] = basis.conj().T @ gN
        done[0] += 1
        if progress is not None:
            progress(done[0], len(cols))

    if workers <= 1 or len(cols) <= 1:
        for i in cols:
            gH, gN, _ = get_es_grad(ts, H, basis[:, i], px, py, cfg, path)
            store(i, gH, gN)
        return effH, effN, envB
    from concurrent.futures import ThreadPoolExecutor
    from .ctm import Context
    nw = min(int(workers), len(cols))
    ctxs = [Context() for _ in range(nw - 1)]
    envs = [ts] + [ts.clone(c) for c in ctxs]

    import threading
    lock = threading.Lock()

    def run(w):
        for i in cols[w::nw]:
            gH, gN = get_es_grad(envs[w], H, basis[:, i], px, py, cfg, path)[:2]
            with lock:
                store(i, gH, gN)

    try:
        with ThreadPoolExecutor(nw) as ex:
            for _ in ex.map(run, range(nw)):
                pass
    finally:
        for e in envs[1:]:
            e.__del__()
        for c in ctxs:
            c.close()
    return effH, effN, envB


def normalize_gs(ts):
    """optim_es.jl:212-222: A -> A / sqrt|<A|env|A>|."""
    nrm = get_gs_norm(ts)
    A = ts.A / np.sqrt(abs(nrm))
    ts.set_A(A)
    return ts


def substract_gs_energy(ts, H):
    """optim_es.jl:224-234: H_i - (E_gs/2) 1."""
    E, _ = get_gs_energy(ts, H)
    e = E.real / 2
    return [np.asarray(h, dtype=np.complex128) - e * np.eye(np.asarray(h).shape[0]) for h in H]


def get_tangent_basis(ts, basis_t="default"):
    """optim_es.jl:236-258.  basis_t (the config key of the same name):
    "default": nullspace of the single-site environment contracted with Ad,
    (D^4 d, D^4 d - 1); "unit": the identity; "gauge": the Q-operator images
    of the unit tensors of a (D/2)^4 d site without flux and with a sigma^z,
    sigma^x, sigma^y flux (get_gauge_basis, optim_es.jl:294-344); "cut" is
    refused: its diag_n_dm (optim_es.jl:261) is defined only in the
    reference's test/test_basis_entropy.jl, not in its package."""
    from .ctm import tcon
    if basis_t == "unit":
        return np.eye(ts.A.size, dtype=np.complex128)
    if basis_t == "gauge":
        from .models import act_Q_op
        D, d = ts.A.shape[0] // 2, ts.A.shape[4]
        M = D ** 4 * d
        cols = []
        for add in (0, 3, 1, 2):  # basis0 (no flux), basis1 (z), basis2 (x), basis3 (y)
            for i in range(M):
                B0 = np.zeros(M)
                B0[i] = 1
                cols.append(act_Q_op(B0.reshape(D, D, D, D, d, order="F"), add=add).reshape(-1, order="F"))
        return np.stack(cols, axis=1).astype(np.complex128)
    if basis_t == "cut":
        raise NotImplementedError("basis_t='cut' needs diag_n_dm, which the reference defines only in its test "
                                  "file test/test_basis_entropy.jl")
    if basis_t != "default":
        raise ValueError(f"unknown basis_t {basis_t!r}")
    Cs, Es = ts.Cs, ts.Es
    C1E1 = tcon([Cs[0], Es[0]], [[-1, 1], [1, -2, -3, -4]])
    CEC2 = tcon([C1E1, Cs[1]], [[-1, 1, -3, -4], [1, -2]])
    CECE4 = tcon([CEC2, Es[3]], [[1, -2, -3, -5], [1, -1, -4, -6]])
    C4E3 = tcon([Cs[3], Es[2]], [[-1, 1], [1, -2, -3, -4]])
    CEC3 = tcon([C4E3, Cs[2]], [[-1, 1, -3, -4], [-2, 1]])
    CECE2 = tcon([CEC3, Es[1]], [[-1, 1, -3, -5], [-2, 1, -4, -6]])
    ndm = tcon([CECE4, CECE2], [[1, 2, -1, -2, -5, -6], [1, 2, -3, -4, -7, -8]])
    Ad = np.conj(ts.A)
    ndm_Ad = tcon([ndm, Ad], [[-1, -2, -3, -4, 1, 2, 3, 4], [1, 2, 3, 4, -5]])
    row = ndm_Ad.reshape(1, -1, order="F")
    return sla.null_space(row)


def prepare_basis(A, H, cfg, dtype="c128", options=None, ctx=None):
    """prepare_basis (optim_es.jl:34-73) without the file I/O: A ->
    renormalize(A); converge the ground-state environment with max_iter
    doubled and rg_tol squared (energy convergence); normalize_gs;
    substract_gs_energy; get_tangent_basis (cfg "basis_t", default
    "default").  Returns (basis, ts, shifted H); ts carries cfg's own
    min/max_iter and rg_tol again, as setproperties(ts, Params = cfg0)."""
    from .ctm import run_ctm
    A = np.asarray(A)
    A = A / np.abs(A).max()
    conv_cfg = dict(cfg)
    conv_cfg["max_iter"] = 2 * int(cfg.get("max_iter", 20))
    conv_cfg["rg_tol"] = float(cfg.get("rg_tol", 1e-6)) ** 2
    ts = CTMTensors(A, conv_cfg, dtype=dtype, options=options, ctx=ctx)
    run_ctm(ts, "energy", H)
    ts.Params = dict(cfg)
    ts = normalize_gs(ts)
    Hs = substract_gs_energy(ts, H)
    basis = get_tangent_basis(ts, cfg.get("basis_t", "default"))
    return basis, ts, Hs


def compute_es(effH, effN, nrmB_cut=1e-3):
    """expectation.jl:13-56: Hermitise, project on the leading norm
    eigenvectors, solve H2 v = e N2 v.  Returns (es ascending, vecs, P)."""
    H = (effH + effH.conj().T) / 2
    N = (effN + effN.conj().T) / 2
    ev_N, P = np.linalg.eigh(N)
    idx = np.argsort(ev_N.real)[::-1]
    ev_N = ev_N[idx]
    if isinstance(nrmB_cut, (int, np.integer)):
        selected = ev_N > ev_N[nrmB_cut]
    else:
        selected = (ev_N / ev_N.max()) > nrmB_cut
    P = P[:, idx][:, selected]
    N2 = P.conj().T @ N @ P
    H2 = P.conj().T @ H @ P
    H2 = (H2 + H2.conj().T) / 2
    N2 = (N2 + N2.conj().T) / 2
    es, vecs = sla.eigh(H2, N2)
    ixs = np.argsort(es.real)
    return es[ixs], vecs[:, ixs], P


def band_structure(ts, H, basis, momenta, cfg, n_bands=4, nrmB_cut=1e-3, workers=1, progress=None, path=True):
    """The excited-state pipeline over a list of momenta (px, py in units of
    pi): at each, the effH / effN columns of every tangent basis vector
    (optim_es, optim_es.jl:91-165) and the generalised eigenproblem
    (compute_es, expectation.jl:13-56).  Returns a list of dicts with the
    momentum, the lowest n_bands energies, effH, effN and the seconds taken
    (the first momentum's include making the shared projector path, which
    every momentum uses: ``path`` as in optim_es)."""
    import time
    out = []
    own_path = path is True
    try:
        for px, py in momenta:
            t0 = time.perf_counter()
            if path is True:
                path = ProjectorPath(ts, cfg)
            effH, effN, _ = optim_es(ts, H, basis, px, py, cfg, workers=workers, progress=progress, path=path)
            es, _, _ = compute_es(effH, effN, nrmB_cut)
            out.append(dict(px=float(px), py=float(py), bands=np.real(es[:n_bands]), effH=effH, effN=effN,
                            seconds=time.perf_counter() - t0))
    finally:
        if own_path and isinstance(path, ProjectorPath):
            path.close()
    return out


def make_es_path(step=0.1):
    """optim_es.jl:1-20: the (kx, ky) path in units of pi."""
    n = int(round(1 / step))
    kx1 = np.ones(n - 1)
    ky1 = np.arange(1, n) * step
    kx2 = 1 - np.arange(0, n + 1) * step
    ky2 = 1 - np.arange(0, n + 1) * step
    kx3 = np.arange(1, n) * step
    ky3 = np.zeros(n - 1)
    m = int(round(0.5 / step))
    kx4 = 1 - np.arange(0, m) * step
    ky4 = np.arange(0, m) * step
    return np.concatenate([kx1, kx2, kx3, kx4]), np.concatenate([ky1, ky2, ky3, ky4])
