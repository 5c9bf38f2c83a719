"""Host-side mirror of the reference CTMRG API over the HIP engine.

Names, arguments and return values follow yitan1/iPEPS.jl so that code (and
tests) written against the reference read the same:

* ``CTMTensors(A, cfg)``             src/ctm/ctm_tensor.jl:63-69 (init_ctm1)
* ``rg_step(ts, chi)``               src/ctm/ctmrg.jl:57-64
* ``run_ctm(ts, conv_fun=...)``      src/ctm/ctmrg.jl:1-55
* ``get_gs_energy(ts, H)``           src/graph/evaluation.jl:15-29
* ``get_gs_norm(ts)``                src/graph/evaluation.jl:47-59
* ``compute_gs_energy(A, H, cfg)``   src/graph/evaluation.jl:1-13
* ``gs_energy_grad(A0, H, cfg)``     fg! of src/optimize/optim_gs.jl:53-91
* ``tcon(xs, ind_xs)``               src/utility/wrapper.jl:15-23
* ``svd1(A)`` / ``svd_back(...)``    src/utility/autodiff.jl:22-75

Arrays are numpy arrays in Julia index order (X[i1, ...] is Julia
X[i1+1, ...]); they are handed to the C ABI in column-major memory order.
Everything runs on the GPU through libipeps_hip.so; there is no CPU path.
"""
import ctypes as C
import os

import numpy as np

from ._lib import C128, F64, MIXED, IpepsError, check, lib

_DT = {"f64": F64, "c128": C128, "mixed": MIXED}


def _dtype_code(dtype, A):
    if dtype is None:
        return C128 if np.iscomplexobj(A) else F64
    if isinstance(dtype, str):
        return _DT[dtype]
    return int(dtype)


def _host_np(code):
    return np.complex128 if code == C128 else np.float64


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i64(seq):
    return (C.c_int64 * len(seq))(*[int(x) for x in seq])


class Context:
    """A device context (stream, allocator, solver handle)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().ipeps_ctx_create(int(device), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().ipeps_ctx_destroy(self.h)
            self.h = None

    def sync(self):
        check(lib().ipeps_ctx_sync(self.h))

    def mem(self):
        live, peak = C.c_int64(), C.c_int64()
        check(lib().ipeps_ctx_mem(self.h, C.byref(live), C.byref(peak)))
        return live.value, peak.value

    def set_timing(self, on=True, prefix=None):
        """on: False/0 off, True/1 per kernel, 2 per kernel and GEMM shape;
        prefix: only kernels whose name starts with it are timed."""
        check(lib().ipeps_ctx_set_timing_filter(self.h, (prefix or "").encode()))
        check(lib().ipeps_ctx_set_timing(self.h, int(on)))

    def kernel_stats(self, reset=True):
        """{kernel: (launches, total_ms, algorithmic_flops, algorithmic_bytes)}."""
        import json
        buf = C.create_string_buffer(1 << 20)
        check(lib().ipeps_ctx_kernel_stats(self.h, buf, len(buf), int(bool(reset))))
        return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = None


def default_context():
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


def _hpair(H):
    hs = [np.asfortranarray(np.asarray(h, dtype=np.complex128)) for h in H]
    if len(hs) == 1:
        hs = hs * 2
    m = hs[0].shape[0]
    for h in hs:
        if h.shape != (m, m):
            raise ValueError("H terms must be square and equal-sized")
    return hs, m


ENV_OPTIONS = ("svd_mode", "oversample", "power_iters", "full_below", "seed", "eta", "mixed_cutoff", "fast_orth",
               "fast_svd",
               "gram_guard", "power_iters_max", "resid_tol", "complement_iters", "complement_tol", "warm_start", "canonical_edge", "mixed", "mixed_back", "mixed_proj", "energy_fp64", "energy_mixed",
               "energy_mixed_back",
               "checkpoint", "fused_absorb", "paired_moves", "draws", "early_mixed", "late_apps", "graph_safe")


class CTMTensors:
    """Device-resident plain environment (src/ctm/ctm_tensor.jl:3-18).

    ``CTMTensors(A, cfg)`` runs init_ctm1 on the GPU.  ``dtype`` is "f64",
    "c128" or "mixed" (FP32 storage, split-bf16 MFMA contractions, FP64
    projector SVD and energy).  ``options`` are engine options (see
    ``ENV_OPTIONS``).
    """

    def __init__(self, A, cfg=None, dtype=None, options=None, ctx=None):
        self.ctx = ctx or default_context()
        self.Params = dict(cfg or {})
        A = np.asarray(A)
        if A.ndim != 5:
            raise ValueError("A must be (D, D, D, D, d)")
        self.code = _dtype_code(dtype, A)
        self._np = _host_np(self.code)
        Af = np.asfortranarray(A, dtype=self._np)
        h = C.c_void_p()
        check(lib().ipeps_env_create(self.ctx.h, self.code, _i64(A.shape), _ptr(Af), C.byref(h)))
        self.h = h
        self.options = {}
        for k, v in (options or {}).items():
            self.set_option(k, v)

    def __del__(self):
        try:
            if self.h:
                lib().ipeps_env_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_option(self, key, value):
        check(lib().ipeps_env_set_option(self.h, key.encode(), float(value)))
        self.options[key] = value

    def clone(self, ctx=None):
        """A copy of this environment (A, C, E, options) on ``ctx`` (another
        context: its own stream, allocator and solver handle)."""
        t = CTMTensors(self.A, self.Params, dtype=self.code, options=self.options, ctx=ctx)
        t.set_env(self.Cs, self.Es)
        t.set_option("draws", self.draws)
        return t

    @property
    def draws(self):
        """Position in the environment's stream of random test matrices."""
        v = C.c_int64()
        check(lib().ipeps_env_get_draws(self.h, C.byref(v)))
        return v.value

    def tensor(self, which):
        dims = (C.c_int64 * 5)()
        nd = C.c_int()
        check(lib().ipeps_env_tensor_dims(self.h, int(which), dims, C.byref(nd)))
        shape = tuple(dims[i] for i in range(nd.value))
        out = np.empty(shape, dtype=self._np, order="F")
        check(lib().ipeps_env_get_tensor(self.h, int(which), _ptr(out)))
        return out

    def set_tensor(self, which, arr):
        a = np.asfortranarray(arr, dtype=self._np)
        check(lib().ipeps_env_set_tensor(self.h, int(which), a.ndim, _i64(a.shape), _ptr(a)))

    @property
    def Cs(self):
        return [self.tensor(i) for i in range(4)]

    @property
    def Es(self):
        return [self.tensor(4 + i) for i in range(4)]

    @property
    def A(self):
        return self.tensor(8)

    def set_A(self, A):
        a = np.asfortranarray(A, dtype=self._np)
        check(lib().ipeps_env_set_A(self.h, _ptr(a)))

    def projector_stats(self, reset=False):
        """Counters of the randomized projectors since the last reset: dict with
        calls, power_iters, power_max, resid_max, checks, cg_calls, cg_iters,
        cg_resid_max, cg_fallbacks, orth_calls, orth_second, orth_householder,
        warm_starts, tri_fallbacks."""
        out = np.zeros(14)
        check(lib().ipeps_env_projector_stats(self.h, _dp(out), int(bool(reset))))
        keys = ("calls", "power_iters", "power_max", "resid_max", "checks", "cg_calls", "cg_iters", "cg_resid_max",
                "cg_fallbacks", "orth_calls", "orth_second", "orth_householder", "warm_starts",
                "tri_fallbacks")
        return {k: (float(v) if k.endswith("max") and "power" not in k else int(v)) for k, v in zip(keys, out)}

    def copy_state_from(self, other):
        """C, E, the test-matrix stream position and the warm-start bases of
        ``other`` (same A shape), so this environment continues as it would."""
        check(lib().ipeps_env_copy_state(self.h, other.h))

    def set_env(self, Cs, Es):
        for i in range(4):
            self.set_tensor(i, Cs[i])
            self.set_tensor(4 + i, Es[i])


def rg_step(ts, chi):
    """One left/right/top/bottom step; returns (ts, left spectrum)."""
    s = np.empty(int(chi) + 2, dtype=np.float64)
    n = C.c_int64()
    check(lib().ipeps_rg_step(ts.h, int(chi), _dp(s), s.size, C.byref(n)))
    return ts, s[: n.value].copy()


def run_graph(ts, chi, steps):
    """``steps`` rg_steps as replays of a captured HIP graph of the step
    (ipeps_env_run_graph); returns dict(captured, replayed, undone, eager)."""
    out = (C.c_int64 * 4)()
    check(lib().ipeps_env_run_graph(ts.h, int(chi), int(steps), out))
    return dict(captured=out[0], replayed=out[1], undone=out[2], eager=out[3])


def run_ctm(ts, conv_fun="energy", H=None):
    """run_ctm with the reference stopping rule.

    conv_fun: "energy" (get_gs_energy with ``H``) or None (left spectrum).
    Params keys chi, min_iter, max_iter, rg_tol as in the reference
    (fallbacks 30, 4, 20, 1e-6).  Returns (ts, conv) where conv is the last
    energy (complex) or spectrum; ts.last_iters / ts.last_diffs are set.
    """
    p = ts.Params
    chi = int(p.get("chi", 30))
    min_iter = int(p.get("min_iter", 4))
    max_iter = int(p.get("max_iter", 20))
    tol = float(p.get("rg_tol", 1e-6))
    mode = 0 if conv_fun == "energy" else 1
    if mode == 0:
        (h1, h2), m = _hpair(H)
        p1, p2 = _ptr(h1), _ptr(h2)
    else:
        p1 = p2 = None
        m = 1
    iters = C.c_int()
    E = np.zeros(2)
    diffs = np.zeros(max_iter + 2)
    nd = C.c_int64()
    check(lib().ipeps_run_ctm(ts.h, chi, min_iter, max_iter, tol, mode, p1, p2, m, C.byref(iters), _dp(E),
                              _dp(diffs), diffs.size, C.byref(nd)))
    ts.last_iters = iters.value
    ts.last_diffs = diffs[: nd.value].copy()
    if mode == 0:
        return ts, complex(E[0], E[1])
    return ts, None


def get_gs_energy(ts, H):
    """(E, Na) as complex numbers."""
    (h1, h2), m = _hpair(H)
    E = np.zeros(2)
    N = np.zeros(2)
    check(lib().ipeps_gs_energy(ts.h, _ptr(h1), _ptr(h2), m, _dp(E), _dp(N)))
    return complex(E[0], E[1]), complex(N[0], N[1])


def get_gs_norm(ts):
    out = np.zeros(2)
    check(lib().ipeps_gs_norm(ts.h, _dp(out)))
    return complex(out[0], out[1])


def compute_gs_energy(A, H, cfg, dtype=None, options=None, ctx=None):
    """src/graph/evaluation.jl:1-13: converge from init_ctm1 and evaluate E."""
    ts = CTMTensors(A, cfg, dtype=dtype, options=options, ctx=ctx)
    ts, _ = run_ctm(ts, "energy", H)
    E, N = get_gs_energy(ts, H)
    return E, N, ts


def step_adjoint(ts, chi, H):
    """One rg_step + energy + the hand-written adjoint; returns (E, dRe(E)/dA)."""
    (h1, h2), m = _hpair(H)
    E = np.zeros(2)
    A = ts.A
    g = np.empty(A.shape, dtype=np.complex128 if ts.code == C128 else np.float64, order="F")
    check(lib().ipeps_step_adjoint(ts.h, int(chi), _ptr(h1), _ptr(h2), m, _dp(E), _ptr(g)))
    return complex(E[0], E[1]), g


def gs_energy_grad(A0, H, cfg, dtype=None, options=None, ctx=None):
    """Energy and gradient as optim_gs's fg! (src/optimize/optim_gs.jl:53-91).

    Returns (E complex, g, (converging iterations, differentiated steps)).
    """
    ctx = ctx or default_context()
    A0 = np.asarray(A0)
    code = _dtype_code(dtype, A0)
    Af = np.asfortranarray(A0, dtype=_host_np(code))
    (h1, h2), m = _hpair(H)
    chi = int(cfg.get("chi", 30))
    min_iter = int(cfg.get("min_iter", 4))
    max_iter = int(cfg.get("max_iter", 20))
    ad_max_iter = int(cfg.get("ad_max_iter", 4))
    tol = float(cfg.get("rg_tol", 1e-6))
    opts = ";".join(f"{k}={float(v)}" for k, v in (options or {}).items()).encode()
    E = np.zeros(2)
    g = np.empty(A0.shape, dtype=_host_np(code), order="F")
    it = (C.c_int * 2)()
    check(lib().ipeps_gs_energy_grad(ctx.h, code, _i64(A0.shape), _ptr(Af), _ptr(h1), _ptr(h2), m, chi,
                                     min_iter, max_iter, ad_max_iter, tol, opts, _dp(E), _ptr(g), it))
    return complex(E[0], E[1]), g, (it[0], it[1])


def run_gs_vjp(ts, x, H, cfg, dE=1.0):
    """run_gs (src/optimize/optim_gs.jl:133-143) and its pullback: from ts's
    environment, held constant, with A = x, the differentiated steps
    (min_iter, max_iter = cfg["ad_max_iter"], rg_tol) and the energy.
    Returns (E complex, dE * dRe(E)/dx, differentiated steps); ts is not
    modified.  The rrule of run_gs is x -> (E, dE -> this gradient)."""
    x = np.asarray(x)
    xf = np.asfortranarray(x, dtype=_host_np(ts.code))
    (h1, h2), m = _hpair(H)
    E = np.zeros(2)
    g = np.empty(x.shape, dtype=_host_np(ts.code), order="F")
    n = C.c_int()
    check(lib().ipeps_run_gs_vjp(ts.h, _ptr(xf), _ptr(h1), _ptr(h2), m, int(cfg.get("chi", 30)),
                                 int(cfg.get("min_iter", 4)), int(cfg.get("ad_max_iter", 4)),
                                 float(cfg.get("rg_tol", 1e-6)), float(dE), _dp(E), _ptr(g), C.byref(n)))
    return complex(E[0], E[1]), g, n.value


def tcon(xs, ind_xs, ctx=None, fp32=False):
    """Pairwise ncon contraction on the device (src/utility/wrapper.jl:15-23).
    fp32=True: real data stored and multiplied in FP32 (the mixed engine's
    storage; FP32 products), float32 result."""
    if len(xs) != 2:
        raise ValueError("tcon contracts two tensors")
    ctx = ctx or default_context()
    A, B = (np.asarray(x) for x in xs)
    cplx = np.iscomplexobj(A) or np.iscomplexobj(B)
    if fp32 and cplx:
        raise ValueError("tcon: fp32 takes real data")
    code = MIXED if fp32 else (C128 if cplx else F64)
    dt = np.float32 if fp32 else _host_np(code)
    Af, Bf = np.asfortranarray(A, dtype=dt), np.asfortranarray(B, dtype=dt)
    la = (C.c_int * max(1, A.ndim))(*ind_xs[0])
    lb = (C.c_int * max(1, B.ndim))(*ind_xs[1])
    nout = sum(1 for l in list(ind_xs[0]) + list(ind_xs[1]) if l < 0)
    dimsC = (C.c_int64 * 16)()
    ndC = C.c_int()
    sizes = {}
    for ind, X in ((ind_xs[0], A), (ind_xs[1], B)):
        for l, d in zip(ind, X.shape):
            if l < 0:
                sizes[-l - 1] = d
    shape = tuple(sizes[k] for k in range(nout))
    out = np.empty(shape, dtype=dt, order="F")
    check(lib().ipeps_tcon(ctx.h, code, A.ndim, _i64(A.shape), _ptr(Af), la, B.ndim, _i64(B.shape), _ptr(Bf), lb,
                           dimsC, C.byref(ndC), _ptr(out), max(1, out.size)))
    return out


def tune(mode=1, min_flops=1e8):
    """GEMM autotuner, process-wide.  mode 0: tiles and split-K by shape rules;
    1: the first product of each unseen shape (>= min_flops) times the
    candidate launch plans and keeps the fastest; 2: table entries only."""
    check(lib().ipeps_tune_config(int(mode), float(min_flops)))


def tune_stats():
    """(entries, shapes tuned, table hits) of the GEMM tuning table."""
    out = (C.c_int64 * 3)()
    check(lib().ipeps_tune_stats(out))
    return dict(entries=out[0], tuned=out[1], hits=out[2])


def tune_clear():
    check(lib().ipeps_tune_clear())


def tune_save(path):
    check(lib().ipeps_tune_save(os.fsencode(path)))


def tune_load(path):
    check(lib().ipeps_tune_load(os.fsencode(path)))


def gemm(A, B, transA=False, conjA=False, transB=False, conjB=False, mixed=None, alpha=1.0, beta=0.0, C=None,
         ctx=None):
    """alpha op(A) op(B) + beta C on the device (FP64 / complex FP64 MFMA, or
    split-bf16 MFMA on float32 data when ``mixed`` is 1, 3 or 6)."""
    ctx = ctx or default_context()
    A, B = np.asarray(A), np.asarray(B)
    if mixed:
        code, dt = MIXED, np.float32
    else:
        cplx = np.iscomplexobj(A) or np.iscomplexobj(B) or (C is not None and np.iscomplexobj(C))
        code = C128 if cplx else F64
        dt = _host_np(code)
    Af, Bf = np.asfortranarray(A, dtype=dt), np.asfortranarray(B, dtype=dt)
    M = A.shape[1] if transA else A.shape[0]
    K = A.shape[0] if transA else A.shape[1]
    N = B.shape[0] if transB else B.shape[1]
    Cf = np.zeros((M, N), dtype=dt, order="F") if C is None else np.asfortranarray(C, dtype=dt).copy(order="F")
    check(lib().ipeps_gemm(ctx.h, code, int(mixed or 0), int(transA), int(conjA), int(transB), int(conjB), M, N, K,
                           float(alpha), _ptr(Af), max(1, Af.shape[0]), _ptr(Bf), max(1, Bf.shape[0]), float(beta),
                           _ptr(Cf), max(1, M)))
    return Cf


def absorb(X, lx, A, la, Ad, lad, G=None, mixed=True, fused=True, ctx=None):
    """Y = X A Ad (ncon labels; one enlarged-corner site absorption) on the
    device; with a cotangent G also (gX, gA, gAd).  mixed: float32 data
    (fused FP32-MFMA kernel when fused and 4 <= D <= 6, d <= 2); else FP64."""
    ctx = ctx or default_context()
    dt = np.float32 if mixed else np.float64
    Xf, Af, Adf = (np.asfortranarray(np.asarray(x), dtype=dt) for x in (X, A, Ad))
    sizes = {}
    for ind, T in ((lx, Xf), (la, Af), (lad, Adf)):
        for l, d in zip(ind, T.shape):
            if l < 0:
                sizes[-l - 1] = d
    Y = np.empty(tuple(sizes[k] for k in range(len(sizes))), dtype=dt, order="F")
    ints = lambda l: (C.c_int * len(l))(*[int(v) for v in l])
    Gf = None if G is None else np.asfortranarray(np.asarray(G), dtype=dt)
    gX = np.empty_like(Xf) if G is not None else None
    gA = np.empty_like(Af) if G is not None else None
    gAd = np.empty_like(Adf) if G is not None else None
    p = lambda a: None if a is None else _ptr(a)
    check(lib().ipeps_absorb(ctx.h, MIXED if mixed else F64, int(bool(fused)), _i64(Xf.shape), _ptr(Xf), ints(lx),
                             _i64(Af.shape), _ptr(Af), ints(la), _ptr(Adf), ints(lad), p(Gf), _ptr(Y), p(gX), p(gA),
                             p(gAd)))
    return Y if G is None else (Y, gX, gA, gAd)


def svd1(A, ctx=None):
    """Full SVD of a square matrix on the device: (U, S, V) with A = U S V^H."""
    ctx = ctx or default_context()
    A = np.asarray(A)
    code = C128 if np.iscomplexobj(A) else F64
    dt = _host_np(code)
    m, n = A.shape
    Af = np.asfortranarray(A, dtype=dt)
    U = np.empty((m, m), dtype=dt, order="F")
    V = np.empty((n, n), dtype=dt, order="F")
    S = np.empty(min(m, n))
    check(lib().ipeps_svd(ctx.h, code, m, n, _ptr(Af), _ptr(U), _dp(S), _ptr(V)))
    return U, S, V


def orth(Y, method=1, ctx=None):
    """Orthonormal basis of the columns of a tall Y on the device (range-basis
    step of the randomized projector).  method 0: Householder QR; 1:
    CholeskyQR2 (rocSOLVER); 2: CholeskyQR2 on the one-workgroup blocked
    Cholesky (F64, n <= 160)."""
    ctx = ctx or default_context()
    Y = np.asarray(Y)
    code = C128 if np.iscomplexobj(Y) else F64
    dt = _host_np(code)
    m, n = Y.shape
    Yf = np.asfortranarray(Y, dtype=dt)
    Q = np.empty((m, n), dtype=dt, order="F")
    check(lib().ipeps_orth(ctx.h, code, int(method), m, n, _ptr(Yf), _ptr(Q)))
    return Q


def svd_thin(A, method=1, guard=1e-4, ctx=None):
    """Thin SVD of a tall matrix on the device: (U, S, V, method_used) with
    A = U diag(S) V^H.  method 0: one-sided Jacobi; 1: Gram eigenpairs when
    S_min/S_max >= guard (else 0), for real n <= 320 by device
    tridiagonalisation, bisection and inverse iteration; 2: as 1 with the Gram
    eigenpairs from the one-workgroup Jacobi kernel (real, n <= 112); 3: as 1
    with rocSOLVER stedc on the tridiagonal form."""
    ctx = ctx or default_context()
    A = np.asarray(A)
    code = C128 if np.iscomplexobj(A) else F64
    dt = _host_np(code)
    m, n = A.shape
    Af = np.asfortranarray(A, dtype=dt)
    U = np.empty((m, n), dtype=dt, order="F")
    V = np.empty((n, n), dtype=dt, order="F")
    S = np.empty(n)
    used = C.c_int(-1)
    check(lib().ipeps_svd_thin(ctx.h, code, int(method), float(guard), m, n, _ptr(Af), _ptr(U), _dp(S), _ptr(V),
                               C.byref(used)))
    return U, S, V, used.value


def svd_back(U, S, V, dU=None, dS=None, dV=None, eta=1e-12, ctx=None):
    """Reference SVD adjoint for a square matrix (src/utility/autodiff.jl:41-75)."""
    ctx = ctx or default_context()
    code = C128 if np.iscomplexobj(U) else F64
    dt = _host_np(code)
    N = U.shape[0]
    n = next((x.shape[-1] if x.ndim == 2 else x.shape[0]) for x in (dU, dS, dV) if x is not None)
    Uf, Vf = np.asfortranarray(U, dtype=dt), np.asfortranarray(V, dtype=dt)
    Sf = np.ascontiguousarray(S, dtype=np.float64)
    args = []
    for x, isS in ((dU, False), (dS, True), (dV, False)):
        if x is None:
            args.append(None)
        else:
            args.append(np.ascontiguousarray(x, dtype=np.float64) if isS else np.asfortranarray(x, dtype=dt))
    out = np.empty((N, N), dtype=dt, order="F")
    check(lib().ipeps_svd_back(ctx.h, code, N, n, _ptr(Uf), _dp(Sf), _ptr(Vf),
                               _ptr(args[0]) if args[0] is not None else None,
                               _dp(args[1]) if args[1] is not None else None,
                               _ptr(args[2]) if args[2] is not None else None, float(eta), _ptr(out)))
    return out


__all__ = ["absorb", "Context", "CTMTensors", "rg_step", "run_ctm", "get_gs_energy", "get_gs_norm", "compute_gs_energy",
           "step_adjoint", "gs_energy_grad", "tcon", "gemm", "svd1", "svd_back", "default_context", "IpepsError",
           "ENV_OPTIONS", "tune", "tune_stats", "tune_clear", "tune_save", "tune_load", "run_graph"]
