"""optim_gs over the HIP engine: the LBFGS loop of src/optimize/optim_gs.jl
on the device-computed energy and gradient.

* ``optim_gs(H, A0, cfg, m=10, f_tol=1e-7, g_tol=1e-6, iterations=200)``
  optim_gs.jl:14-99.  The objective is fg! (optim_gs.jl:53-91): x ->
  renormalize(x), CTMRG converged without AD, ad_max_iter differentiated
  steps from the converged environment held constant, Re E and its gradient
  -- one ``gs_energy_grad`` call on the GPU, cached for a repeated x as the
  reference caches it.  The optimiser is limited-memory BFGS on the unit
  sphere (Optim's ``LBFGS(m=m, manifold=Sphere())``): the gradient projected
  on the tangent space, the two-loop recursion with the initial inverse
  Hessian scaled by s.y / y.y, steps retracted to the sphere; the line
  search is backtracking on the Armijo condition (Optim's default is
  Hager-Zhang, so the iterates differ from the reference's while the
  objective is the same).  Stops when max|projected g| <= g_tol, when the
  relative change of the energy is <= f_tol, or after ``iterations``.
* Per-iteration checkpoint (optim_gs.jl:27-51, ``gs.jld2``): with
  ``save_path`` the iterate A, the energies and gradient norms of the
  accepted iterates are written to an .npz after every iteration, and
  ``resume=True`` restarts from that file (A, energies, gradnorms), as the
  reference's ``cfg["resume"]``.
"""
import os

import numpy as np

from .ctm import gs_energy_grad


def _rdot(a, b):
    return float(np.real(np.vdot(a, b)))


def _sphere_retract(x):
    return x / np.linalg.norm(x)


def _sphere_project(g, x):
    return g - _rdot(x, g) * x


class _Objective:
    """fg! with its cache (optim_gs.jl:53-91)."""

    def __init__(self, H, cfg, dtype, options):
        self.H, self.cfg, self.dtype, self.options = H, cfg, dtype, options
        self.cached = None
        self.calls = 0

    def __call__(self, x):
        xr = x / np.abs(x).max()  # renormalize(x)
        if self.cached is not None and np.linalg.norm(xr - self.cached[0]) < 1e-14:
            return self.cached[1], self.cached[2]
        E, g, _ = gs_energy_grad(x, self.H, self.cfg, dtype=self.dtype, options=self.options)
        self.calls += 1
        self.cached = (xr, float(E.real), g)
        return float(E.real), g


def optim_gs(H, A0, cfg, m=10, f_tol=1e-7, g_tol=1e-6, iterations=200, dtype=None, options=None, save_path=None,
             resume=False, c1=1e-4, max_backtracks=12, verbose=False):
    """Returns dict(A, energy, energies, gradnorms, iterations, converged,
    fg_calls).  ``energies`` / ``gradnorms`` hold the energy and |grad| of
    each accepted iterate, the first being A0's (the reference's
    gs.jld2 fields of the same names)."""
    energies, gradnorms = [], []
    x = np.array(A0, dtype=np.complex128 if np.iscomplexobj(A0) else np.float64)
    if resume and save_path and os.path.exists(save_path):
        z = np.load(save_path)
        x = z["A"]
        energies, gradnorms = list(z["energies"]), list(z["gradnorms"])
    fg = _Objective(H, cfg, dtype, options)

    def save():
        if save_path:
            tmp = save_path + ".tmp.npz"
            np.savez(tmp, A=x, energies=np.array(energies), gradnorms=np.array(gradnorms))
            os.replace(tmp, save_path)

    x = _sphere_retract(x)
    f, graw = fg(x)
    g = _sphere_project(graw, x)
    if not energies:
        energies.append(f)
        gradnorms.append(float(np.linalg.norm(graw)))
        save()
    S, Y = [], []
    converged = False
    it = 0
    for it in range(1, iterations + 1):
        if np.abs(g).max() <= g_tol:
            converged = True
            break
        # two-loop recursion on the tangent vectors
        q = g.copy()
        al = []
        for s, y in reversed(list(zip(S, Y))):
            rho = 1.0 / _rdot(y, s)
            a = rho * _rdot(s, q)
            al.append((a, rho, s, y))
            q = q - a * y
        if S:
            q = q * (_rdot(S[-1], Y[-1]) / _rdot(Y[-1], Y[-1]))
        for a, rho, s, y in reversed(al):
            b = rho * _rdot(y, q)
            q = q + (a - b) * s
        d = _sphere_project(-q, x)
        slope = _rdot(g, d)
        if not slope < 0:  # not a descent direction: restart from steepest descent
            S.clear()
            Y.clear()
            d = -g
            slope = _rdot(g, d)
        alpha = 1.0 if S else min(1.0, 0.1 / max(np.linalg.norm(d), 1e-300))
        accepted = False
        for _ in range(max_backtracks):
            xn = _sphere_retract(x + alpha * d)
            fn, gn_raw = fg(xn)
            if fn <= f + c1 * alpha * slope:
                accepted = True
                break
            alpha *= 0.5
        if not accepted:
            break
        gn = _sphere_project(gn_raw, xn)
        s_vec, y_vec = xn - x, gn - g
        if _rdot(s_vec, y_vec) > 1e-300:
            S.append(s_vec)
            Y.append(y_vec)
            if len(S) > m:
                S.pop(0)
                Y.pop(0)
        f_prev = f
        x, f, g = xn, fn, gn
        energies.append(f)
        gradnorms.append(float(np.linalg.norm(gn_raw)))
        save()
        if verbose:
            print(f" Step {it:3d}  E: {f:.10f}  |grad|: {gradnorms[-1]:.3e}")
        if abs(f - f_prev) <= f_tol * abs(f):
            converged = True
            break
    return dict(A=x, energy=f, energies=np.array(energies), gradnorms=np.array(gradnorms), iterations=it,
                converged=converged, fg_calls=fg.calls)
