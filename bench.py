"""CTMRG step + adjoint throughput on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3]

Workload (default "c3"): spin-1/2 Heisenberg (Jz=1, d=2), A of shape
(6,6,6,6,2) from numpy default_rng(2).standard_normal (Fortran order),
chi = 144 (half-system matrices 5184 x 5184).  One step = rg_step (4 moves)
+ the energy and its VJP + the hand-written adjoint of the step, from a
chi-saturated environment.  The measured path is the mixed-precision engine
(FP32 storage, split-bf16 MFMA contractions, FP64 projector SVD /
orthogonalisation / energy reduction); the FP64 GPU path runs the same steps
for the speed ratio and the energy / gradient error.  Prints ONE JSON line.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

WORKLOADS = {
    "c3": dict(D=6, chi=144, d=2, seed=2, unit_max=False),
    "c2": dict(D=4, chi=64, d=2, seed=1, unit_max=True),
    "small": dict(D=3, chi=32, d=2, seed=2, unit_max=False),
}
PEAK_TFLOPS = {"f64": 78.6, "bf16": 2500.0, "f32": 157.3}
PEAK_HBM_GBS = 8000.0


def load_pkg():
    if "ipeps_amd" in sys.modules:
        return sys.modules["ipeps_amd"]
    pkg_dir = os.path.join(ROOT, "ipeps.jl_amd")
    spec = importlib.util.spec_from_file_location("ipeps_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ipeps_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def flops_per_unit(D, chi, d=2, k_extra=16, q=1):
    """Algorithmic flops of one step + adjoint (real arithmetic, canonical order)."""
    N = chi * D * D
    k = chi + k_extra
    corners = 2 * (chi ** 3 * D ** 2 + chi ** 3 * D ** 4 + 2 * chi ** 2 * D ** 6 * d)
    proj = 2 * N * N * k * (2 * q + 2) + 4 * N * k * k + 2 * N * N * chi
    renorm = 2 * N * chi ** 2 + 2 * chi ** 3 * D ** 4 + 2 * chi ** 2 * D ** 6 * d
    move = corners + proj + renorm
    f_fwd = 4 * move * 2
    return 3 * f_fwd


def kernel_roofline(stats):
    """Dominant kernel (largest total time) and its achieved rate."""
    name, (launches, ms, flops, bytes_) = max(stats.items(), key=lambda kv: kv[1][1])
    avg_ms = ms / max(1, launches)
    if name.startswith("gemm_"):
        unit = "TFLOP/s"
        peak = PEAK_TFLOPS["bf16"] if name.startswith("gemm_f32") else PEAK_TFLOPS["f64"]
        achieved = (flops / launches) / (avg_ms * 1e-3) / 1e12
        bound = "mfma"
    else:
        unit = "GB/s"
        peak = PEAK_HBM_GBS
        achieved = (bytes_ / launches) / (avg_ms * 1e-3) / 1e9 if bytes_ else 0.0
        bound = "hbm"
    return dict(kernel=name, bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak,
                traffic=None, launches=launches, avg_ms=avg_ms, total_ms=ms)


def cpu_baseline_sample(D, chi, seed):
    """Oracle (reference op order) on one left move at the workload's shape."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctmrg_oracle as O

    threads = min(16, len(os.sched_getaffinity(0)))
    torch.set_num_threads(threads)
    A = O.seeded_tensor((D, D, D, D, 2), seed)
    g = torch.Generator().manual_seed(0)
    Cs = [torch.randn(chi, chi, generator=g, dtype=torch.float64) for _ in range(4)]
    Es = [torch.randn(chi, chi, D, D, generator=g, dtype=torch.float64) for _ in range(4)]
    ts = O.CTMTensors(A, A.conj(), Cs, Es, dict(chi=chi))
    with torch.no_grad():
        t0 = time.perf_counter()
        P, Pd, _ = O.get_projector_left(ts, chi)
        O.proj_left(ts, P, Pd)
        t = time.perf_counter() - t0
    per_unit = 12.0 * t  # 4 moves x (forward + adjoint ~ 2x forward)
    return dict(value=1.0 / per_unit, unit="step+grad/s", cores=threads, kind="port",
                sample=(f"one left move of the D={D} chi={chi} step in the reference op order (R1, R2, "
                        f"M = R1 R2, full LAPACK gesdd of the {chi*D*D}^2 M, P1/P2, proj_left with the "
                        f"chi^2 D^6 edge intermediate), {t:.1f} s on {threads} threads; step+adjoint "
                        f"extrapolated as 4 moves x 3"),
                move_seconds=t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--mixed", type=int, default=3, choices=[1, 3, 6])
    ap.add_argument("--oversample", type=int, default=16)
    ap.add_argument("--power-iters", type=int, default=1)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent instances shard with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[a.workload]
    D, chi, d = w["D"], w["chi"], w["d"]
    pkg = load_pkg()
    M = pkg.models
    A = M.seeded_tensor((D, D, D, D, d), w["seed"], unit_max=w["unit_max"])
    H = M.heisenberg(1.0)
    ctx = pkg.default_context()
    base_opts = dict(svd_mode=1, oversample=a.oversample, power_iters=a.power_iters)
    sat_steps = 2  # chi saturates after the first step (N = D^4 >= chi)

    def run_path(dtype, n_warm, n_time, extra):
        ts = pkg.CTMTensors(A, {}, dtype=dtype, options=dict(base_opts, **extra))
        for _ in range(sat_steps):
            pkg.rg_step(ts, chi)
        E = g = None
        for _ in range(n_warm):
            E, g = pkg.step_adjoint(ts, chi, H)
        ctx.sync()
        ctx.set_timing(True)
        ctx.kernel_stats(reset=True)
        t0 = time.perf_counter()
        for _ in range(n_time):
            E, g = pkg.step_adjoint(ts, chi, H)
        ctx.sync()
        t = time.perf_counter() - t0
        stats = ctx.kernel_stats(reset=True)
        ctx.set_timing(False)
        return t, E, g, stats, ts

    log(f"[bench] workload {a.workload}: D={D} chi={chi} N={chi*D*D}; mixed bf16x{a.mixed}")
    t_m, E_m, g_m, st_m, ts_m = run_path("mixed", a.warmup, a.steps, dict(mixed=a.mixed))
    log(f"[bench] mixed: {a.steps} steps in {t_m:.3f} s, E={E_m.real:.12f}")
    out_extra = {}
    if not a.no_fp64:
        import numpy as np
        t_64, E_64, g_64, st_64, ts_64 = run_path("f64", a.warmup, a.steps, {})
        log(f"[bench] fp64 : {a.steps} steps in {t_64:.3f} s, E={E_64.real:.12f}")
        # Precision error of one step + adjoint: both engines start from the
        # same FP64 environment (the FP64 run's last state) with the same
        # random test matrices, so the difference is the mixed arithmetic.
        Cs, Es = ts_64.Cs, ts_64.Es
        same = {}
        for dt, extra in (("f64", {}), ("mixed", dict(mixed=a.mixed))):
            ts = pkg.CTMTensors(A, {}, dtype=dt, options=dict(base_opts, **extra))
            ts.set_env(Cs, Es)
            same[dt] = pkg.step_adjoint(ts, chi, H)
        dE = abs(same["mixed"][0].real - same["f64"][0].real)
        gerr = float(np.linalg.norm(same["mixed"][1] - same["f64"][1]) / np.linalg.norm(same["f64"][1]))
        log(f"[bench] same-input step+adjoint: |dE/site| mixed vs fp64 = {dE:.2e}, grad rel = {gerr:.2e}")
        out_extra["fp64_gpu"] = dict(value=a.steps / t_64, ms_per_step=1e3 * t_64 / a.steps,
                                     roofline=kernel_roofline(st_64))
        out_extra["speedup_vs_fp64_gpu"] = t_64 / t_m
        out_extra["dE_per_site_vs_fp64"] = dE
        out_extra["grad_rel_err_vs_fp64"] = gerr
    for k, v in sorted(st_m.items(), key=lambda kv: -kv[1][1])[:12]:
        log(f"[bench]   {k:48s} n={v[0]:5d} {v[1]:9.3f} ms")
    roof = kernel_roofline(st_m)
    value = a.steps / t_m
    F = flops_per_unit(D, chi, d, a.oversample, a.power_iters)
    cpu = None
    if not a.no_cpu and a.workload in ("c3", "c2", "small"):
        log("[bench] cpu baseline sample ...")
        cpu = cpu_baseline_sample(D, chi, w["seed"])
        log(f"[bench] cpu: {cpu['sample']}")
    res = {
        "metric": "CTMRG step+grad/s at D=6 chi=144, mixed vs FP64 GPU, dE/site <= 1e-7",
        "value": value,
        "unit": "step+grad/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * t_m / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"bf16x{a.mixed}-fp32acc+f64",
        "data": "synthetic",
        "config": {"workload": f"CTMRG rg_step + energy VJP + step adjoint, Heisenberg, A (D,D,D,D,d)="
                               f"({D},{D},{D},{D},{d}) randn seed {w['seed']}, chi={chi}",
                   "D": D, "chi": chi, "N": chi * D * D, "svd": f"randomized k=chi+{a.oversample} q={a.power_iters}",
                   "mixed_products": a.mixed},
        "algorithmic_flops_per_step": F,
        "model_flops_rate_tflops": F * value / 1e12,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    res.update(out_extra)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
