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
# GEMM plans tuned on an MI355X for this workload (bench.py --tune 1 --save-tune):
# replaying them makes the arithmetic, and so the mixed-vs-FP64 differences,
# the same from run to run.
TUNE_TABLE = os.path.join(ROOT, "ipeps.jl_amd", "tune_tables", "gfx950_ctmrg_D6_chi144.txt")
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


LIBRARY_PREFIXES = ("rocsolver_", "rocblas_")


def dominant_kernel(stats):
    """Hand-written throughput kernel with the largest total time.  Library
    calls carry no algorithmic counts, and the one-workgroup FP64
    factorisations (Cholesky + inverse, tridiagonalisation: kernels that
    report no algorithmic bytes) are latency chains with no roofline; both
    are reported separately (library_ms, latency_bound)."""
    own = {k: v for k, v in stats.items() if not k.startswith(LIBRARY_PREFIXES) and v[3] > 0}
    return max(own.items(), key=lambda kv: kv[1][1])[0] if own else None


def latency_bound(stats, top=4):
    """The hand-written kernels without algorithmic bytes: {name: [launches, ms]}."""
    lat = {k: v for k, v in stats.items() if not k.startswith(LIBRARY_PREFIXES) and v[3] <= 0}
    return {k: [v[0], round(v[1], 3)] for k, v in sorted(lat.items(), key=lambda kv: -kv[1][1])[:top]}


def kernel_roofline(stats, name=None):
    """Achieved rate of one kernel (default: the dominant hand-written one)
    from its algorithmic flops / bytes per launch and its average launch time."""
    name = name or dominant_kernel(stats)
    launches, ms, flops, bytes_ = stats[name]
    avg_ms = ms / max(1, launches)
    mfma_peak = None
    if name.startswith(("gemm_f32mfma", "gemm_lowk_f32", "absorb_")):
        mfma_peak = PEAK_TFLOPS["f32"]  # FP32 MFMA products
    elif name.startswith("gemm_f32"):
        # split-bf16: NSPLIT planes give NSPLIT (NSPLIT + 1) / 2 bf16 products per flop
        ns = int(name.split("<")[1].split(",")[2].strip(" >")) if name.count(",") >= 2 else 2
        mfma_peak = PEAK_TFLOPS["bf16"] / (ns * (ns + 1) // 2)
    elif name.startswith("gemm_"):
        mfma_peak = PEAK_TFLOPS["f64"]
    # The bound follows the launches' algorithmic arithmetic intensity: below
    # the ridge point (peak flops / peak HBM bandwidth) the kernel is HBM-bound.
    intensity = flops / bytes_ if bytes_ else float("inf")
    if mfma_peak is not None and intensity >= mfma_peak * 1e12 / (PEAK_HBM_GBS * 1e9):
        unit, peak, bound = "TFLOP/s", mfma_peak, "mfma"
        achieved = (flops / launches) / (avg_ms * 1e-3) / 1e12
    else:
        unit, peak, bound = "GB/s", PEAK_HBM_GBS, "hbm"
        achieved = (bytes_ / launches) / (avg_ms * 1e-3) / 1e9 if bytes_ else 0.0
    traffic, source = measured_traffic(name)
    return dict(kernel=name, bound=bound, achieved=achieved, peak=peak, unit=unit, frac=achieved / peak,
                traffic=traffic, traffic_source=source, algorithmic_bytes=bytes_ / max(1, launches),
                algorithmic_flops=flops / max(1, launches), intensity_flop_per_byte=intensity,
                mfma_peak_tflops=mfma_peak,
                mfma_tflops=(flops / launches) / (avg_ms * 1e-3) / 1e12 if mfma_peak else None,
                launches=launches, avg_ms=avg_ms, total_ms=ms)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summaries (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), as a
    line 'traffic_bytes_per_launch <kernel>: <bytes>' in profiles/*.md."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch ") and ":" in line:
                k, v = line[len("traffic_bytes_per_launch "):].rsplit(":", 1)
                if k.strip() == kernel:
                    return float(v), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline_sample(D, chi, seed):
    """The oracle (FP64, CPU) on one left move of the workload's step, forward
    and adjoint (torch autograd through the reference's svd_back), in the
    reference op order (M = R1 R2, full LAPACK SVD of the N x N M, the
    chi^2 D^6 edge intermediate) and in the canonical order (the edge update
    P2-first), plus the energy with its VJP; step + adjoint = 4 moves + 1
    energy.  Threads: OMP_NUM_THREADS if set (the CPU share of a GPU box),
    else every core of the process's affinity mask; both counts reported."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctmrg_oracle as O

    host_cores = len(os.sched_getaffinity(0))
    threads = min(host_cores, int(os.environ.get("OMP_NUM_THREADS") or host_cores))
    torch.set_num_threads(threads)
    A = O.seeded_tensor((D, D, D, D, 2), seed)
    H = O.heisenberg(1.0)
    g = torch.Generator().manual_seed(0)
    Cs = [torch.randn(chi, chi, generator=g, dtype=torch.float64) for _ in range(4)]
    Es = [torch.randn(chi, chi, D, D, generator=g, dtype=torch.float64) for _ in range(4)]
    w = [torch.randn(chi, chi, generator=g, dtype=torch.float64), torch.randn(chi, chi, D, D, generator=g,
                                                                              dtype=torch.float64),
         torch.randn(chi, chi, generator=g, dtype=torch.float64)]

    def left_move(order):
        O.EDGE_ORDER = order
        try:
            x = A.clone().requires_grad_(True)
            ts = O.CTMTensors(x, x.conj(), Cs, Es, dict(chi=chi))
            t0 = time.perf_counter()
            ts2, _ = O.left_rg(ts, chi)
            loss = (w[0] * ts2.Cs[0]).sum() + (w[1] * ts2.Es[3]).sum() + (w[2] * ts2.Cs[3]).sum()
            t1 = time.perf_counter()
            loss.backward()
            t2 = time.perf_counter()
        finally:
            O.EDGE_ORDER = "reference"
        return t1 - t0, t2 - t1

    ref_f, ref_b = left_move("reference")
    can_f, can_b = left_move("canonical")
    x = A.clone().requires_grad_(True)
    ts = O.CTMTensors(x, x.conj(), Cs, Es, dict(chi=chi))
    t0 = time.perf_counter()
    E = O.get_gs_energy(ts, H)[0].real
    t1 = time.perf_counter()
    E.backward()
    t2 = time.perf_counter()
    en_f, en_b = t1 - t0, t2 - t1
    unit_ref = 4 * (ref_f + ref_b) + en_f + en_b
    unit_can = 4 * (can_f + can_b) + en_f + en_b
    return dict(value=1.0 / unit_ref, unit="step+grad/s", cores=threads, host_cores=host_cores, kind="port",
                sample=(f"oracle, FP64, {threads} threads ({host_cores} cores in the affinity mask): one left move "
                        f"of the D={D} chi={chi} step forward + adjoint in the reference op order (full LAPACK SVD "
                        f"of the {chi*D*D}^2 M, chi^2 D^6 edge intermediate) {ref_f:.1f} + {ref_b:.1f} s, canonical "
                        f"order {can_f:.1f} + {can_b:.1f} s; energy + VJP {en_f:.1f} + {en_b:.1f} s; step + adjoint "
                        f"= 4 moves + 1 energy"),
                reference_order=dict(move_forward_s=ref_f, move_adjoint_s=ref_b, unit_s=unit_ref,
                                     value=1.0 / unit_ref),
                canonical_order=dict(move_forward_s=can_f, move_adjoint_s=can_b, unit_s=unit_can,
                                     value=1.0 / unit_can),
                energy_forward_s=en_f, energy_adjoint_s=en_b)


def batch_c4(pkg, steps=10, workers=16):
    """Uneven 64-instance workload (D in {2,3,4}, chi in {16..96}): mixed and
    FP64 engines advance all instances `steps` rg_steps after one warm-up
    step, eagerly and in graph mode (each instance's step captured into a HIP
    graph once its shapes settle, then replayed; captures happen inside the
    timed steps); aggregate instance-steps/s and |dE/site| mixed vs FP64
    (eager FP64 as the reference)."""
    import numpy as np
    ins = pkg.c4_instances()
    As, chis, Hs = [x[0] for x in ins], [x[1] for x in ins], [x[2] for x in ins]
    out = {}
    Es = {}
    for dt, graph in (("mixed", 1), ("mixed", 0), ("f64", 1), ("f64", 0)):
        key = dt + ("_graph" if graph else "")
        b = pkg.BatchedCTM(As, chis, Hs, dtype=dt, workers=workers, options=dict(svd_mode=1, graph=graph))
        b.run(1)
        E, sec = b.run(steps)
        Es[key] = E
        out[key] = dict(instance_steps_per_s=steps * len(ins) / sec, seconds=sec)
        if graph:
            out[key]["graph_counters"] = b.stats()
        b.close()
        log(f"[bench] c4 batch {key}: {steps} steps x {len(ins)} instances in {sec:.2f} s"
            + (f" {out[key]['graph_counters']}" if graph else ""))
    for key in ("mixed", "mixed_graph", "f64_graph"):
        d = np.abs(Es[key].real - Es["f64"].real)
        out[key].update(dE_per_site_vs_f64_median=float(np.median(d)), dE_per_site_vs_f64_max=float(d.max()))
    out.update(instances=len(ins), steps=1 + steps, workers=workers)
    return out


C5_MOMENTA = [(1.0, 0.0), (1.0, 1.0), (0.5, 0.5), (0.0, 0.0), (1.0, 0.0), (0.5, 0.5)]  # units of pi


def excited_c5(pkg, args):
    """Config 5: Ising h=2, A (3,3,3,3,2) from default_rng(4) optimised in FP64
    (optim_gs, LBFGS), chi = 48; prepare_basis gives D^4 d - 1 = 161 tangent
    vectors; at each momentum every column is a nested CTMRG + adjoint on the
    GPU (optim_es on `workers` contexts) and compute_es gives the lowest 4
    bands.  Every column and momentum takes its projectors from one shared
    path (the plain parts of the nested runs do not depend on B or the
    momentum), made once per projector setting.  The production projector
    (randomized, k = chi + 300, q = 8: made once, its cost is that of a few
    columns) is compared with the full SVD
    (svd_mode=0, reference semantics) at the first --c5-ref momenta of
    (pi/2, pi/2), (pi, 0), (pi, pi), (0, 0)."""
    import numpy as np
    M = pkg.models
    H = M.ising(2.0)
    cfg = dict(chi=48, min_iter=4, max_iter=20, ad_max_iter=4, rg_tol=1e-6)
    A0 = M.seeded_tensor((3, 3, 3, 3, 2), 4)
    t0 = time.perf_counter()
    r = pkg.optim_gs(H, A0, cfg, iterations=args.c5_gs_iters, dtype="f64",
                     options=dict(svd_mode=1, oversample=48, power_iters=2))
    t_gs = time.perf_counter() - t0
    log(f"[c5] ground state: {r['iterations']} LBFGS iterations, E = {r['energy']:.12f}, {t_gs:.1f} s")
    prod = dict(svd_mode=1, oversample=args.c5_oversample, power_iters=args.c5_power_iters)
    t0 = time.perf_counter()
    basis, ts, Hs = pkg.excited.prepare_basis(r["A"].astype(np.complex128), H, cfg, dtype="c128", options=prod)
    log(f"[c5] basis {basis.shape} in {time.perf_counter() - t0:.1f} s")
    # the nested runs' stopping rule (rg_tol) may be set apart from the ground state's
    cfg = dict(cfg, rg_tol=args.c5_rg_tol, max_iter=args.c5_max_iter)
    momenta = C5_MOMENTA[: args.c5_momenta]

    def progress(done, total):
        if done % 20 == 0 or done == total:
            log(f"[c5]   {done}/{total} columns")

    t0 = time.perf_counter()
    bands = pkg.excited.band_structure(ts, Hs, basis, momenta, cfg, workers=args.c5_workers, progress=progress)
    t_bands = time.perf_counter() - t0
    for b in bands:
        log(f"[c5] ({b['px']}, {b['py']}) pi: bands {np.round(b['bands'], 8).tolist()} in {b['seconds']:.1f} s")
    ref = []
    if args.c5_ref:
        # reference semantics: the full SVD of M (svd_mode=0)
        ref_opts = dict(svd_mode=0)
        ts_ref = pkg.CTMTensors(ts.A, cfg, dtype="c128", options=ref_opts)
        ts_ref.set_env(ts.Cs, ts.Es)
        order = [(0.5, 0.5), (1.0, 0.0), (1.0, 1.0), (0.0, 0.0)]
        idx = [momenta.index(m) for m in order if m in momenta][: args.c5_ref]
        t0 = time.perf_counter()
        ref_path = pkg.excited.ProjectorPath(ts_ref, cfg)
        log(f"[c5] full-SVD projector path ({ref_path.moves} moves) in {time.perf_counter() - t0:.1f} s")
        for i in idx:
            rb = pkg.excited.band_structure(ts_ref, Hs, basis, [momenta[i]], cfg, workers=args.c5_workers,
                                            progress=progress, path=ref_path)[0]
            b = bands[i]
            err = float(np.abs(rb["bands"] - b["bands"]).max())
            # agreement of the columns themselves, the norm-space dimension compute_es keeps, and the band
            # change that a Hermitian perturbation of the reference's effH / effN of the measured size causes
            eH = float(np.abs(rb["effH"] - b["effH"]).max() / np.abs(rb["effH"]).max())
            eN = float(np.abs(rb["effN"] - b["effN"]).max() / np.abs(rb["effN"]).max())
            rng = np.random.default_rng(0)

            def perturbed(X, rel):
                Z = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
                Z = (Z + Z.conj().T) / 2
                return X + Z * (rel * np.abs(X).max() / np.abs(Z).max())

            es_p, _, _ = pkg.excited.compute_es(perturbed(rb["effH"], eH), perturbed(rb["effN"], eN))
            sens = float(np.abs(np.real(es_p[:4]) - rb["bands"]).max())
            nsel = [int(pkg.excited.compute_es(x["effH"], x["effN"])[2].shape[1]) for x in (b, rb)]
            ref.append(dict(px=rb["px"], py=rb["py"], bands_ref=rb["bands"].tolist(), max_band_err=err,
                            effH_rel_err=eH, effN_rel_err=eN, band_change_same_size_perturbation=sens,
                            kept_norm_dims=nsel, seconds=rb["seconds"], projector=ref_opts))
            log(f"[c5] reference at ({rb['px']}, {rb['py']}) pi: {np.round(rb['bands'], 8).tolist()}, "
                f"max |band error| {err:.2e}, effH / effN rel {eH:.1e} / {eN:.1e}, band change under a "
                f"perturbation of that size {sens:.1e}, kept norm dims {nsel}, {rb['seconds']:.1f} s")
        ref_path.close()
    ncol = basis.shape[1] * len(momenta)
    print(json.dumps({
        "metric": "config-5 excited-state columns/s (nested CTMRG + adjoint per tangent vector)",
        "value": ncol / t_bands, "unit": "columns/s", "n_gpus": 1, "higher_is_better": True,
        "config": {"workload": "Ising h=2, A (3,3,3,3,2) seed 4 optimised in FP64, chi=48, 161 tangent vectors",
                   "momenta_pi": momenta, "workers": args.c5_workers, "projector": prod},
        "nested_cfg": cfg, "seconds": {"ground_state": t_gs, "bands": t_bands}, "columns": ncol,
        "gs_energy": r["energy"], "gs_iterations": r["iterations"],
        "bands": [dict(px=b["px"], py=b["py"], lowest=b["bands"].tolist()) for b in bands],
        "reference": ref}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3", choices=sorted(WORKLOADS))
    ap.add_argument("--mixed", type=int, default=3, choices=[1, 3, 6])
    ap.add_argument("--mixed-back", type=int, default=0, choices=[0, 1, 3, 6],
                    help="bf16 products of the VJPs (0: same as --mixed)")
    ap.add_argument("--mixed-proj", type=int, default=6, choices=[0, 1, 3, 6],
                    help="bf16 products inside the projector (0: same as --mixed; 6: FP32 products)")
    ap.add_argument("--early-mixed", type=int, default=1, choices=[0, 1, 3, 6],
                    help="products of the range finder's early applications of M (all but the last --late-apps "
                         "before B = M^H Q; 0: same as --mixed-proj)")
    ap.add_argument("--late-apps", type=int, default=1)
    ap.add_argument("--tune", type=int, default=2, choices=[0, 1, 2, 3],
                    help="GEMM launch plans: 0 shape rules, 1 autotune now (unseen shapes are timed during the "
                         "warm-up steps), 2 replay the tuning table --tune-file (shapes missing from it: rules), "
                         "3 replay the table and autotune the shapes it lacks (with --save-tune: extend it)")
    ap.add_argument("--tune-file", default=TUNE_TABLE)
    ap.add_argument("--save-tune", default=None, help="write the tuning table after the run")
    # k = chi + 144 = 2 chi and 4 power iterations resolve the kept subspace of
    # this workload's flat spectrum at the cut (FP64 energy within ~1e-10 of
    # the full SVD's, independent of the test matrices)
    ap.add_argument("--oversample", type=int, default=144)
    ap.add_argument("--power-iters", type=int, default=4)
    ap.add_argument("--complement-iters", type=int, default=0,
                    help="CG iterations of the adjoint's discarded-subspace term (0: its thin-SVD limit)")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--shape-stats", default=None,
                    help="after the timed mixed steps, time one more step + adjoint per kernel and GEMM shape "
                         "(HIP events) and write {name: [launches, ms, flops, bytes]} to this JSON file")
    ap.add_argument("--c5", action="store_true",
                    help="run the config-5 excited-state band structure instead (one JSON line of its own)")
    ap.add_argument("--c5-momenta", type=int, default=6)
    ap.add_argument("--c5-ref", type=int, default=1, help="momenta compared with the full-SVD engine")
    ap.add_argument("--c5-workers", type=int, default=8)
    ap.add_argument("--c5-gs-iters", type=int, default=40)
    ap.add_argument("--c5-oversample", type=int, default=300)
    ap.add_argument("--c5-power-iters", type=int, default=8)
    ap.add_argument("--c5-rg-tol", type=float, default=1e-6, help="nested runs' stopping tolerance")
    ap.add_argument("--c5-max-iter", type=int, default=20, help="nested runs' step limit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(independent instances shard with no exchange; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if a.c5:
        excited_c5(load_pkg(), a)
        return
    w = WORKLOADS[a.workload]
    D, chi, d = w["D"], w["chi"], w["d"]
    pkg = load_pkg()
    M = pkg.models
    A = M.seeded_tensor((D, D, D, D, d), w["seed"], unit_max=w["unit_max"])
    H = M.heisenberg(1.0)
    ctx = pkg.default_context()
    base_opts = dict(svd_mode=1, oversample=a.oversample, power_iters=a.power_iters,
                     complement_iters=a.complement_iters)
    sat_steps = 2  # chi saturates after the first step (N = D^4 >= chi)

    def run_path(dtype, n_warm, n_time, extra):
        """Warm-up steps (the last one with every launch timed, to find the
        kernel mix and the dominant hand-written kernel), then the timed steps
        with HIP events around the dominant kernel's launches only."""
        ts = pkg.CTMTensors(A, {}, dtype=dtype, options=dict(base_opts, **extra))
        for _ in range(sat_steps):
            pkg.rg_step(ts, chi)
        E = g = None
        # while tuning, one more untimed step tunes the adjoint's shapes
        n_warm = max(1, n_warm) + (1 if a.tune == 1 else 0)
        for i in range(n_warm):
            if i == n_warm - 1:
                ctx.sync()
                ctx.set_timing(True)
                ctx.kernel_stats(reset=True)
            E, g = pkg.step_adjoint(ts, chi, H)
        ctx.sync()
        mix = ctx.kernel_stats(reset=True)
        dom = dominant_kernel(mix)
        ctx.set_timing(True, prefix=dom)
        t0 = time.perf_counter()
        for _ in range(n_time):
            E, g = pkg.step_adjoint(ts, chi, H)
        ctx.sync()
        t = time.perf_counter() - t0
        stats = ctx.kernel_stats(reset=True)
        ctx.set_timing(False)
        stats = {k: v for k, v in stats.items() if k == dom}
        if a.shape_stats and dtype == "mixed":
            # one more unit with every launch timed per kernel and GEMM shape
            ctx.set_timing(2)
            pkg.step_adjoint(ts, chi, H)
            ctx.sync()
            with open(a.shape_stats, "w") as f:
                json.dump(ctx.kernel_stats(reset=True), f)
            ctx.set_timing(False)
        return t, E, g, stats, ts, mix

    mixed_opts = dict(mixed=a.mixed, mixed_back=a.mixed_back, mixed_proj=a.mixed_proj, early_mixed=a.early_mixed,
                      late_apps=a.late_apps)
    if a.tune == 2 and not os.path.exists(a.tune_file):
        log(f"[bench] no tuning table at {a.tune_file}: tuning now")
        a.tune = 1
    if a.tune == 2:
        pkg.tune_load(a.tune_file)
        pkg.tune(2)
    elif a.tune == 3:
        # the committed table, and the shapes it lacks tuned in the run
        pkg.tune_load(a.tune_file)
        pkg.tune(1, 1e8)
    elif a.tune == 1:
        pkg.tune(1, 1e8)
    log(f"[bench] workload {a.workload}: D={D} chi={chi} N={chi*D*D}; mixed {mixed_opts}")
    t_m, E_m, g_m, st_m, ts_m, mix_m = run_path("mixed", a.warmup, a.steps, mixed_opts)
    log(f"[bench] mixed: {a.steps} steps in {t_m:.3f} s, E={E_m.real:.12f}")
    out_extra = {}
    if not a.no_fp64:
        import numpy as np
        t_64, E_64, g_64, st_64, ts_64, _ = run_path("f64", a.warmup, a.steps, {})
        log(f"[bench] fp64 : {a.steps} steps in {t_64:.3f} s, E={E_64.real:.12f}")
        # Precision error of one step + adjoint: both engines start from the
        # same FP64 environment (the FP64 run's last state) with the same
        # random test matrices, so the difference is the mixed arithmetic.
        Cs, Es = ts_64.Cs, ts_64.Es
        same = {}
        for dt, extra in (("f64", {}), ("mixed", mixed_opts)):
            ts = pkg.CTMTensors(A, {}, dtype=dt, options=dict(base_opts, **extra))
            ts.set_env(Cs, Es)
            same[dt] = pkg.step_adjoint(ts, chi, H)
        # the method's own spread: FP64 with another seed of the random test matrices
        ts = pkg.CTMTensors(A, {}, dtype="f64", options=dict(base_opts, seed=98765))
        ts.set_env(Cs, Es)
        same["f64_seed"] = pkg.step_adjoint(ts, chi, H)
        dE = abs(same["mixed"][0].real - same["f64"][0].real)
        gnorm = np.linalg.norm(same["f64"][1])
        gerr = float(np.linalg.norm(same["mixed"][1] - same["f64"][1]) / gnorm)
        dE_seed = abs(same["f64_seed"][0].real - same["f64"][0].real)
        gerr_seed = float(np.linalg.norm(same["f64_seed"][1] - same["f64"][1]) / gnorm)
        log(f"[bench] same-input step+adjoint: |dE/site| mixed vs fp64 = {dE:.2e}, grad rel = {gerr:.2e}; "
            f"fp64 vs fp64 with other test matrices: {dE_seed:.2e}, {gerr_seed:.2e}")
        out_extra["fp64_gpu"] = dict(value=a.steps / t_64, ms_per_step=1e3 * t_64 / a.steps,
                                     roofline=kernel_roofline(st_64))
        out_extra["speedup_vs_fp64_gpu"] = t_64 / t_m
        out_extra["dE_per_site_vs_fp64"] = dE
        out_extra["grad_rel_err_vs_fp64"] = gerr
        out_extra["fp64_test_matrix_spread"] = dict(dE_per_site=dE_seed, grad_rel=gerr_seed)
        # against a resolved FP64 projector (k = chi + 256, 8 power iterations,
        # the adjoint's discarded-subspace term by CG): at this state within
        # ~1e-14 in energy and ~3e-6 in gradient of the full-SVD reference
        # semantics, which take about a minute per step + adjoint here
        ts = pkg.CTMTensors(A, {}, dtype="f64", options=dict(svd_mode=1, oversample=256, power_iters=8,
                                                           complement_iters=60, complement_tol=1e-10))
        ts.set_env(Cs, Es)
        E_r, g_r = pkg.step_adjoint(ts, chi, H)
        gn_r = np.linalg.norm(g_r)
        out_extra["vs_fp64_resolved"] = {
            k: dict(dE_per_site=abs(same[k][0].real - E_r.real),
                    grad_rel=float(np.linalg.norm(same[k][1] - g_r) / gn_r)) for k in ("mixed", "f64")}
        log(f"[bench] against the resolved FP64 projector: mixed {out_extra['vs_fp64_resolved']['mixed']}, "
            f"fp64 {out_extra['vs_fp64_resolved']['f64']}")
    tot_mix = sum(v[1] for v in mix_m.values())
    log(f"[bench] kernel mix of one mixed step + adjoint (all launches timed): {tot_mix:.1f} ms")
    for k, v in sorted(mix_m.items(), key=lambda kv: -kv[1][1])[:14]:
        log(f"[bench]   {k:48s} n={v[0]:5d} {v[1]:9.3f} ms")
    lib_ms = sum(v[1] for k, v in mix_m.items() if k.startswith(LIBRARY_PREFIXES))
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
                   "D": D, "chi": chi, "N": chi * D * D,
                   "svd": f"randomized k=chi+{a.oversample} q={a.power_iters}"
                          + (f" complement CG {a.complement_iters}" if a.complement_iters else ""),
                   "mixed_products": a.mixed, "mixed_back": a.mixed_back or a.mixed,
                   "mixed_proj": a.mixed_proj or a.mixed,
                   "early_mixed": a.early_mixed, "late_apps": a.late_apps,
                   "gemm_plans": {0: "shape rules", 1: "autotuned in this run",
                                  2: "tuning table " + os.path.relpath(a.tune_file, ROOT),
                                  3: "tuning table " + os.path.relpath(a.tune_file, ROOT) + " + autotuned in this run"}[
                                      a.tune]},
        "algorithmic_flops_per_step": F,
        "model_flops_rate_tflops": F * value / 1e12,
        "roofline": roof,
        "cpu_baseline": cpu,
        "step_mix": {"timed_kernels_ms": tot_mix, "rocsolver_rocblas_ms": lib_ms, "latency_bound": latency_bound(mix_m),
                     "top": {k: round(v[1], 3) for k, v in sorted(mix_m.items(), key=lambda kv: -kv[1][1])[:8]}},
    }
    if a.tune:
        res["autotune"] = pkg.tune_stats()
    if a.save_tune:
        pkg.tune_save(a.save_tune)
    if not a.no_batch:
        res["batched_c4"] = batch_c4(pkg)
    res.update(out_extra)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
