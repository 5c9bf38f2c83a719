"""Batched independent environments (BatchedCTM / ipeps_batch_*): the uneven
64-instance workload, per-instance agreement with separate single-environment
runs and with FP64, and the worker pool's error handling."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _subset(pkg, idx):
    ins = pkg.c4_instances()
    return [ins[i] for i in idx]


def test_batch_matches_separate_runs(pkg):
    """Six uneven instances (D 2..4, chi 16..96, both models) advanced 4 steps
    by 3 workers agree with the same instances run one by one through
    CTMTensors / rg_step / get_gs_energy.  Same kernels and the same random
    test matrices (per-environment stream); the only differences are the
    run-to-run rounding of the rocSOLVER reductions inside the projector,
    amplified over the steps: |dE| <= 1e-8 (FP64) and 1e-6 (mixed)."""
    sel = _subset(pkg, [0, 4, 7, 16, 32, 52])
    As, chis, Hs = [s[0] for s in sel], [s[1] for s in sel], [s[2] for s in sel]
    for dtype, tol in (("f64", 1e-8), ("mixed", 1e-6)):
        b = pkg.BatchedCTM(As, chis, Hs, dtype=dtype, workers=3, options=dict(svd_mode=1))
        Eb, sec = b.run(4)
        assert sec > 0
        for i, (A, chi, H, _) in enumerate(sel):
            ts = pkg.CTMTensors(A, {}, dtype=dtype, options=dict(svd_mode=1))
            for _ in range(4):
                pkg.rg_step(ts, chi)
            E, _ = pkg.get_gs_energy(ts, H)
            print(dtype, i, Eb[i].real, E.real, abs(Eb[i] - E))
            assert abs(Eb[i] - E) < tol
            # the environment tensors come back with the instance's shapes
            C1 = b.tensor(i, 0)
            assert C1.shape == ts.Cs[0].shape
        b.close()


def test_batch_c4_workload_vs_fp64(pkg):
    """The 64-instance uneven workload (D in {2,3,4}, chi in {16..96}, Ising /
    Heisenberg, A float32 from seeds 3+i), 30 steps, mixed engine vs the FP64
    engine.  Per-instance energy per site: median |dE| <= 1e-5 and every
    instance <= 1e-3 (random tensors; the mixed engine truncates at its 1e-6
    noise floor, which shifts chi where the spectrum falls below it)."""
    ins = pkg.c4_instances()
    As, chis, Hs = [s[0] for s in ins], [s[1] for s in ins], [s[2] for s in ins]
    bm = pkg.BatchedCTM(As, chis, Hs, dtype="mixed", workers=8, options=dict(svd_mode=1))
    Em, sm = bm.run(30)
    b64 = pkg.BatchedCTM(As, chis, Hs, dtype="f64", workers=8, options=dict(svd_mode=1))
    E64, s64 = b64.run(30)
    d = np.abs(Em.real - E64.real)
    print(f"mixed {30 * len(ins) / sm:.1f} instance-steps/s, fp64 {30 * len(ins) / s64:.1f}; "
          f"|dE| median {np.median(d):.2e} max {d.max():.2e}")
    assert np.all(np.isfinite(Em)) and np.all(np.isfinite(E64))
    assert np.median(d) < 1e-5
    assert d.max() < 1e-3


def test_batch_rejects_bad_arguments(pkg):
    A = np.ones((2, 2, 2, 2, 2))
    H = pkg.models.ising(2.0)
    with pytest.raises(pkg.IpepsError):
        pkg.BatchedCTM([A], [8], [H], workers=0)
    with pytest.raises(pkg.IpepsError):
        pkg.BatchedCTM([np.ones((2, 3, 2, 2, 2))], [8], [H])
    with pytest.raises(ValueError):
        pkg.BatchedCTM([A, A], [8], [H])
