/*
 * ipeps_hip -- C ABI of the MI355X (gfx950) CTMRG engine.
 *
 * Conventions
 *  - Tensors are dense, column-major (Julia layout), with explicit int64 dims.
 *  - dtype codes: IPEPS_F64 (double), IPEPS_C128 (interleaved re,im doubles =
 *    ComplexF64 = C double _Complex), IPEPS_MIXED (environment stored in FP32,
 *    contractions on split-bf16 MFMA with FP32 accumulation, FP64 projector
 *    SVD / orthogonalisation / energy reduction; host buffers are FP64).
 *  - Hamiltonian terms H1 (horizontal bond) and H2 (vertical bond) are
 *    hdim x hdim ComplexF64 matrices (hdim = d^2), site-1 index fastest as
 *    tout() builds them (src/model/basis.jl:16-21).
 *  - Host pointers are borrowed for the duration of a call only.  Device state
 *    lives behind the opaque handles and is freed by the destroy calls.
 *  - Every entry point returns 0 on success or a nonzero status; the message
 *    is available from ipeps_last_error() on the calling thread.
 *  - One HIP stream per context; calls on one handle are not re-entrant,
 *    separate contexts may be used from separate threads.
 *
 * Each entry point names the reference (yitan1/iPEPS.jl) function it
 * replaces.
 */
#ifndef IPEPS_HIP_H
#define IPEPS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IPEPS_F64 0
#define IPEPS_C128 1
#define IPEPS_MIXED 2

typedef struct ipeps_ctx ipeps_ctx;
typedef struct ipeps_env ipeps_env;
typedef struct ipeps_batch ipeps_batch;

/* Library version (major*10000 + minor*100 + patch). */
int ipeps_version(void);
/* Message of the last failed call on this thread ("" if none). */
const char* ipeps_last_error(void);

/* Context: device, stream, caching allocator, solver handle.  Environments
 * share ownership of the context they were created in: ipeps_ctx_destroy
 * releases the handle, and the device resources are freed once the last
 * environment made from it is destroyed too (handles may go in any order). */
int ipeps_ctx_create(int device, ipeps_ctx** out);
int ipeps_ctx_destroy(ipeps_ctx* ctx);
int ipeps_ctx_sync(ipeps_ctx* ctx);
/* Live / peak device bytes of the context allocator. */
int ipeps_ctx_mem(ipeps_ctx* ctx, int64_t* live, int64_t* peak);
/* Kernel timing: when on, every GEMM / permute / solver launch is bracketed
 * by hipEvents on the context stream.  Engine instrumentation, no reference
 * counterpart. */
int ipeps_ctx_set_timing(ipeps_ctx* ctx, int on);
/* Restrict the timing brackets to launches whose kernel name starts with
 * prefix (NULL or "" = every launch): a measured region then carries events
 * around the kernel of interest only. */
int ipeps_ctx_set_timing_filter(ipeps_ctx* ctx, const char* prefix);
/* Per-kernel totals since the last reset as JSON
 * {"name": [launches, total_ms, algorithmic_flops, algorithmic_bytes], ...};
 * reset != 0 clears them. */
int ipeps_ctx_kernel_stats(ipeps_ctx* ctx, char* buf, int64_t cap, int reset);

/* GEMM autotuner (process-wide; engine tuning, no reference counterpart:
 * the reference's contraction order and kernels are OMEinsum's choice,
 * src/utility/wrapper.jl:23).  mode 0 = tile and split-K by shape rules
 * (default), 1 = the first product of each unseen shape with at least
 * min_flops times every candidate plan (FP32 block tiles of 64 / 128 / 192
 * columns x split-K slice counts) on the caller's stream and keeps the
 * fastest, 2 = use loaded / tuned entries only.  stats: entries, shapes
 * tuned, table hits.  save / load: a text table (one "key tile_n splitk" per
 * line) so a later run replays the same plans and hence the same arithmetic. */
int ipeps_tune_config(int mode, double min_flops);
int ipeps_tune_stats(int64_t out[3]);
int ipeps_tune_clear(void);
int ipeps_tune_save(const char* path);
int ipeps_tune_load(const char* path);

/* CTMTensors(A, cfg) -> init_ctm1 (src/ctm/ctm_tensor.jl:63-69, 97-130).
 * dims = (D, D, D, D, d); A is F64 or C128 host data (MIXED takes F64). */
int ipeps_env_create(ipeps_ctx* ctx, int dtype, const int64_t dims[5], const void* A, ipeps_env** out);
int ipeps_env_destroy(ipeps_env* env);
/* Engine options: "svd_mode" (0 full SVD of M = R1 R2 as the reference,
 * 1 randomized range finder), "oversample", "power_iters", "full_below",
 * "seed", "eta", "mixed_cutoff" (relative singular-value cutoff of MIXED
 * environments, default 1e-6; F64 / C128 keep the reference's 1e-12),
 * "canonical_edge" (0/1), "mixed" (1 or 3 bf16 products per FP32 product, or
 * 6 = FP32 products on FP32 MFMA),
 * "mixed_back" (0, 1, 3 or 6: bf16 products of the step's VJPs; default
 * 0 = "mixed"), "mixed_proj" (0, 1, 3 or 6: bf16 products inside the
 * projector -- range finder, P1 and P2 -- default 0 = "mixed"),
 * "energy_fp64" (0/1: energy contractions of a MIXED environment in FP64),
 * "energy_mixed" (1, 3 or 6: bf16 products of those contractions when not
 * FP64, default 6; the density-matrix traces and ratios are FP64 always),
 * "energy_mixed_back" (0, 1, 3 or 6: bf16 products of the VJP of those
 * contractions; default 0 = the step's VJP setting),
 * "checkpoint" (0/1), "paired_moves" (0/1, default 1: the left and right
 * moves, then the top and bottom moves, of rg_step run concurrently on two
 * streams; results are identical either way), "fused_absorb" (0/1, default 1: MIXED environments
 * with 4 <= D <= 6 and d <= 2 absorb A and Ad into the enlarged corners in
 * one FP32-MFMA kernel instead of two split-bf16 GEMMs), and for the
 * randomized mode
 * "fast_orth" (0: Householder QR range bases; 1: CholeskyQR2 with Householder
 * fallback; 2, the default: CholeskyQR2 on the one-workgroup blocked FP64
 * Cholesky + inverse when k <= 160, on its 2 x 2 block split when k <= 320,
 * else as 1),
 * "fast_svd" (0/1: small SVD from Gram eigenpairs with one-sided Jacobi
 * fallback, for k <= 320 the eigenpairs by device tridiagonalisation,
 * bisection and inverse iteration; 2: the Gram eigenpairs from the
 * one-workgroup Jacobi kernel, k <= 112; 3: as 1 with rocSOLVER stedc on the
 * tridiagonal form), "early_mixed" / "late_apps" (MIXED: products of all but the
 * last late_apps applications of M before B = M^H Q at early_mixed = 1, 3
 * or 6 instead of the projector's setting; 0 = off), "graph_safe" (0/1: each
 * rg_step in the form a captured graph runs -- no host round trips,
 * truncation at the largest rank the cutoff rule admits, the small SVD's
 * Gram eigenpairs by the device tridiagonalisation, bisection and inverse
 * iteration (k <= 320), or the Jacobi kernel with fast_svd=2 --
 * with its deferred checks validated after the step, an error if one fails;
 * needs paired_moves=0), "gram_guard" (s_min/s_max below which the Gram route is not
 * taken, default 1e-4), "power_iters_max" and "resid_tol" (adaptive power
 * iterations: after "power_iters", while the a-posteriori residual
 * max_i ||M v_i - s_i u_i|| / s_1 of the kept triplets exceeds resid_tol and
 * fewer than power_iters_max have run, one more; resid_tol 0 = off), and
 * "complement_iters" / "complement_tol" (the adjoint's couplings between the
 * kept triplets and those beyond the computed k: 0 = their s_j -> 0 limit,
 * the thin-SVD terms; > 0 = solved exactly by conjugate gradients, at most
 * that many iterations, to the relative residual complement_tol, default
 * 1e-8), and "warm_start" (0/1, default 0: each move's range finder starts
 * from that move's previous left Ritz basis, kept in the environment, instead
 * of a random test matrix; the power iterations then refine it).  "draws" sets the position in the environment's
 * stream of random test matrices (keyed by "seed"; it advances by one per
 * randomized projector): environments with equal seed, draws and tensors
 * build identical projectors whatever their dtype. */
int ipeps_env_set_option(ipeps_env* env, const char* key, double value);
/* The environment's current position in that stream (the value "draws"
 * sets), for a copy that must draw the same test matrices: the field the
 * Julia-side copy of CTMTensors would carry (src/ctm/ctm_tensor.jl:3-18 has
 * no counterpart; its projectors are deterministic). */
int ipeps_env_get_draws(ipeps_env* env, int64_t* out);
/* Counters of the environment's randomized projectors since the last reset
 * (engine instrumentation; the reference's projectors are exact SVDs,
 * src/ctm/ctmrg.jl:254-265): out = [projectors built, power iterations
 * (sum), power iterations (largest in one projector), largest final
 * a-posteriori residual / s_1, checks made, discarded-subspace solves in the
 * adjoint, their CG iterations (sum), largest final relative CG residual,
 * solves abandoned for a non-positive-definite operator (thin terms used),
 * range-basis orthonormalisations, those that took CholeskyQR2's second
 * pass, those that fell back to Householder QR, projectors warm-started,
 * tridiagonal eigenpairs redone by rocSOLVER stedc after the inverse-iteration
 * vectors failed their orthogonality check].
 * reset != 0 clears them after the read. */
int ipeps_env_projector_stats(ipeps_env* env, double out[14], int reset);
/* Copy the environment state of src into dst (same A shape): C, E
 * (converted to dst's dtype), the position in the test-matrix stream and the
 * warm-start bases, so dst continues exactly as src would -- the state a
 * copied CTMTensors carries (src/ctm/ctm_tensor.jl:3-18 has no projector
 * state; its projectors are exact SVDs).  A is not copied. */
int ipeps_env_copy_state(ipeps_env* dst, const ipeps_env* src);
/* which: 0..3 = C1..C4, 4..7 = E1..E4, 8 = A.  dims has room for 5. */
int ipeps_env_tensor_dims(ipeps_env* env, int which, int64_t dims[5], int* nd);
/* Copy out / in (host data in the env's host dtype: F64 or C128). */
int ipeps_env_get_tensor(ipeps_env* env, int which, void* host);
int ipeps_env_set_tensor(ipeps_env* env, int which, int nd, const int64_t* dims, const void* host);
/* setproperties(ts, A=A, Ad=conj(A)) keeping the environment
 * (src/optimize/optim_gs.jl:134). */
int ipeps_env_set_A(ipeps_env* env, const void* A);

/* rg_step (src/ctm/ctmrg.jl:57-64): one left/right/top/bottom step; writes
 * the left move's truncated spectrum (s_len values, at most s_cap). */
int ipeps_rg_step(ipeps_env* env, int64_t chi, double* s_out, int64_t s_cap, int64_t* s_len);

/* run_ctm (src/ctm/ctmrg.jl:1-55) with its stopping rule; conv_mode 0 uses
 * get_gs_energy as conv_fun, 1 the left spectrum.  E_out = (re, im) of the
 * last conv energy; diffs_out receives the diff trace. */
int ipeps_run_ctm(ipeps_env* env, int64_t chi, int min_iter, int max_iter, double tol, int conv_mode,
                  const double* H1, const double* H2, int64_t hdim, int* iters, double* E_out,
                  double* diffs_out, int64_t diffs_cap, int64_t* n_diffs);

/* get_gs_energy (src/graph/evaluation.jl:15-29): E and Na as (re, im). */
int ipeps_gs_energy(ipeps_env* env, const double* H1, const double* H2, int64_t hdim, double* E_out,
                    double* N_out);
/* get_gs_energy_4x4 (src/graph/evaluation.jl:340-355: energy_norm_4x4
 * :443-496 with get_op_Ad4 :504-519) for a four-site plaquette Hamiltonian
 * in the factored form get_local_h returns (src/model/local_h.jl:1-31):
 * s1 (d,d,r1), s2 (r1,d,d,r2), s3 (r2,d,d,r3), s4 (r3,d,d), ComplexF64,
 * column-major; s_dims is 4 x 4 int64 (row i: the dims of s_i, trailing
 * entries ignored).  E_out = energy / norm, N_out = norm, each (re, im);
 * FP64 arithmetic whatever the environment's dtype. */
int ipeps_gs_energy_4x4(ipeps_env* env, const int64_t* s_dims, const double* s1, const double* s2,
                        const double* s3, const double* s4, double* E_out, double* N_out);
/* get_gs_norm (src/graph/evaluation.jl:47-59): (re, im). */
int ipeps_gs_norm(ipeps_env* env, double* out);

/* One CTMRG step plus its adjoint: forward rg_step, the energy and its VJP,
 * then the step VJP (the hand-written adjoint used by the gradient).  The
 * environment advances by one step.  gradA_host gets dRe(E)/dA (A's dtype;
 * MIXED returns F64). */
int ipeps_step_adjoint(ipeps_env* env, int64_t chi, const double* H1, const double* H2, int64_t hdim,
                       double* E_out, void* gradA_host);

/* fg! of optim_gs (src/optimize/optim_gs.jl:53-91, 120-143):
 * x = renormalize(A0); converge CTMRG from init_ctm1(x) (min/max_iter, tol,
 * energy conv); then ad_max_iter differentiated steps from the converged
 * environment held constant; E = get_gs_energy.  grad_out = dRe(E)/dRe(x) +
 * i dRe(E)/dIm(x) (A's dtype; MIXED returns F64).  options: "key=value;..."
 * as ipeps_env_set_option.  iters_out = (converging iterations, ad steps). */
int ipeps_gs_energy_grad(ipeps_ctx* ctx, int dtype, const int64_t dims[5], const void* A0, const double* H1,
                         const double* H2, int64_t hdim, int64_t chi, int min_iter, int max_iter,
                         int ad_max_iter, double tol, const char* options, double* E_out, void* grad_out,
                         int* iters_out);

/* run_gs (src/optimize/optim_gs.jl:133-143) and its pullback: the function
 * whose ChainRulesCore.rrule replaces Zygote's pass through the
 * differentiated CTMRG run of fg! (optim_gs.jl:72-82).  From env's current
 * environment, held constant (env is not modified), with A = x and
 * Ad = conj(x) (x: env's host dtype, A's dims): run_ctm's stopping rule with
 * min_iter, max_iter = ad_max_iter and tol on the energy, then
 * E = get_gs_energy (E_out = re, im).  grad_out = dE (dRe(E)/dRe(x) +
 * i dRe(E)/dIm(x)) for the cotangent dE of Re(E); steps_out = differentiated
 * steps.  The engine options of env apply (checkpoint=1: step
 * recomputation in the reverse sweep). */
int ipeps_run_gs_vjp(ipeps_env* env, const void* x, const double* H1, const double* H2, int64_t hdim, int64_t chi,
                     int min_iter, int ad_max_iter, double tol, double dE, double* E_out, void* grad_out,
                     int* steps_out);

/* get_es_grad (src/optimize/optim_es.jl:167-194) for one tangent tensor B:
 * env must be a converged, normalised ground-state environment in IPEPS_C128;
 * the momentum is (px, py) pi as optim_es's arguments (optim_es.jl:124-125).
 * Converges the nested environment [X, X_B, X_Bd, X_BB] from env with
 * conv = get_es_energy / get_all_norm[1] (min/max_iter, tol), differentiates
 * ad_max_iter nested steps + get_es_energy w.r.t. B and writes gH (dRe(E4)/dB),
 * envB (get_all_norm's B environment), E4 and Nb.  B, gH, envB: (D,D,D,D,d)
 * ComplexF64. */
int ipeps_es_grad(ipeps_env* env, const double* H1, const double* H2, int64_t hdim, double px, double py,
                  const void* B, int64_t chi, int min_iter, int max_iter, int ad_max_iter, double tol, void* gH_out,
                  void* envB_out, double* E4_out, double* Nb_out);

/* Projector path of nested runs, for optim_es (src/optimize/optim_es.jl:127-153),
 * whose columns all start from one environment.  A nested product's first
 * component involves only first components (src/nested_tensor/nested_tensor.jl:
 * 48-63), so the plain parts of every nested run from env evolve alike,
 * whatever B and the momentum, and move t (= 4 step + direction) of each run
 * uses the same projector (src/ctm/ctmrg.jl:122-265).  ipeps_es_path_create
 * computes the projectors of the first `steps` steps at bond dimension chi
 * once (a run with B = EmptyT, on env's context); ipeps_es_grad_path is
 * ipeps_es_grad with the moves the path covers taking their projectors from
 * it (later moves compute their own).  The path must have been made from an
 * environment in the same state as env -- equal tensors, test-matrix
 * position (ipeps_env_get_draws) and options -- and the same chi, else
 * IPEPS_ERR_ARG; environments on other contexts may use it (ipeps_env_copy_state
 * copies the state).  Destroy it after the last call that uses it has
 * returned. */
typedef struct ipeps_es_path ipeps_es_path;
int ipeps_es_path_create(ipeps_env* env, int64_t chi, int steps, ipeps_es_path** out);
int ipeps_es_path_destroy(ipeps_es_path* path);
int ipeps_es_path_moves(const ipeps_es_path* path, int64_t* moves);
int ipeps_es_grad_path(ipeps_env* env, const ipeps_es_path* path, const double* H1, const double* H2, int64_t hdim,
                       double px, double py, const void* B, int64_t chi, int min_iter, int max_iter,
                       int ad_max_iter, double tol, void* gH_out, void* envB_out, double* E4_out, double* Nb_out);

/* tcon (src/utility/wrapper.jl:15-23): ncon-style pairwise contraction on the
 * device.  Positive labels are summed, negative ones are output legs -1, -2,
 * ...  dimsC receives the output dims (room for 16), C_host the result
 * (capacity C_cap elements).  dtype IPEPS_F64 / IPEPS_C128, or IPEPS_MIXED
 * for float host data stored and multiplied in FP32 (the mixed engine's
 * storage; FP32 products). */
int ipeps_tcon(ipeps_ctx* ctx, int dtype, int ndA, const int64_t* dimsA, const void* A, const int* la, int ndB,
               const int64_t* dimsB, const void* B, const int* lb, int64_t* dimsC, int* ndC, void* C_host,
               int64_t C_cap);

/* The site absorption of an enlarged corner (two tcon calls of
 * get_projector_left etc., src/ctm/ctmrg.jl:139-140) as one contraction:
 * Y = X A Ad with ncon labels lx (6), la (5), lad (5): X has two open legs,
 * two legs shared with A and two with Ad; A and Ad share their physical leg.
 * dtype IPEPS_MIXED (float host data; fused != 0 runs the fused FP32-MFMA
 * kernel when 4 <= D <= 6 and d <= 2, else two split-bf16 contractions) or
 * IPEPS_F64 (two contractions).  Y_out receives Y (output legs -1..-6).  With
 * a cotangent G (Y's shape) the VJP is written to gX_out / gA_out / gAd_out
 * (each may be NULL). */
int ipeps_absorb(ipeps_ctx* ctx, int dtype, int fused, const int64_t* dimsX, const void* X, const int* lx,
                 const int64_t* dimsA, const void* A, const int* la, const void* Ad, const int* lad, const void* G,
                 void* Y_out, void* gX_out, void* gA_out, void* gAd_out);

/* No counterpart in the reference: an allocation-free variant of
 * ipeps_absorb for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * fused FP32 path only (4 <= D <= 6, d <= 2; other shapes are rejected), on
 * caller-owned DEVICE pointers of FP32 data, with the labels of ipeps_absorb:
 * Y (may be NULL when only the VJP is wanted) = X A Ad; with the cotangent G
 * the VJP goes to gX / gA / gAd (each may be NULL; gA / gAd in A's leg order).
 * Enqueued on the context's stream (ipeps_ctx_sync waits for it).  Calls no
 * hipMalloc / hipFree: the weight gradient's partial sums live in a
 * file-scope device array of at most 5.6 MiB that each GPU holds until the
 * process ends, shared by every call on that GPU -- calls that overlap on one
 * GPU (two contexts, two threads) race and must be serialised by the caller. */
int ipeps_absorb_noalloc(ipeps_ctx* ctx, const int64_t* dimsX, const float* X, const int* lx, const int64_t* dimsA,
                         const float* A, const int* la, const float* Ad, const int* lad, const float* G, float* Y,
                         float* gX, float* gA, float* gAd);

/* Matrix product on the device, the GEMM that OMEinsum hands to BLAS for a
 * two-index contraction (wrapper.jl:23): C = alpha op(A) op(B) + beta C,
 * column-major, op = X, X^T, conj(X) or X^H.  dtype IPEPS_F64 / IPEPS_C128
 * (FP64 MFMA) or IPEPS_MIXED (float host data, split-bf16 MFMA with
 * `mixed` = 1, 3 or 6 products, FP32 accumulation). */
int ipeps_gemm(ipeps_ctx* ctx, int dtype, int mixed, int transA, int conjA, int transB, int conjB, int64_t M,
               int64_t N, int64_t K, double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
               double beta, void* C, int64_t ldc);

/* svd1 (src/utility/autodiff.jl:22-25): full SVD A = U diag(S) V^H of an
 * m x n matrix; U is m x m, V is n x n (V, not V^H), S has min(m,n). */
int ipeps_svd(ipeps_ctx* ctx, int dtype, int64_t m, int64_t n, const void* A, void* U, double* S, void* V);

/* Thin SVD of a tall m x n matrix (m >= n): A = U diag(S) V^H, S descending,
 * U m x n, V n x n.  The small-SVD step of the randomized projector
 * (B = M^H Q; src/utility/autodiff.jl:22-25 is svd1, which this replaces for
 * tall operands).  method 0: one-sided Jacobi (rocSOLVER gesvdj); 1: Gram
 * eigenpairs B^H B with one Newton-Schulz orthonormalisation of U, taken
 * only when S_min/S_max >= guard, else method 0; 2: as 1 with the Gram
 * eigenpairs from the one-workgroup Jacobi kernel (FP64, n <= 112: the form
 * a captured graph holds); 3: as 1 with the Gram eigenpairs' tridiagonal
 * form solved by rocSOLVER stedc (method 1 solves it by bisection and inverse
 * iteration on the device for n <= 320, stedc only as the fallback).
 * *method_used reports it. */
int ipeps_svd_thin(ipeps_ctx* ctx, int dtype, int method, double guard, int64_t m, int64_t n, const void* A,
                   void* U, double* S, void* V, int* method_used);
/* Orthonormal basis Q (m x n) of the columns of a tall Y (m >= n): the range
 * basis step of the randomized projector that replaces svd1's full SVD
 * (src/utility/autodiff.jl:22-25; the reference has no separate routine).
 * method 0: Householder QR; 1: CholeskyQR2 on rocSOLVER potrf / rocBLAS trsm;
 * 2: CholeskyQR2 on the one-workgroup blocked Cholesky + triangular inverse
 * (F64, n <= 320: n > 160 through a 2 x 2 block split).  Methods 1 and 2 fall
 * back to 0 when a Cholesky fails or
 * the second Gram matrix is farther than 1/2 from I. */
int ipeps_orth(ipeps_ctx* ctx, int dtype, int method, int64_t m, int64_t n, const void* Y, void* Q);
/* svd_back (src/utility/autodiff.jl:41-75) for a square N x N matrix with
 * full U, S, V and cotangents of the first n columns dU (N x n), dS (n),
 * dV (N x n) (any may be NULL); eta is the Lorentzian broadening (1e-12 in
 * the reference).  dA_out is N x N. */
int ipeps_svd_back(ipeps_ctx* ctx, int dtype, int64_t N, int64_t n, const void* U, const double* S, const void* V,
                   const void* dU, const double* dS, const void* dV, double eta, void* dA_out);

/* correlation_spin (src/graph/expectation.jl:246-271): direction 0 = "h"
 * (get_envs_TM_h: L0 = C1 E4 C4, R0 = C2 E2 C3, columns E1 T E3), 1 = "v"
 * (get_envs_TM_v: U0 = C1 E1 C2, D0 = C4 E3 C3, rows E4 T E2).  op1 / op2:
 * d x d ComplexF64 (column-major), inserted as A[..,p1] op[p1,p2] Ad[..,p2].
 * s2s_out receives max_iter complex values (re, im), m1 / m2 = con_LR(L, R0)
 * and con_LR(L0, R).  Complex FP64 on the device for every env dtype. */
int ipeps_correlation_spin(ipeps_env* env, const double* op1, const double* op2, int64_t d, int max_iter,
                           int direction, double* s2s_out, double* m1_out, double* m2_out);
/* Transfer matrix of correlation_TM (src/graph/expectation.jl:193-220):
 * TM = tcon([E4, E2], [[-1,-3,1,2], [-2,-4,1,2]]) for direction 0 ("h") or
 * [E1, E3] for 1 ("v"), reshaped column-major to rows x cols.  tm_apply
 * writes TM X for X (cols x nvec) in the env's host dtype (the operator the
 * reference hands to eigsolve); tm_dense writes TM itself (rows x cols). */
int ipeps_tm_dims(ipeps_env* env, int direction, int64_t* rows, int64_t* cols);
int ipeps_tm_apply(ipeps_env* env, int direction, int64_t nvec, const void* X, void* Y);
int ipeps_tm_dense(ipeps_env* env, int direction, void* TM_out);

/* `steps` rg_steps (src/ctm/ctmrg.jl:57-64) of the environment as replays
 * of a HIP graph: the step is captured on the first call (shapes must be
 * stable, chi saturated; paired_moves=0; randomized projector with
 * chi + oversample <= 112) in the graph-safe form that "graph_safe" names,
 * and each replay's deferred checks are validated; a replay that fails them
 * is undone and the step run eagerly.  The graph is dropped when an option,
 * A or a tensor is set, and recaptured when the environment's buffers have
 * moved (an eager step).  out (may be NULL) = [captures, replays kept,
 * replays undone, eager steps].  No counterpart in the reference. */
int ipeps_env_run_graph(ipeps_env* env, int64_t chi, int steps, int64_t out[4]);

/* Batched independent environments (many instances of CTMTensors +
 * rg_step, src/ctm/ctm_tensor.jl:63-69 and src/ctm/ctmrg.jl:57-64; the
 * reference runs such instances one after another, e.g. the per-basis-vector
 * loop of optim_es.jl:127-153 and its es_resume / part_save job split).
 * n instances with their own A (dims: n x 5, A[i] host data in the dtype's
 * host format), chi[i], and Hamiltonian terms H1[i], H2[i] (hdim[i] =
 * d^2); init_ctm1 runs for each.  `workers` contexts (one HIP stream and host
 * thread each, 1..64) share the instances by a longest-processing-time
 * assignment on the per-step cost.  options: "key=value;..." as
 * ipeps_env_set_option, applied to every instance; with "graph=1" each
 * instance's rg_step is captured into a HIP graph once its shapes are stable
 * (chi saturated) and replayed from then on: the step in graph-safe form (no
 * host round trip; zero fills and copies as kernels; the projector's small
 * SVD from the Gram eigenpairs by the device tridiagonalisation, bisection
 * and inverse iteration, their orthogonality check deferred; the truncation at the
 * largest rank the cutoff rule admits) with its checks (factorisation infos,
 * the cutoff rule replayed on the spectrum) validated after each replay; a
 * replay that fails them is undone and run eagerly.  Needs paired_moves=0
 * (the batch default) and the randomized projector with chi + oversample <=
 * 320; other instances stay eager. */
int ipeps_batch_create(int device, int n, int dtype, const int64_t* dims, const void* const* A, const int64_t* chi,
                       const double* const* H1, const double* const* H2, const int64_t* hdim, int workers,
                       const char* options, ipeps_batch** out);
int ipeps_batch_destroy(ipeps_batch* b);
/* Advance every instance by `steps` rg_steps, then evaluate get_gs_energy
 * (src/graph/evaluation.jl:15-29) of each: E_out (n x 2: re, im); wall time
 * of the call in seconds_out. */
int ipeps_batch_run(ipeps_batch* b, int steps, double* E_out, double* seconds_out);
/* Copy out tensor `which` (as ipeps_env_get_tensor) of instance `inst`;
 * dims (room for 5) / nd receive its shape; host may be NULL. */
int ipeps_batch_get_tensor(ipeps_batch* b, int inst, int which, int64_t* dims, int* nd, void* host);
/* Counters of graph mode since creation: out = [steps captured, replays
 * kept, replays undone (and run eagerly), eager steps].  No counterpart in
 * the reference (its instances run eagerly). */
int ipeps_batch_stats(ipeps_batch* b, int64_t out[4]);

#ifdef __cplusplus
}
#endif

#endif
