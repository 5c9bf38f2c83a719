// Raw kernel launchers (no autodiff).  All run on the context stream.
#pragma once

#include "common.h"

namespace ipeps {
namespace k {

// Which split-bf16 arithmetic the FP32 GEMM uses on MFMA:
//   1 = plain bf16 products, 3 = hi*hi + hi*lo + lo*hi (two-term split),
//   6 = full FP32 accuracy: FP32 MFMA products (v_mfma_f32_16x16x4f32) in
//       place of the three-term split (hi, mid, lo) it stood for, which kept
//       every product above 2^-24.
// Products with K, N <= 96 over >= 4096 rows use FP32 MFMA products at every
// setting (HBM-bound, nothing to gain from bf16).
enum class Mixed : int { BF16 = 1, BF16X3 = 3, BF16X6 = 6 };

// Column-major BLAS-style GEMM:  C = alpha * op(A) * op(B) + beta * C
// op(X) = X, X^T (trans), conj(X) (conj) or X^H (trans && conj).
// op(A) is M x K, op(B) is K x N, C is M x N.  Batched over `batch` with
// element strides sA, sB, sC.  dt selects the arithmetic:
//   F64  -> FP64 MFMA (v_mfma_f64_16x16x4)
//   C128 -> complex FP64 as four real FP64 MFMA products
//   F32  -> FP32 storage, split-bf16 MFMA products, FP32 accumulation
struct GemmArgs {
  DType dt = DType::F64;
  bool transA = false, conjA = false, transB = false, conjB = false;
  int64_t M = 0, N = 0, K = 0;
  double alpha = 1.0, beta = 0.0;
  const void* A = nullptr;
  int64_t lda = 0, sA = 0;
  const void* B = nullptr;
  int64_t ldb = 0, sB = 0;
  void* C = nullptr;
  int64_t ldc = 0, sC = 0;
  int64_t batch = 1;
  Mixed mixed = Mixed::BF16X3;
  // output offsets (F32 / F64, batch 1): element (m, n) of the product goes to
  // C[offM[m] + offN[n]] (device int arrays of M and N entries) instead of
  // C[m + n ldc]: a product written straight into a permuted layout
  const int* offM = nullptr;
  const int* offN = nullptr;
};
void gemm(Ctx& c, const GemmArgs& g);
// offsets[i] = sum_a digit_a(i) stride[a] for i < count, the digits of i in
// the mixed radix ext[0..nd-1] (first fastest); int32 offsets (< 2^31)
void mixed_radix_offsets(Ctx& c, int nd, const int64_t* ext, const int64_t* stride, int64_t count, int* out);
// GEMM autotuner (process-wide table of launch plans keyed by dtype, split,
// transposes, shape, beta and operand alignment).  mode 0: shape rules,
// 1: time the candidate plans of each unseen shape with at least min_flops
// and keep the fastest, 2: use table entries only (no timing runs).
void tune_config(int mode, double min_flops);
void tune_stats(long out[3]);  // entries, shapes tuned, table hits
void tune_clear();
void tune_save(const std::string& path);
void tune_load(const std::string& path);

// out = conj?(in) permuted: out dims[i] = in dims[perm[i]].  Column-major.
void permute(Ctx& c, DType dt, const std::vector<int64_t>& in_dims, const std::vector<int>& perm,
             const void* in, void* out, bool conj);

// Elementwise helpers.  n counts scalars of dt (complex = 1 element).
void fill_zero(Ctx& c, void* p, size_t bytes);
void axpy(Ctx& c, DType dt, int64_t n, double a, const void* x, void* y);        // y += a x
void add(Ctx& c, DType dt, int64_t n, const void* x, const void* y, void* out);  // out = x + y
void scale_by_dev(Ctx& c, DType dt, int64_t n, const void* x, void* y,           // y = x * f(*s)
                  const double* s, int mode);                                    // mode 0: 1/s, 1: s
void conj_copy(Ctx& c, DType dt, int64_t n, const void* x, void* y);
void convert(Ctx& c, DType from, DType to, int64_t n, const void* x, void* y);   // F64<->F32
// Column / row scaling of a column-major matrix X (rows x cols) by a real
// vector s (FP64): Y = X diag(s) (cols) or Y = diag(s) X (rows).
void scale_cols(Ctx& c, DType dt, int64_t rows, int64_t cols, const void* X, const double* s,
                void* Y, bool rowwise);
// Column-wise updates of a block of vectors with per-column ratios
// a_c = num[c] / den[c] (FP64 device vectors; a_c = 0 where den[c] = 0):
//   col_axpy_ratio: X[:, c] += sign a_c P[:, c]
//   col_xpby_ratio: P[:, c] = R[:, c] + a_c P[:, c]
// (the steps of a conjugate-gradient solve with one system per column).
void col_axpy_ratio(Ctx& c, DType dt, int64_t rows, int64_t cols, const double* num, const double* den, double sign,
                    const void* P, void* X);
void col_xpby_ratio(Ctx& c, DType dt, int64_t rows, int64_t cols, const double* num, const double* den, const void* R,
                    void* P);
// Gaussian N x k test matrix from a counter-based generator (deterministic).
void randn(Ctx& c, DType dt, int64_t n, uint64_t seed, void* out);
// The same with the seed formed on the device: seed = draw_seed(base, *draws +
// off) (ops.h), draws a device counter (what a captured graph replays with).
void randn_drawn(Ctx& c, DType dt, int64_t n, uint64_t base, const uint64_t* draws, uint64_t off, void* out);
// *counter += add (one thread).
void counter_add(Ctx& c, uint64_t* counter, uint64_t add);
// Byte copies and zero fills as kernels (graph-safe forms of hipMemcpyAsync /
// hipMemsetAsync): rows of `width` bytes, `rows` of them, pitches in bytes.
void copy_bytes(Ctx& c, void* dst, const void* src, size_t bytes);
void copy_bytes_2d(Ctx& c, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows);
void zero_bytes_2d(Ctx& c, void* p, size_t pitch, size_t width, size_t rows);

// Reductions into device scalars.
//   maxabs: out[0] = max |x|, out_idx[0] = first column-major index of it.
void maxabs(Ctx& c, DType dt, int64_t n, const void* x, double* out, int64_t* out_idx);
//   re_dot: out[0] = Re sum conj(x) y
void re_dot(Ctx& c, DType dt, int64_t n, const void* x, const void* y, double* out);

// Renormalisation backward:
//   gx = gy / m  +  e_idx * (-(Re<gy, x>) / m^2) * x_idx / |x_idx|
// with m = *mval, idx = *midx, rd = *dotv (= Re<gy, x>).
void renorm_back(Ctx& c, DType dt, int64_t n, const void* x, const void* gy, const double* mval,
                 const int64_t* midx, const double* dotv, void* gx);

// svd_back core: builds Gamma (N x N, U's dtype) from UdU = U^H dU (N x n),
// VdV = V^H dV (N x n), dS (n or null), S (N, FP64):
//   Gamma = (J + J^H) S + S (K + K^H) + diag(dS) + diag(i Im(diag UdU) Sinv)
// where J = F o UdU, K = F o VdV zero-padded beyond column n and
// F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eta).
void svd_gamma(Ctx& c, DType dt, int64_t N, int64_t n, const double* S, const void* UdU,
               const void* VdV, const double* dS, double eta, void* Gamma);
// Same for the thin (truncated) factorisation, Gamma is n x n and the
// diagonal also receives nothing extra (the non-square terms are GEMMs).
// Writes also Sinv_n (n) = s/(s^2+eta).
void svd_gamma_thin(Ctx& c, DType dt, int64_t n, const double* S, const void* UdU,
                    const void* VdV, const double* dS, double eta, void* Gamma, double* Sinv);

// Sinv[i] = s_i / (s_i^2 + eta), i < n
void svd_sinv(Ctx& c, int64_t n, const double* S, double eta, double* Sinv);
// s (FP64, n) -> out = s^(-1/2)
void inv_sqrt(Ctx& c, int64_t n, const double* s, double* out);
// gs += -0.5 * s^(-3/2) * g_isq   (backward of inv_sqrt)
void inv_sqrt_back(Ctx& c, int64_t n, const double* s, const double* g_isq, double* gs);
// out[j] = Re sum_i conj(X_ij) * Y_ij over columns of (rows x cols) matrices
// (used for the gradient of a column scaling).
void col_re_dot(Ctx& c, DType dt, int64_t rows, int64_t cols, const void* X, const void* Y,
                double* out, bool rowwise);
// Copy a real FP64 vector into the real parts of dt (F64/C128) storage.
void real_to(Ctx& c, DType dt, int64_t n, const double* x, void* y);
// Real parts (FP64) of dt storage.
void real_of(Ctx& c, DType dt, int64_t n, const void* x, double* y);

// Small energy kernel: given rho_h, rho_v (m x m, dt), H1, H2 (m x m, C128
// on device), computes out = [Re E, Im E, Re Na, Im Na, trh re, im, trv re, im]
// where E = tr(H1 rho_h)/tr(rho_h) + tr(H2 rho_v)/tr(rho_v), all in FP64.
void energy2(Ctx& c, DType dt, int64_t m, const void* rho_h, const void* rho_v, const double* H1,
             const double* H2, double* out);
// Its backward for dL/dRe(E) = 1: grho_h, grho_v (dt, m x m).
void energy2_back(Ctx& c, DType dt, int64_t m, const void* rho_h, const void* rho_v,
                  const double* H1, const double* H2, void* grho_h, void* grho_v);

// y = z x for a complex scalar z (C128 data; F64/F32 data need zi == 0).
void scale_cplx(Ctx& c, DType dt, int64_t n, double zr, double zi, const void* x, void* y);
// Excited-state energy (src/graph/evaluation.jl:31-45) from the component-1
// and component-4 density matrices (C128, m x m):
//   out = [Re E4, Im E4, Nh, Nv] with Nh = Re tr(rh1), Nv = Re tr(rv1),
//   E4 = tr(H1 rh4)/Nh + tr(H2 rv4)/Nv.
void es_energy(Ctx& c, int64_t m, const void* rh1, const void* rv1, const void* rh4, const void* rv4, const double* H1,
               const double* H2, double* out);
// Cotangents of rh4, rv4 for L = Re E4 (Nh, Nv from out): H^H / N.
void es_energy_back(Ctx& c, int64_t m, const double* H1, const double* H2, const double* out, void* g_rh4,
                    void* g_rv4);

// ---- small dense helpers of the projector solvers (n x n, F64 or C128) ----
// out[0] = max_ij |G_ij - delta_ij|  (one workgroup; n up to a few thousand)
void eye_dev(Ctx& c, DType dt, int64_t n, const void* G, double* out);
// Eigen pairs (ascending lam, columns of W) of a Gram matrix B^H B to the
// singular values of B in descending order: s_i = sqrt(max(lam_{n-1-i}, 0)),
// Wr[:, i] = W[:, n-1-i], sinv_i = 1/s_i (0 where s_i = 0).
void eig_to_svd(Ctx& c, DType dt, int64_t n, const double* lam, const void* W, double* s, void* Wr, double* sinv);
// G <- (3 I - G) / 2  (one Newton-Schulz step X <- X (3I - X^H X)/2)
void ns_coeff(Ctx& c, DType dt, int64_t n, void* G);
// Fused site absorption (absorb.hip): Y[r,a,b] = sum X[r,u,v] A[u,a,p] Ad[v,b,p]
// on FP32 data.  A tensor with two row legs (extents n1, n2, strides s1, s2)
// and two groups of two legs of extent D (group x: index i = i1 + D i2 at
// strides sx1, sx2; group y likewise).
struct AbsorbLegs {
  int n1 = 0, n2 = 0;
  int64_t s1 = 0, s2 = 0, sx1 = 0, sx2 = 0, sy1 = 0, sy2 = 0;
};
struct AbsorbP {
  int U = 0, P = 0, D = 0;    // U = D^2, P = physical extent
  AbsorbLegs x, y;            // X: rows, (u, v); Y: rows, (a, b)
  const float* X = nullptr;
  float* Y = nullptr;
  // A and Ad in any layout: element (u1 + D u2, a1 + D a2, p) of A sits at
  // Am[u1 sa[0] + u2 sa[1] + a1 sa[2] + a2 sa[3] + p sa[4]], likewise Ad
  // with sd for (v1, v2, b1, b2, p); the kernels gather them into LDS, so the
  // site tensor needs no permute (its transpose is a swap of strides)
  const float* Am = nullptr;
  const float* Adm = nullptr;
  int sa[5] = {0, 0, 0, 0, 0}, sd[5] = {0, 0, 0, 0, 0};
};
struct AbsorbWP {
  int U = 0, P = 0, D = 0;
  AbsorbLegs x, g;            // X: rows, (u, v); G = dL/dY: rows, (a, b)
  const float* X = nullptr;
  const float* G = nullptr;
  const float* Am = nullptr;
  const float* Adm = nullptr;
  int sa[5] = {0, 0, 0, 0, 0}, sd[5] = {0, 0, 0, 0, 0};  // as in AbsorbP
  float* partial = nullptr;   // set by absorb_wgrad
};
bool absorb_supported(int D, int P);
void absorb_fwd(Ctx& c, const AbsorbP& p);
// out (device, 2 U^2 P floats): Abar as [u + U (a + U p)], then Adbar as [v + U (b + U p)]
void absorb_wgrad(Ctx& c, AbsorbWP p, float* out);
// The same with its partial sums and its result in the library's per-device
// scratch array (no allocation; calls that overlap on one GPU race for it):
// returns the device address of the result, valid until the next call on the
// device.  The device's scratch address is taken once by absorb_scratch_init,
// which ctx_create calls with the device current.
float* absorb_wgrad_scratch(Ctx& c, AbsorbWP p);
void absorb_scratch_init(int device);
int64_t absorb_scratch_floats();

constexpr int SCHOL_MAXN_HOST = 160;  // largest n of small_chol_inv
// L^-1 of the Cholesky factor G = L L^T of an SPD n x n FP64 matrix (n <= 160)
// in one workgroup (blocked, 32 x 32 tiles in LDS); info (device int) = 0, or
// j + 1 if pivot j is not positive.
void small_chol_inv(Ctx& c, int n, const double* G, int64_t ldg, double* Linv, int64_t ldl, int* info);

constexpr int TRI_MAXN_HOST = 320;  // largest n of sytrd_q
// Householder tridiagonalisation Q^T G Q = tridiag(d, e) of a symmetric n x n
// FP64 matrix (n <= 320; lower triangle read): d (n), e (n - 1) and the
// orthogonal Q (n x n, leading dimension ldq), the input rocSOLVER's stedc
// takes for the eigenvectors of G.
void sytrd_q(Ctx& c, int n, const double* G, int64_t ldg, double* d, double* e, double* Q, int64_t ldq);

constexpr int TRIEIG_MAXN_HOST = 320;  // largest n of tri_eig
// Eigenpairs of the symmetric tridiagonal matrix with diagonal d (n) and
// off-diagonal e (n - 1): lam ascending by bisection (one wave per
// eigenvalue), unit eigenvectors by inverse iteration (one thread per
// eigenvalue) as the rows of Zt (n x n, Zt[j + i n] = component i of vector
// j).  Vectors of (nearly) equal eigenvalues are independent but not
// orthogonal: the caller orthonormalises and checks the set.  Kernels only.
void tri_eig(Ctx& c, int n, const double* d, const double* e, double* lam, double* Zt);

constexpr int JAC_MAXN_HOST = 112;  // largest n of jacobi_eig
// Eigenpairs G = Q diag(lam) Q^T of a symmetric n x n FP64 matrix (n <= 112;
// lower triangle read) in one workgroup by cyclic two-sided Jacobi: lam
// ascending, Q (n x n, leading dimension ldq) the matching eigenvectors;
// info (device int) = 0, or 1 if the off-diagonal part did not fall below the
// relative threshold within the sweep limit.  Kernels only (no library call,
// no memset): the routine a captured graph can hold.
void jacobi_eig(Ctx& c, int n, const double* G, int64_t ldg, double* lam, double* Q, int64_t ldq, int* info);

}  // namespace k
}  // namespace ipeps
