// Differentiable tensor operations on the device and the reverse-mode tape.
//
// Every op runs its forward kernels on the context stream.  When the input
// carries a tape node and a tape is recording, the op records a closure that
// maps the output cotangent to input cotangents (hand-written VJPs).  The
// cotangent convention is the one of Zygote/ChainRules: for a real loss L of
// complex z the gradient is dL/dRe z + i dL/dIm z, so the VJP of a linear map
// is its adjoint.
#pragma once

#include <functional>
#include <memory>
#include <mutex>
#include <vector>

#include "common.h"
#include "kernels.h"

namespace ipeps {

struct Node {
  Tensor grad;                                  // accumulated cotangent of the output
  std::vector<Tensor> slots;                    // per-output cotangents of multi-output ops
  std::function<void(Ctx&, Node&)> back;        // propagate to inputs (null for leaves)
  bool leaf = false;
  int seg = 0;                                  // tape segment (a lane of a concurrent pair), 0 outside
};

struct Tape {
  bool active = false;
  std::vector<std::shared_ptr<Node>> order;
  std::mutex mu;  // two lanes of a CTMRG step record concurrently
  // Two independent computations recorded concurrently (a pair of CTMRG
  // moves): order[m0, m1) by the main lane (segment smain), order[m1, l1) by
  // the second lane (segment slane).  The reverse sweep runs them on two lanes.
  struct Pair {
    size_t m0, m1, l1;
    int smain, slane;
  };
  std::vector<Pair> pairs;
  int next_seg = 1;
  void push(std::shared_ptr<Node> n) {
    std::lock_guard<std::mutex> g(mu);
    order.push_back(std::move(n));
  }
  void clear() {
    order.clear();
    pairs.clear();
  }
};

// Current tape of this thread (null when not recording).
Tape*& current_tape();
// Tape segment of the nodes this thread records (0: none).
int& current_seg();

struct TapeScope {
  Tape* prev;
  explicit TapeScope(Tape* t) : prev(current_tape()) { current_tape() = t; }
  ~TapeScope() { current_tape() = prev; }
};

namespace op {

// Precision knobs of the contraction engine.
struct Prec {
  k::Mixed mixed = k::Mixed::BF16X3;  // products of the forward ops
  k::Mixed back = k::Mixed::BF16X3;   // products of the VJPs recorded meanwhile
  bool fused_absorb = true;  // op::absorb may use the fused FP32-MFMA kernel
};
Prec& prec();
// Sets prec().mixed and prec().back for a scope; each VJP runs with the
// `back` setting in force when its forward op was recorded.
struct PrecScope {
  k::Mixed prev, prev_back;
  explicit PrecScope(k::Mixed m) : PrecScope(m, m) {}
  PrecScope(k::Mixed m, k::Mixed b) : prev(prec().mixed), prev_back(prec().back) {
    prec().mixed = m;
    prec().back = b;
  }
  ~PrecScope() {
    prec().mixed = prev;
    prec().back = prev_back;
  }
};

bool recording();
// Mark t as a differentiable leaf on the current tape.
Tensor leaf(Tensor t);
// Reverse sweep over the tape (the loss node's grad must be seeded).
void backward(Ctx& c, Tape& t);
// Add g to n's cotangent.
void acc(Ctx& c, const std::shared_ptr<Node>& n, const Tensor& g);
// Fresh node on the current tape.
std::shared_ptr<Node> make_node(std::function<void(Ctx&, Node&)> back);

// ---- raw (non-recording) helpers -----------------------------------------
Tensor clone(Ctx& c, const Tensor& t);
Tensor permute_raw(Ctx& c, const Tensor& t, const std::vector<int>& perm, bool conj = false);
// With `into` (same dtype and element count as the result): when the product
// comes out of its GEMM in the result's own leg order, it is added to *into by
// the GEMM (beta = 1) and *into is returned; otherwise a new tensor as usual.
Tensor tcon_raw(Ctx& c, const Tensor& A, const std::vector<int>& la, bool conjA, const Tensor& B,
                const std::vector<int>& lb, bool conjB, Tensor* into = nullptr);
Tensor to_dtype(Ctx& c, const Tensor& t, DType dt);
// C = op(A) op(B) for matrices (column-major), raw.
Tensor matmul_raw(Ctx& c, const Tensor& A, bool tA, bool cA, const Tensor& B, bool tB, bool cB);

// ---- differentiable ops -------------------------------------------------------
// ncon-style contraction: positive labels are summed, negative labels are the
// output legs in the order -1, -2, ...  (src/utility/wrapper.jl:15-23)
Tensor tcon(Ctx& c, const Tensor& A, const std::vector<int>& la, const Tensor& B, const std::vector<int>& lb);
Tensor reshape(Ctx& c, const Tensor& t, const std::vector<int64_t>& dims);
Tensor permute(Ctx& c, const Tensor& t, const std::vector<int>& perm);
Tensor conj(Ctx& c, const Tensor& t);
// X ./ maximum(abs.(X)) with the gradient of maximum at the first maximum
// in column-major order (src/utility/wrapper.jl:57).
Tensor renormalize(Ctx& c, const Tensor& t);
// x ./ maximum(abs.(ref)): the nested renormalisation of the B components by
// the plain component's factor (src/nested_tensor/nested_tensor.jl:41-46).
Tensor renormalize_by(Ctx& c, const Tensor& x, const Tensor& ref);
// X diag(s) (rowwise=false) or diag(s) X (rowwise=true), s a real FP64 vector.
Tensor scale_vec(Ctx& c, const Tensor& X, const Tensor& s, bool rowwise);
// s .^ (-1/2) for a real FP64 vector.
Tensor inv_sqrt(Ctx& c, const Tensor& s);
// Y = X A Ad contracted as the labels say (ncon convention), where X has two
// open legs, two legs shared with A and two with Ad, and A / Ad share their
// physical leg: the enlarged-corner absorption of a site.  FP32 data with
// D^2 <= 36, D >= 4 and a physical extent <= 2 run the fused FP32-MFMA kernel
// (absorb.hip) with its own VJPs; anything else two tcon contractions.
Tensor absorb(Ctx& c, const Tensor& X, const std::vector<int>& lx, const Tensor& A, const std::vector<int>& la,
              const Tensor& Ad, const std::vector<int>& lad);
// absorb's fused FP32 kernels on caller-owned device memory, allocating
// nothing: Y (if non-null) = X A Ad, and with the cotangent G the VJP into
// gX / gA / gAd (each may be null).  The weight gradient's partial sums live
// in the library's per-device scratch array, so calls that overlap on one GPU
// race.  Shapes outside the fused kernel's are rejected.
void absorb_noalloc(Ctx& c, const std::vector<int64_t>& dx, const float* X, const std::vector<int>& lx,
                    const std::vector<int64_t>& da, const float* A, const std::vector<int>& la, const float* Ad,
                    const std::vector<int>& lad, const float* G, float* Y, float* gX, float* gA, float* gAd);
Tensor matmul(Ctx& c, const Tensor& A, const Tensor& B);       // A * B
Tensor adjoint_matmul(Ctx& c, const Tensor& A, const Tensor& B);  // A^H * B
Tensor cast(Ctx& c, const Tensor& t, DType dt);                   // differentiable F64 <-> F32
Tensor add(Ctx& c, const Tensor& a, const Tensor& b);             // a + b
Tensor scale_cplx(Ctx& c, const Tensor& t, double zr, double zi); // z t, z a host constant

struct SVDResult {
  Tensor U, S, V;   // U: N x n, S: n (FP64), V: N x n
  std::vector<double> s_host;  // all computed singular values (descending)
  int64_t n = 0;
};

// Counters of the randomized projectors of one environment (shared by the
// two lanes of a paired step, hence the lock).
struct SvdStats {
  std::mutex mu;
  long calls = 0;          // randomized projectors built
  long power_iters = 0;    // power iterations, summed over calls
  long power_max = 0;      // largest count in one call
  double resid_max = 0.0;  // largest final a-posteriori residual (relative to s_1)
  long checks = 0;         // a-posteriori checks made
  long cg_calls = 0;       // discarded-subspace solves in the adjoint
  long cg_iters = 0;       // their CG iterations, summed
  double cg_resid_max = 0.0;  // largest final relative CG residual
  long cg_fallbacks = 0;   // solves abandoned (operator not positive definite): thin term used
  long orth_calls = 0;     // range-basis orthonormalisations
  long orth_second = 0;    // ... that needed CholeskyQR2's second pass
  long orth_householder = 0;  // ... that fell back to Householder QR
  long warm_starts = 0;    // projectors started from the previous one's basis
  long tri_fallbacks = 0;  // tridiagonal eigenpairs redone by rocSOLVER stedc (inverse iteration rejected)
  void reset() {
    std::lock_guard<std::mutex> g(mu);
    calls = power_iters = power_max = checks = cg_calls = cg_iters = cg_fallbacks = 0;
    orth_calls = orth_second = orth_householder = warm_starts = tri_fallbacks = 0;
    resid_max = cg_resid_max = 0.0;
  }
};

// Options of the truncated SVD of the half-system matrix.
struct SvdOpts {
  int mode = 0;          // 0 = full SVD of M = R1 R2 (reference semantics), 1 = randomized on (R1, R2)
  int oversample = 16;   // k = n_target + oversample
  int power_iters = 1;   // q
  int full_below = 0;    // use the full SVD when N <= full_below (randomized mode)
  uint64_t seed = 1234;
  double eta = 1e-12;    // Lorentzian broadening of the adjoint (src/utility/autodiff.jl:41)
  // Relative singular-value cutoff of FP32 (mixed) engines; F64 / C128 use the
  // reference's 1e-12 (src/utility/wrapper.jl:38-39).
  double mixed_cutoff = 1e-6;
  // Randomized mode only:
  //   fast_orth = 1: range bases by CholeskyQR2 (two Gram + Cholesky + triangular
  //                  solve passes), falling back to Householder QR when a
  //                  Cholesky fails or the second Gram is farther than 1/2 from I;
  //                  2: the same with the one-workgroup blocked FP64 Cholesky +
  //                  inverse (F64, k <= 160; a 2 x 2 block split of it for
  //                  k <= 320; otherwise as 1);
  //   fast_svd  = 1: the small SVD of B = M^H Q from the eigenpairs of B^H B
  //                  plus one Newton-Schulz orthonormalisation of B's singular
  //                  vectors, falling back to one-sided Jacobi (gesvdj) when
  //                  s_min / s_max < gram_guard (the Gram eigenvalues then lose
  //                  more than ~1e-8 relative accuracy);
  //                  the Gram eigenpairs (F64, k <= 320) by the device
  //                  tridiagonalisation, bisection and inverse iteration with
  //                  Newton-Schulz orthonormalisation, rocSOLVER stedc when
  //                  the vectors fail their check;
  //                  2: the same with the Gram eigenpairs from the
  //                  one-workgroup Jacobi kernel (F64, k <= 112);
  //                  3: as 1 with rocSOLVER stedc for every tridiagonal form.
  int fast_orth = 2;
  int fast_svd = 1;
  double gram_guard = 1e-4;
  // Adaptive power iterations: after the first power_iters, while
  // resid_tol > 0 and fewer than power_iters_max have run, the kept triplets
  // are checked a posteriori -- max_i ||M v_i - s_i u_i|| / s_1 over the n
  // triplets the cutoff keeps (M^H u_i = s_i v_i holds by construction) --
  // and one more power iteration runs while the check exceeds resid_tol.
  int power_iters_max = 0;  // <= power_iters: no adaptive iterations
  double resid_tol = 0.0;
  // Adjoint: the discarded-subspace term of svd_back beyond the k computed
  // triplets.  0: its s_j -> 0 limit (the thin-SVD terms); > 0: solved by
  // conjugate gradients, at most complement_iters iterations, to a relative
  // residual complement_tol.
  int complement_iters = 0;
  double complement_tol = 1e-8;
  // Range-finder products of FP32 engines: the first applications of M / M^H
  // (all but the last `late_apps` before B = M^H Q) run with `early_mixed`
  // products (1, 3 or 6; 0 = the projector's setting).  Subspace iteration
  // damps an error made in an early application by (s_{k+1} / s_n)^2 per
  // later application pair, so only the last applications and B need the
  // projector's accuracy.
  int early_mixed = 0;
  int late_apps = 2;
  // Test matrix seed formed on the device from a draw counter (graph
  // capture): draw_seed(seed_base, *draws_dev + draw_off) instead of `seed`.
  const uint64_t* draws_dev = nullptr;
  uint64_t seed_base = 0, draw_off = 0;
  // Graph-safe mode: the rank to keep (> 0; the cutoff rule is checked
  // against it after the step), else the largest the rule admits.
  int64_t fixed_n = 0;
  std::shared_ptr<SvdStats> stats;  // optional counters
  // Warm start: with warm_start set and warm_in holding an m x k basis in the
  // solver dtype, the range finder starts from it (the previous projector's
  // left Ritz vectors of the same move) instead of M Omega; warm_out receives
  // this projector's left Ritz vectors (m x k) for the next one.
  int warm_start = 0;
  const Tensor* warm_in = nullptr;
  Tensor* warm_out = nullptr;
};

// Truncated SVD with the reference cutoff rule (src/utility/wrapper.jl:25-55):
// n = min(chi, #{s/smax > 1e-12}); if |s[n+1]-s[n]| < 1e-10 then n += 1.
// mode 0: M is formed, full SVD, adjoint = full svd_back (autodiff.jl:41-75).
// mode 1: randomized range finder on M = R1 R2, adjoint = thin svd_back
//         propagated to R1 and R2 without forming M.
// o.seed is used as is for the test matrix (see draw_seed).
SVDResult svd_trunc(Ctx& c, const Tensor& R1, const Tensor& R2, int64_t chi, const SvdOpts& o);
// Seed of the i-th draw of a stream keyed by `base` (splitmix64 finaliser).
uint64_t draw_seed(uint64_t base, uint64_t i);

// Orthonormal basis of the columns of Y (N x k, F64 or C128): Householder QR
// (fast = 0), CholeskyQR2 with rocSOLVER Cholesky factors and triangular
// solves (1) or with the one-workgroup Cholesky + inverse for FP64, k <= 192
// (2); 1 and 2 fall back to Householder as described at orth_cholqr2.
Tensor orth(Ctx& c, const Tensor& Y, int fast = 0);
// Thin SVD of a tall F64/C128 matrix B = W diag(s) X^H (s descending).
// method 0: one-sided Jacobi (rocSOLVER gesvdj); 1: Gram eigenpairs when
// s_min/s_max >= guard, else method 0.  Returns the method used.
int svd_tall(Ctx& c, const Tensor& B, int method, double guard, Tensor& W, Tensor& s, Tensor& X);
// Full SVD X = U diag(s) V^H (U m x m, V n x n, V returned, not V^H).
void svd_full_raw(Ctx& c, const Tensor& X, Tensor& U, Tensor& s, Tensor& V);
// Thin SVD of a tall matrix X (m x n, m >= n): X = W diag(s) Z^H.
void svd_thin_raw(Ctx& c, const Tensor& X, Tensor& W, Tensor& s, Tensor& Z);

}  // namespace op
}  // namespace ipeps
