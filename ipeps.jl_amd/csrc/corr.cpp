// Transfer-matrix spectrum and spin correlations on the device: the networks
// of src/graph/expectation.jl:193-361 written as pairwise tcon contractions.
//
// Legs: C1 (down, right), C2 (left, down), C3 (up, left), C4 (up, right);
// E1/E3 (left, right, ket, bra), E2/E4 (up, down, ket, bra);
// A (up, left, down, right, phys).  The double-layer tensor
// T[m1..m8] = A[m1,m3,m5,m7,p1] op[p1,p2] Ad[m2,m4,m6,m8,p2] is formed once
// (D^8 elements) and absorbed column by column.
#include "corr.h"

namespace ipeps {
namespace ctm {

using op::tcon;

namespace {

// Correlations run in complex FP64 whatever the environment's dtype (the
// operators may be complex).
Env as_c128(Ctx& c, const Env& e) {
  Env r = e;
  r.dt = DType::C128;
  r.A = op::to_dtype(c, e.A, DType::C128);
  r.Ad = op::to_dtype(c, e.Ad, DType::C128);
  for (int i = 0; i < 4; ++i) {
    r.C[i] = op::to_dtype(c, e.C[i], DType::C128);
    r.E[i] = op::to_dtype(c, e.E[i], DType::C128);
  }
  return r;
}

// T (op == nullptr) or T_op (expectation.jl:310-312)
Tensor double_layer(Ctx& c, const Env& e, const Tensor* op) {
  if (!op) return tcon(c, e.A, {-1, -3, -5, -7, 1}, e.Ad, {-2, -4, -6, -8, 1});
  Tensor AO = tcon(c, e.A, {-1, -2, -3, -4, 1}, *op, {1, -5});
  return tcon(c, AO, {-1, -3, -5, -7, 1}, e.Ad, {-2, -4, -6, -8, 1});
}

// con_LT (expectation.jl:273-281):
// L[p1,p2,p3,p4] = L0[m1,m4,m5,m8] Et[m1,p1,m2,m3] T[..] Ed[m8,p4,m6,m7]
Tensor con_LT(Ctx& c, const Tensor& Et, const Tensor& Ed, const Tensor& L0, const Tensor& T, int dir) {
  Tensor X1 = tcon(c, L0, {1, -1, -2, -3}, Et, {1, -4, -5, -6});  // (m4, m5, m8, p1, m2, m3)
  Tensor X2 = dir == 0 ? tcon(c, X1, {3, 4, -1, -2, 1, 2}, T, {1, 2, 3, 4, -3, -4, -5, -6})   // T[m2,m3,m4,m5,m6,m7,p2,p3]
                       : tcon(c, X1, {3, 4, -1, -2, 1, 2}, T, {3, 4, 1, 2, -5, -6, -3, -4});  // T[m4,m5,m2,m3,p2,p3,m6,m7]
  // X2: (m8, p1, m6, m7, p2, p3)
  return tcon(c, X2, {1, -1, 2, 3, -2, -3}, Ed, {1, -4, 2, 3});
}

// con_TR (expectation.jl:283-291):
// R[p1,p2,p3,p4] = R0[m1,m4,m5,m8] Et[p1,m1,m2,m3] T[..] Ed[p4,m8,m6,m7]
Tensor con_TR(Ctx& c, const Tensor& Et, const Tensor& Ed, const Tensor& T, const Tensor& R0, int dir) {
  Tensor Y1 = tcon(c, R0, {1, -1, -2, -3}, Et, {-4, 1, -5, -6});  // (m4, m5, m8, p1, m2, m3)
  Tensor Y2 = dir == 0 ? tcon(c, Y1, {3, 4, -1, -2, 1, 2}, T, {1, 2, -5, -6, -3, -4, 3, 4})   // T[m2,m3,p2,p3,m6,m7,m4,m5]
                       : tcon(c, Y1, {3, 4, -1, -2, 1, 2}, T, {-5, -6, 1, 2, 3, 4, -3, -4});  // T[p2,p3,m2,m3,m4,m5,m6,m7]
  // Y2: (m8, p1, m6, m7, p2, p3)
  return tcon(c, Y2, {1, -1, 2, 3, -2, -3}, Ed, {-4, 1, 2, 3});
}

// con_LR (expectation.jl:293-297) as a device scalar
Tensor con_LR(Ctx& c, const Tensor& L, const Tensor& R) { return tcon(c, L, {1, 2, 3, 4}, R, {1, 2, 3, 4}); }

void read_c(Ctx& c, const Tensor& s, double out[2]) { to_host(c, s, out); }

}  // namespace

void correlation_spin(Ctx& c, const Env& e0, const Tensor& op1, const Tensor& op2, int max_iter, int dir,
                      std::vector<double>& s2s, double m1[2], double m2[2]) {
  IPEPS_CHECK(dir == 0 || dir == 1, "correlation_spin: direction must be 0 (h) or 1 (v)");
  IPEPS_CHECK(max_iter >= 1, "correlation_spin: max_iter must be >= 1");
  const int64_t d = e0.A.dims[4];
  IPEPS_CHECK(op1.nd() == 2 && op1.dims[0] == d && op1.dims[1] == d && op2.dims == op1.dims,
              "correlation_spin: operators must be d x d");
  TapeScope none(nullptr);
  Env e = as_c128(c, e0);
  const auto& C = e.C;
  const auto& E = e.E;
  Tensor Et, Ed, L0, R0;
  if (dir == 0) {
    // get_envs_TM_h (expectation.jl:299-327)
    Tensor X = tcon(c, C[0], {1, -1}, E[3], {1, -2, -3, -4});  // C1[p1,m1] E4[p1,p2,m2,m3]
    L0 = tcon(c, X, {-1, 1, -2, -3}, C[3], {1, -4});          // C4[p2,m4]
    Tensor Y = tcon(c, C[1], {-1, 1}, E[1], {1, -2, -3, -4});  // C2[m1,p1] E2[p1,p2,m2,m3]
    R0 = tcon(c, Y, {-1, 1, -2, -3}, C[2], {1, -4});          // C3[p2,m4]
    Et = E[0];
    Ed = E[2];
  } else {
    // get_envs_TM_v (expectation.jl:329-361)
    Tensor X = tcon(c, C[0], {-1, 1}, E[0], {1, -2, -3, -4});  // C1[m1,p1] E1[p1,p2,m2,m3]
    L0 = tcon(c, X, {-1, 1, -2, -3}, C[1], {1, -4});          // C2[p2,m4]
    Tensor Y = tcon(c, C[3], {-1, 1}, E[2], {1, -2, -3, -4});  // C4[m1,p1] E3[p1,p2,m2,m3]
    R0 = tcon(c, Y, {-1, 1, -2, -3}, C[2], {-4, 1});          // C3[m4,p2]
    Et = E[3];
    Ed = E[1];
  }
  Tensor T = double_layer(c, e, nullptr);
  Tensor T1 = double_layer(c, e, &op1);
  Tensor T2 = double_layer(c, e, &op2);
  // correlation_spin (expectation.jl:246-271)
  Tensor L = con_LT(c, Et, Ed, L0, T1, dir);
  Tensor R = con_TR(c, Et, Ed, T2, R0, dir);
  s2s.assign(2 * max_iter, 0.0);
  read_c(c, con_LR(c, L, R), &s2s[0]);
  read_c(c, con_LR(c, L, R0), m1);
  read_c(c, con_LR(c, L0, R), m2);
  for (int r = 1; r < max_iter; ++r) {
    L = con_LT(c, Et, Ed, L, T, dir);
    read_c(c, con_LR(c, L, R), &s2s[2 * r]);
  }
}

// TM of expectation.jl:199-211: TM[i1,i2,i3,i4] = sum_kl Ea[i1,i3,k,l] Eb[i2,i4,k,l]
// with (Ea, Eb) = (E4, E2) for h and (E1, E3) for v.
static void tm_pair(const Env& e, int dir, const Tensor*& Ea, const Tensor*& Eb) {
  IPEPS_CHECK(dir == 0 || dir == 1, "transfer matrix: direction must be 0 (h) or 1 (v)");
  Ea = dir == 0 ? &e.E[3] : &e.E[0];
  Eb = dir == 0 ? &e.E[1] : &e.E[2];
}

void tm_dims(const Env& e, int dir, int64_t& rows, int64_t& cols) {
  const Tensor *Ea, *Eb;
  tm_pair(e, dir, Ea, Eb);
  rows = Ea->dims[0] * Eb->dims[0];
  cols = Ea->dims[1] * Eb->dims[1];
}

Tensor tm_dense(Ctx& c, const Env& e, int dir) {
  TapeScope none(nullptr);
  const Tensor *Ea, *Eb;
  tm_pair(e, dir, Ea, Eb);
  Tensor TM = tcon(c, *Ea, {-1, -3, 1, 2}, *Eb, {-2, -4, 1, 2});
  return op::reshape(c, TM, {TM.dims[0] * TM.dims[1], TM.dims[2] * TM.dims[3]});
}

Tensor tm_apply(Ctx& c, const Env& e, int dir, const Tensor& X) {
  TapeScope none(nullptr);
  const Tensor *Ea, *Eb;
  tm_pair(e, dir, Ea, Eb);
  const int64_t n3 = Ea->dims[1], n4 = Eb->dims[1];
  IPEPS_CHECK(X.nd() == 2 && X.dims[0] == n3 * n4, "tm_apply: X must be (cols, nvec)");
  Tensor Xr = op::reshape(c, X, {n3, n4, X.dims[1]});
  // W[i2,k,l,i3,v] = sum_i4 Eb[i2,i4,k,l] X[i3,i4,v]; Y[i1,i2,v] = sum Ea[i1,i3,k,l] W[i2,k,l,i3,v]
  Tensor W = tcon(c, *Eb, {-1, 1, -2, -3}, Xr, {-4, 1, -5});
  Tensor Y = tcon(c, *Ea, {-1, 3, 1, 2}, W, {-2, 1, 2, 3, -3});
  return op::reshape(c, Y, {Y.dims[0] * Y.dims[1], X.dims[1]});
}

}  // namespace ctm
}  // namespace ipeps
