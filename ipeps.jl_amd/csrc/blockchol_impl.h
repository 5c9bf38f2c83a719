// Blocked Cholesky factor and triangular inverse of a small SPD matrix in one
// workgroup (FP64, n <= 160): the tile layout and the per-element steps (b),
// (c) and the inverse phase of the device kernel small_chol_inv_kernel in
// elem.hip; step (a), the one-wave diagonal-tile factor, is in the kernel.
//
// G (n x n, lower triangle read) = L L^T; the kernel returns L^-1 (lower,
// zeros above).  The matrix is padded to NB 32 x 32 tiles with the identity
// (the padded factor and inverse are block-diagonal, so the leading n x n
// block is exact).  The lower tiles live in LDS, tile (I, J), I >= J, at
// index I (I + 1) / 2 + J, each column-major with leading dimension 33.
//
// Cholesky, right-looking over tile columns K:
//  (a) one wave factors the diagonal tile in registers (lane i holds row i,
//      columns broadcast with readlane) and replaces it by its inverse;
//  (b) the panel below: L_IK = A_IK Dinv_K^T, one element per thread pass;
//  (c) the trailing tiles: A_IJ -= L_IK L_JK^T for K < J <= I.
// Inverse, over tile columns J = NB-2 .. 0 and rows I = NB-1 .. J+1:
//  X_IJ = -(sum_{M=J+1..I} X_IM L_MJ) X_JJ, the sum staged in a scratch tile;
//  rows are taken in descending order so that L_MJ (M <= I) is still in place.
#pragma once

#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif

namespace ipeps {
namespace k {

constexpr int BCH_T = 32;                // tile size
constexpr int BCH_LD = 33;               // tile leading dimension (LDS banks)
constexpr int BCH_TE = BCH_T * BCH_LD;   // doubles per tile
constexpr int BCH_NBMAX = 5;             // tiles per dimension
constexpr int BCH_MAXN = BCH_T * BCH_NBMAX;
constexpr int BCH_NTILES = BCH_NBMAX * (BCH_NBMAX + 1) / 2;
constexpr int BCH_THREADS = 512;

IPEPS_HD inline int bch_tix(int I, int J) { return I * (I + 1) / 2 + J; }
IPEPS_HD inline int bch_at(int I, int J, int r, int c) { return bch_tix(I, J) * BCH_TE + c * BCH_LD + r; }

// load the lower tiles of G (padding: identity)
IPEPS_HD inline void bch_load(int n, int NB, const double* G, int64_t ld, double* T, int tid, int nt) {
  const int nel = NB * (NB + 1) / 2 * BCH_T * BCH_T;
  for (int e = tid; e < nel; e += nt) {
    const int t = e / (BCH_T * BCH_T), rc = e % (BCH_T * BCH_T);
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    const int J = t - I * (I + 1) / 2;
    const int r = rc % BCH_T, c = rc / BCH_T;
    const int gi = I * BCH_T + r, gj = J * BCH_T + c;
    double v;
    if (gi < n && gj < n) v = G[gi + gj * ld];
    else v = gi == gj ? 1.0 : 0.0;
    T[bch_at(I, J, r, c)] = v;
  }
}

// (b) element e of the panel below diagonal tile K: returns the new value
// of L_IK[r][j] = sum_{m <= j} A_IK[r][m] Dinv_K[j][m]
IPEPS_HD inline double bch_panel(const double* T, int NB, int K, int e, int* I_out, int* r_out, int* j_out) {
  const int I = K + 1 + e / (BCH_T * BCH_T), rc = e % (BCH_T * BCH_T);
  const int r = rc % BCH_T, j = rc / BCH_T;
  double s = 0.0;
  for (int m = 0; m <= j; ++m) s += T[bch_at(I, K, r, m)] * T[bch_at(K, K, j, m)];
  *I_out = I;
  *r_out = r;
  *j_out = j;
  (void)NB;
  return s;
}

// (c) trailing element e over the tiles (I, J), K < J <= I < NB
IPEPS_HD inline void bch_update(double* T, int NB, int K, int e) {
  const int m = NB - K - 1;  // trailing tile rows
  int t = e / (BCH_T * BCH_T);
  const int rc = e % (BCH_T * BCH_T);
  int a = 0;  // tile (K+1+a, K+1+b), b <= a
  while ((a + 1) * (a + 2) / 2 <= t) ++a;
  const int b = t - a * (a + 1) / 2;
  (void)m;
  const int I = K + 1 + a, J = K + 1 + b;
  const int r = rc % BCH_T, c = rc / BCH_T;
  double s = 0.0;
  for (int q = 0; q < BCH_T; ++q) s += T[bch_at(I, K, r, q)] * T[bch_at(J, K, c, q)];
  T[bch_at(I, J, r, c)] -= s;
}

// inverse phase, sum tile: W[r][c] = sum_{M=J+1..I} X_IM[r][:] L_MJ[:][c]
IPEPS_HD inline double bch_isum(const double* T, int I, int J, int r, int c) {
  double s = 0.0;
  for (int M = J + 1; M <= I; ++M)
    for (int q = 0; q < BCH_T; ++q) s += T[bch_at(I, M, r, q)] * T[bch_at(M, J, q, c)];
  return s;
}

// inverse phase, product: X_IJ[r][c] = -sum_{q >= c} W[r][q] X_JJ[q][c]
IPEPS_HD inline double bch_iprod(const double* T, const double* W, int J, int r, int c) {
  double s = 0.0;
  for (int q = c; q < BCH_T; ++q) s += W[q * BCH_LD + r] * T[bch_at(J, J, q, c)];
  return -s;
}

}  // namespace k
}  // namespace ipeps
