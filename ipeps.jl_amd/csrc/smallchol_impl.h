// Cholesky factor and triangular inverse of a small SPD matrix in one
// workgroup, shared by the device kernel and host-side checks.
//
// G (n x n, FP64, lower triangle read) = L L^T; the kernel returns L^-1
// (lower, zeros above the diagonal).  The lower triangle lives in LDS packed
// by columns: element (i, j), i >= j, at off(j) + i - j with
// off(j) = j n - j (j - 1) / 2, so n (n + 1) / 2 doubles (n <= 192: 148 KB).
//
// Each phase below is run by every thread of the workgroup between barriers;
// within a phase the threads write disjoint elements.
//  * Cholesky, right-looking, column j: (a) thread 0 takes the square root of
//    the pivot (or records failure), (b) the column below it is scaled,
//    (c) the trailing triangle gets the rank-1 update, waves over columns and
//    lanes over rows.
//  * Inverse in place (LAPACK trti2 order), j = n-1 .. 0: (a) the column
//    below the diagonal is copied to a scratch vector and the diagonal
//    inverted, (b) rows i > j get -L^-1_jj * sum_{k=j+1..i} Linv[i][k] x[k]
//    from groups of 4 lanes per row.
#pragma once

#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif

namespace ipeps {
namespace k {

constexpr int SCHOL_MAXN = 192;
constexpr int SCHOL_THREADS = 1024;

IPEPS_HD inline int schol_off(int n, int j) { return j * n - j * (j - 1) / 2; }

// load the lower triangle of G (column-major, leading dimension ld)
IPEPS_HD inline void schol_load(int n, const double* G, int64_t ld, double* P, int tid, int nt) {
  for (int j = 0; j < n; ++j) {
    const int o = schol_off(n, j);
    for (int i = j + tid; i < n; i += nt) P[o + i - j] = G[i + j * ld];
  }
}

// Cholesky phase (a): returns false (and sets *fail) if the pivot is not positive
IPEPS_HD inline void schol_pivot(int n, int j, double* P, int* fail, int tid) {
  if (tid != 0) return;
  const int o = schol_off(n, j);
  const double d = P[o];
  if (!(d > 0.0) || !(d < 1.7e308)) {
    *fail = j + 1;
    return;
  }
  P[o] = sqrt(d);
}

// Cholesky phase (b)
IPEPS_HD inline void schol_scale(int n, int j, double* P, int tid, int nt) {
  const int o = schol_off(n, j);
  const double inv = 1.0 / P[o];
  for (int i = j + 1 + tid; i < n; i += nt) P[o + i - j] *= inv;
}

// Cholesky phase (c): P[i][k] -= P[i][j] P[k][j] for j < k <= i < n
IPEPS_HD inline void schol_update(int n, int j, double* P, int tid, int nt) {
  const int lanes = 64, waves = nt / lanes;
  const int w = tid / lanes, l = tid % lanes;
  const int oj = schol_off(n, j);
  for (int k = j + 1 + w; k < n; k += waves) {
    const double pkj = P[oj + k - j];
    const int ok = schol_off(n, k);
    for (int i = k + l; i < n; i += lanes) P[ok + i - k] -= P[oj + i - j] * pkj;
  }
}

// Inverse phase (a): x = column j below the diagonal, diagonal inverted
IPEPS_HD inline void sinv_prep(int n, int j, double* P, double* x, int tid, int nt) {
  const int o = schol_off(n, j);
  for (int i = j + 1 + tid; i < n; i += nt) x[i] = P[o + i - j];
  if (tid == 0) P[o] = 1.0 / P[o];
}

// Inverse phase (b) for the 4-lane group g (row j+1+g) and lane r in the
// group: partial sum of Linv[i][k] x[k], k = j+1..i, over k = j+1+r (step 4)
IPEPS_HD inline double sinv_partial(int n, int j, const double* P, const double* x, int i, int r) {
  double s = 0.0;
  for (int k = j + 1 + r; k <= i; k += 4) s += P[schol_off(n, k) + i - k] * x[k];
  return s;
}

// store of phase (b): row i of column j
IPEPS_HD inline void sinv_store(int n, int j, double* P, int i, double s) {
  const int o = schol_off(n, j);
  P[o + i - j] = -P[o] * s;
}

// write L^-1 as a full column-major n x n matrix (ld) with zeros above
IPEPS_HD inline void schol_store(int n, const double* P, double* out, int64_t ld, int tid, int nt) {
  for (int e = tid; e < n * n; e += nt) {
    const int i = e % n, j = e / n;
    out[i + j * ld] = i >= j ? P[schol_off(n, j) + i - j] : 0.0;
  }
}

}  // namespace k
}  // namespace ipeps
