// Householder reduction of a small real symmetric matrix to tridiagonal form
// in one workgroup (FP64, n <= 320), and the formation of the orthogonal
// factor: constants and index helpers of the device kernels in elem.hip
// (sytrd_kernel, orgtr_kernel; the packed-triangle helpers serve the Jacobi
// eigensolver there).
//
// LAPACK dsytd2 'L' convention: for i = 0 .. n-3 the reflector
// H(i) = I - tau_i v v^T (v_0 = 1) maps A(i+1:n, i) to e_i e_1 and is applied
// to the trailing block from both sides,
//   p = tau A v,  w = p - (tau / 2)(p . v) v,  A <- A - v w^T - w v^T.
// The reflectors go to global memory as a column-major n x n array V
// (V(i+1, i) = 1, V(r, i) = v_{r-i-1}) and tau.
//
// sytrd_kernel stores the matrix as a square.  While the trailing block is
// larger than TRI_SQ_LDS a side it lives in a global-memory workspace (ld =
// n + 64: each column has 64 padding rows, so a lane's 64-row block never
// needs a bounds test), read and written through the L2 of the workgroup's
// XCD, and only its lower triangle is kept: lane rows accumulate the lower
// part of p, a wave sum per column its transposed part.  Then the trailing
// square (both triangles, TRI_SQ_LDS a side at most) moves to LDS, where p is
// a plain row-per-lane product with no per-column wave sums.  Step j - 1's
// rank-2 update is deferred into step j: column j is updated by the wave that
// forms reflector j, the rest of the trailing block in the same pass that
// forms p = tau A v, so each element is read and written once per step.
//
// Q = H(0) H(1) ... H(n-3) acts on rows / columns 1 .. n-1 (Q(0, 0) = 1).
// Column c of Q is independent of the others: starting from e_c, apply
// H(c-1), H(c-2), ..., H(0) in that order (the reflectors H(i), i >= c, leave
// e_c unchanged), so one wave forms one column.
#pragma once

#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif

namespace ipeps {
namespace k {

constexpr int TRI_MAXN = 160;  // orgtr_kernel: 3 or 5 row blocks per lane
constexpr int TRI_BIG_MAXN = 320;
constexpr int TRI_SQ_LDS = 128;  // sytrd_kernel works in LDS once the trailing square is at most this a side
constexpr int TRI_THREADS = 1024;

IPEPS_HD inline int tri_off(int n, int c) { return c * n - c * (c - 1) / 2; }
IPEPS_HD inline int tri_at(int n, int r, int c) {
  return r >= c ? tri_off(n, c) + r - c : tri_off(n, r) + c - r;
}

}  // namespace k
}  // namespace ipeps
