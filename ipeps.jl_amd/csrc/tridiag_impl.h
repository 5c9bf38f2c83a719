// Householder reduction of a small real symmetric matrix to tridiagonal form
// in one workgroup (FP64, n <= 320), and the formation of the orthogonal
// factor: the storage scheme and index helpers of the device kernels in
// elem.hip (sytrd_kernel for n <= 160, sytrd_big_kernel for n <= 320,
// orgtr_kernel).
//
// Lower storage, LAPACK dsytd2 'L' convention: for i = 0 .. n-3 the reflector
// H(i) = I - tau_i v v^T (v_0 = 1) maps A(i+1:n, i) to e_i e_1 and is applied
// to the trailing block from both sides,
//   p = tau A v,  w = p - (tau / 2)(p . v) v,  A <- A - v w^T - w v^T,
// the product and the update column by column (one wave per column).
// The lower triangle lives in LDS packed by columns: (r, c), r >= c, at
// tri_off(n, c) + r - c.  The reflectors go to global memory as a
// column-major n x n array V (V(i+1, i) = 1, V(r, i) = v_{r-i-1}) and tau.
//
// sytrd_big_kernel keeps the triangle in a global-memory workspace (the same
// packing, read and written through the L2 of the workgroup's XCD) while the
// block a step touches is larger than TRI_LDS_M + 1, then moves A(j.., j..)
// into LDS (packed with its own dimension) for the remaining steps.  Step
// j - 1's rank-2 update is deferred into step j: column j is updated by the
// wave that forms reflector j, and the rest of the trailing block in the
// same pass that forms p = tau A v, so each element is read and written once
// per step.  Lane rows accumulate the lower part of p and a wave sum gives
// each column's transposed part.
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

constexpr int TRI_MAXN = 160;
constexpr int TRI_BIG_MAXN = 320;
constexpr int TRI_LDS_M = 159;  // sytrd_big_kernel works in LDS once A(j.., j..) is at most TRI_LDS_M + 1 a side
constexpr int TRI_THREADS = 1024;

IPEPS_HD inline int tri_off(int n, int c) { return c * n - c * (c - 1) / 2; }
IPEPS_HD inline int tri_at(int n, int r, int c) {
  return r >= c ? tri_off(n, c) + r - c : tri_off(n, r) + c - r;
}

}  // namespace k
}  // namespace ipeps
