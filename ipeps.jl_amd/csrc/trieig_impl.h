// Eigenpairs of a real symmetric tridiagonal matrix T (diagonal d[0..n-1],
// off-diagonal e[0..n-2], T(i+1, i) = T(i, i+1) = e[i]) by bisection and
// inverse iteration.  The steps below are plain inline functions compiled
// for the device kernels in elem.hip (tri_bisect_kernel, tri_invit_kernel)
// and, with IPEPS_HD defined empty, for the host test driver
// tests/trieig_host.cpp, which checks them against LAPACK on the CPU.
//
// Eigenvalues (ascending, index j = 0 .. n-1): Sturm counts
//   q_0 = d_0 - s,  q_i = (d_i - s) - e_{i-1}^2 / q_{i-1},
// count(s) = #{i : q_i < 0} = number of eigenvalues below s, with |q_i| kept
// above pivmin (Kahan's safeguard, as LAPACK dstebz).  The device runs one
// wave per eigenvalue: each step its 64 lanes count at TRIEIG_PTS points each
// (64 TRIEIG_PTS interior points of [lo, hi]) and the interval shrinks to the
// two neighbouring points around the first point whose count exceeds j.
//
// Eigenvectors: for each eigenvalue lam, a few solves (T - lam I) x = b from
// a pseudo-random start b (the sign of the result fixed by x . b > 0), each
// with the LU factorisation of T - lam I with
// partial pivoting (the carried row never has more than two non-zeros; a
// row swap puts a second superdiagonal in U), pivots below pert replaced by
// +-pert.  Vectors of eigenvalues closer than the solve's backward error
// (~ u |T|) come out as independent combinations of their eigenvectors; the
// caller orthonormalises the set (elem.hip: Newton-Schulz steps on the
// Gram matrix of the vectors, checked against the identity).
#pragma once

#include <math.h>
#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif
// Inlined into the kernels, so that d and e stay LDS (not generic) pointers
// and the solve's unrolled sweeps keep their values in registers.
#define TRIEIG_FN IPEPS_HD inline __attribute__((always_inline))

namespace ipeps {
namespace k {

constexpr int TRIEIG_MAXN = 320;    // LDS copies of d and e^2 in the kernels
// Multisection points per lane: the counts are bound by the issue rate of
// their FP64 reciprocals and FMAs per SIMD, so more points per lane (more
// work per step, fewer steps) measured slower (4: 454 us against 206 us for
// 288 eigenvalues).
constexpr int TRIEIG_PTS = 1;
constexpr int TRIEIG_MAXIT = 20;    // multisection steps (65^20 ~ 2e36)
constexpr int TRIEIG_SOLVES = 2;    // inverse-iteration solves per eigenvalue
constexpr int TRIEIG_BLK = 16;      // rows per block of the sweeps' batched loads

// Gershgorin interval [gl, gu] widened as dstebz does, |T|_1 bound tnorm and
// the pivot safeguard pivmin.
TRIEIG_FN void tri_bounds(int n, const double* d, const double* e, double& gl, double& gu, double& tnorm,
                                double& pivmin) {
  double lo = d[0], hi = d[0], emax2 = 0.0;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    if (i < n - 1) emax2 = fmax(emax2, e[i] * e[i]);
  }
  const double eps = 2.220446049250313e-16, safemin = 2.2250738585072014e-308;
  tnorm = fmax(fabs(lo), fabs(hi));
  pivmin = safemin * fmax(1.0, emax2);
  const double w = 2.1 * tnorm * eps * n + 4.2 * pivmin;
  gl = lo - w;
  gu = hi + w;
}

// 1 / x for x != 0 (finite, normal): on the device the hardware reciprocal
// refined by two Newton steps r <- r + r (1 - x r), each doubling its correct
// bits (the latency of a few FMAs instead of the IEEE division sequence); on
// the host a division.  Within an ulp or two of the correctly rounded value,
// which the Sturm counts and the solves below absorb in their backward error.
TRIEIG_FN double tri_rcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(x);
  double t = fma(-x, r, 1.0);
  r = fma(r, t, r);
  t = fma(-x, r, 1.0);
  return fma(r, t, r);
#else
  return 1.0 / x;
#endif
}

// Numbers of eigenvalues of T below each of the P shifts s[0..P-1]; e2[i] =
// e[i]^2.  The P recurrences are independent, so their latency chains overlap.
template <int P>
TRIEIG_FN void tri_sturm(int n, const double* d, const double* e2, const double* s, double pivmin, int* cnt) {
  double q[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    q[k] = d[0] - s[k];
    if (fabs(q[k]) <= pivmin) q[k] = -pivmin;
    cnt[k] = q[k] < 0.0;
  }
  for (int i = 1; i < n; ++i) {
    const double di = d[i], ei = e2[i - 1];
#pragma unroll
    for (int k = 0; k < P; ++k) {
      q[k] = (di - s[k]) - ei * tri_rcp(q[k]);
      if (fabs(q[k]) <= pivmin) q[k] = -pivmin;
      cnt[k] += q[k] < 0.0;
    }
  }
}

// Multisection point idx (0 .. npts-1) of [lo, hi]: the npts interior points
// of npts + 1 equal parts.  Every lane forms any point from the same inputs,
// so the interval's new ends are the very values that were counted.
TRIEIG_FN double tri_point(double lo, double hi, int idx, int npts) {
  return lo + (hi - lo) * ((double)(idx + 1) / (double)(npts + 1));
}

// Stop when the interval is at the rounding of its end points: relative
// accuracy for every eigenvalue, not u |T| absolute.  The Gram matrices of a
// graded B reduce to tridiagonal forms whose small eigenvalues are determined
// far more accurately than u |T|, and the truncation rule compares singular
// values (square roots) by their absolute differences, so the small ones must
// come out as accurately as the form allows.  atol only ends the search at
// (numerically) zero eigenvalues.
TRIEIG_FN double tri_atol(double pivmin) { return 4.0 * pivmin; }
TRIEIG_FN bool tri_converged(double lo, double hi, double atol) {
  return hi - lo <= fmax(atol, 4.440892098500626e-16 * fmax(fabs(lo), fabs(hi)));
}

// Pseudo-random start entry in [-1, 1) for row i of eigenvector j (an
// integer mix of (i, j): xorshift-multiply rounds of the lowbias32 hash).
TRIEIG_FN double tri_start(int i, int j) {
  uint32_t h = (uint32_t)i * 0x9E3779B1u + (uint32_t)j * 0x85EBCA77u + 0x165667B1u;
  h ^= h >> 16;
  h *= 0x7feb352du;
  h ^= h >> 15;
  h *= 0x846ca68bu;
  h ^= h >> 16;
  return (double)(h >> 11) * (1.0 / 1048576.0) - 1.0;
}

// One solve (T - lam I) x = bscale * b.  b is read at b[i * bs] (or formed by
// tri_start when b is null), x written at x[i * xs] (x may alias b: the
// forward sweep reads all of b before the backward sweep writes x).  w holds
// the factor rows: w[(4 i + f) * ws], f = 0 the reciprocal of U's diagonal
// entry of row i (pivots below pert taken as +-pert), f = 1, 2 its two
// superdiagonals, f = 3 the eliminated right-hand side.  One reciprocal per
// row (the pivot's: it forms the multiplier and is stored for the backward
// sweep, which then only multiplies).  Returns sum_i x_i^2.
TRIEIG_FN double tri_invit_solve(int n, int j, const double* d, const double* e, double lam, double pert,
                                       double bscale, const double* b, int64_t bs, double* x, int64_t xs, double* w,
                                       int64_t ws) {
  auto pivot_rcp = [pert](double u, double ru) {  // 1 / u with |u| raised to pert
    return fabs(u) >= pert ? ru : (u < 0.0 ? -1.0 : 1.0) / pert;
  };
  // forward sweep: the carried row (p, q) sits in columns (i, i + 1)
  double p = d[0] - lam, q = n > 1 ? e[0] : 0.0;
  double y = bscale * (b ? b[0] : tri_start(0, j));
  for (int i0 = 0; i0 < n - 1; i0 += TRIEIG_BLK) {
    double bb[TRIEIG_BLK];
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) {
      const int i = i0 + t + 1;
      bb[t] = i < n ? (b ? b[(int64_t)i * bs] : tri_start(i, j)) : 0.0;
    }
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) {
      const int i = i0 + t;
      if (i < n - 1) {
        const double c = e[i], a1 = d[i + 1] - lam, b1 = i + 1 < n - 1 ? e[i + 1] : 0.0;
        const double beta = bscale * bb[t];
        double* wr = w + (int64_t)(4 * i) * ws;
        // partial pivoting: the pivot row is the carried one unless |c| > |p|
        const bool keep = fabs(p) >= fabs(c);
        const double piv = keep ? p : c;
        const double rp = piv != 0.0 ? tri_rcp(piv) : 0.0;
        const double m = (keep ? c : p) * rp;  // 0 when p = c = 0
        wr[0] = pivot_rcp(piv, rp);
        wr[ws] = keep ? q : a1;
        wr[2 * ws] = keep ? 0.0 : b1;
        wr[3 * ws] = keep ? y : beta;
        const double pn = keep ? a1 - m * q : q - m * a1;
        const double qn = keep ? b1 : -m * b1;
        const double yn = keep ? beta - m * y : y - m * beta;
        p = pn;
        q = qn;
        y = yn;
      }
    }
  }
  // backward sweep
  double ss = 0.0;
  double x1 = 0.0, x2 = 0.0;  // x_{i+1}, x_{i+2}
  {
    x1 = y * pivot_rcp(p, p != 0.0 ? tri_rcp(p) : 0.0);
    x[(int64_t)(n - 1) * xs] = x1;
    ss = x1 * x1;
  }
  for (int i0 = n - 2; i0 >= 0; i0 -= TRIEIG_BLK) {
    double r0[TRIEIG_BLK], r1[TRIEIG_BLK], r2[TRIEIG_BLK], r3[TRIEIG_BLK];
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) {
      const int i = i0 - t;
      const double* wr = w + (int64_t)(4 * (i >= 0 ? i : 0)) * ws;
      r0[t] = wr[0];
      r1[t] = wr[ws];
      r2[t] = wr[2 * ws];
      r3[t] = wr[3 * ws];
    }
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) {
      const int i = i0 - t;
      if (i >= 0) {
        const double xi = (r3[t] - r1[t] * x1 - r2[t] * x2) * r0[t];
        x[(int64_t)i * xs] = xi;
        ss += xi * xi;
        x2 = x1;
        x1 = xi;
      }
    }
  }
  return ss;
}

// The final scale of vector j: 1 / |x| with the sign that makes x . b > 0
// for the start vector b.  The sign of an inverse-iteration vector is that of
// (lam_j - lam) (b . z_j), and the bisection's lam falls on either side of
// lam_j at random, so without this the sign (the gauge of the renormalised
// bond) would jump between runs on inputs equal to rounding.
TRIEIG_FN double tri_sign_scale(int n, int j, const double* x, int64_t xs, double ss) {
  double dot = 0.0;
  for (int i0 = 0; i0 < n; i0 += TRIEIG_BLK) {
    double v[TRIEIG_BLK];
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) v[t] = i0 + t < n ? x[(int64_t)(i0 + t) * xs] : 0.0;
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t)
      if (i0 + t < n) dot += v[t] * tri_start(i0 + t, j);
  }
  return (dot < 0.0 ? -1.0 : 1.0) / sqrt(ss);
}

}  // namespace k
}  // namespace ipeps
