// Host driver of the tridiagonal eigensolver steps in
// ipeps.jl_amd/csrc/trieig_impl.h, compiled with g++ (IPEPS_HD empty) by
// tests/test_trieig_host.py.  It runs the same steps as tri_bisect_kernel and
// tri_invit_kernel: the 64 TRIEIG_PTS points of a wave's multisection step become a loop,
// the threads of the inverse iteration a loop over eigenvalues sharing the
// kernel's n x n work and output layouts.
#define IPEPS_HD
#include "trieig_impl.h"

#include <vector>

using namespace ipeps::k;

extern "C" int trieig_host(int n, const double* d, const double* e, double* lam, double* Zt, int* steps) {
  if (n < 1 || n > TRIEIG_MAXN) return 1;
  std::vector<double> e2(n, 0.0), ee(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) {
    ee[i] = e[i];
    e2[i] = e[i] * e[i];
  }
  double gl, gu, tnorm, pivmin;
  tri_bounds(n, d, ee.data(), gl, gu, tnorm, pivmin);
  const double atol = tri_atol(pivmin);
  for (int j = 0; j < n; ++j) {
    double lo = gl, hi = gu;
    int it = 0;
    constexpr int NP = 64 * TRIEIG_PTS;
    for (; it < TRIEIG_MAXIT && !tri_converged(lo, hi, atol); ++it) {
      int f = NP;
      for (int idx = 0; idx < NP && f == NP; ++idx) {
        const double s = tri_point(lo, hi, idx, NP);
        int cnt;
        tri_sturm<1>(n, d, e2.data(), &s, pivmin, &cnt);
        if (cnt > j) f = idx;
      }
      const double nlo = f > 0 ? tri_point(lo, hi, f - 1, NP) : lo;
      const double nhi = f < NP ? tri_point(lo, hi, f, NP) : hi;
      lo = nlo;
      hi = nhi;
    }
    if (steps) steps[j] = it;
    lam[j] = 0.5 * (lo + hi);
  }
  const double pert = fmax(2.220446049250313e-16 * tnorm, 1e-300);
  std::vector<double> w((size_t)4 * n * n);
  for (int j = 0; j < n; ++j) {
    double* x = Zt + j;
    double* wj = w.data() + j;
    double ss = tri_invit_solve(n, j, d, ee.data(), lam[j], pert, 1.0, nullptr, n, x, n, wj, n);
    for (int it = 1; it < TRIEIG_SOLVES; ++it)
      ss = tri_invit_solve(n, j, d, ee.data(), lam[j], pert, 1.0 / sqrt(ss), x, n, x, n, wj, n);
    const double sc = tri_sign_scale(n, j, x, n, ss);
    for (int i = 0; i < n; ++i) x[(int64_t)i * n] *= sc;
  }
  return 0;
}
