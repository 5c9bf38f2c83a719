// Permute, elementwise and reduction kernels (HBM-bound helpers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"
#include "permute_impl.h"
#include "blockchol_impl.h"
#include "tridiag_impl.h"
#include "trieig_impl.h"

namespace ipeps {
namespace k {

namespace {

constexpr int TPB = 256;

inline unsigned blocks_for(int64_t n, int per_thread = 1, unsigned cap = 65535u * 4) {
  int64_t b = (n + (int64_t)TPB * per_thread - 1) / ((int64_t)TPB * per_thread);
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

// complex helpers on double2
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ inline double2 cdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

template <typename T>
__device__ inline T conj_if(T v, bool c) { return v; }
template <>
__device__ inline double2 conj_if<double2>(double2 v, bool c) { return c ? cconj(v) : v; }

__device__ inline double absval(double v) { return fabs(v); }
__device__ inline double absval(float v) { return fabs((double)v); }
__device__ inline double absval(double2 v) { return hypot(v.x, v.y); }

struct PermP {
  int nd;
  int64_t odims[8];
  int64_t istr[8];  // input stride of output axis i
  int64_t n;
};

template <typename T>
__global__ __launch_bounds__(TPB) void permute_kernel(PermP p, const T* __restrict__ in, T* __restrict__ out,
                                                      int conj) {
  for (int64_t idx = (int64_t)blockIdx.x * TPB + threadIdx.x; idx < p.n; idx += (int64_t)gridDim.x * TPB) {
    int64_t rem = idx, off = 0;
#pragma unroll 8
    for (int d = 0; d < p.nd; ++d) {
      int64_t c = rem % p.odims[d];
      rem /= p.odims[d];
      off += c * p.istr[d];
    }
    out[idx] = conj_if(in[off], conj != 0);
  }
}

// Rows kernel (fastest output axis = fastest input axis); see permute_impl.h.
// A thread's PER elements are loaded unconditionally (out-of-range ones read
// element 0) before any is stored, so its loads are in flight together.
constexpr int PERM_ROWS_MAX = 4096;  // rb * chunk of a block (plan_permute's target)
template <typename T>
__global__ __launch_bounds__(TPB) void permute_rows_kernel(PermRowsP p, const T* __restrict__ in,
                                                           T* __restrict__ out, int conj) {
  constexpr int PER = PERM_ROWS_MAX / TPB;
  __shared__ uint32_t row_off[256];
  const uint32_t blk = blockIdx.x;
  perm_rows_phase1(p, blk, threadIdx.x, row_off);
  __syncthreads();
  const uint32_t total = p.rb * p.chunk;
  T v[PER];
  uint32_t oo[PER];
  bool ok[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t e = threadIdx.x + k * TPB;
    uint32_t io = 0, o = 0;
    ok[k] = e < total && perm_rows_elem(p, blk, e, row_off, io, o);
    oo[k] = o;
    v[k] = in[ok[k] ? io : 0u];
  }
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if (ok[k]) out[oo[k]] = conj_if(v[k], conj != 0);
}

// Tiled kernel (the input-contiguous axis j is not the first output axis);
// TS x TS tiles through LDS, see permute_impl.h.  Loads as in the rows kernel.
template <typename T, int TS>
__global__ __launch_bounds__(TPB) void permute_tiled_kernel(PermTileP p, const T* __restrict__ in,
                                                            T* __restrict__ out, int conj) {
  constexpr int PER = TS * TS / TPB;
  __shared__ T tile[TS][TS + 1];
  __shared__ uint32_t lead_off[TS];
  for (uint32_t rest = blockIdx.y; rest < p.nrest_total; rest += gridDim.y) {
    uint32_t x0, y0, in_rest, out_rest;
    perm_tile_block(p, blockIdx.x, rest, x0, y0, in_rest, out_rest);
    if (threadIdx.x < TS) lead_off[threadIdx.x] = y0 + threadIdx.x < p.wl ? perm_tile_lead(p, y0 + threadIdx.x) : 0u;
    __syncthreads();
    T v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t e = threadIdx.x + k * TPB, x = e % TS, y = e / TS;
      const bool ok = x0 + x < p.wj && y0 + y < p.wl;
      v[k] = in[in_rest + (ok ? lead_off[y] + x0 + x : 0u)];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t e = threadIdx.x + k * TPB;
      tile[e / TS][e % TS] = v[k];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < TS * TS; e += TPB) {
      const uint32_t y = e % TS, x = e / TS;
      if (x0 + x < p.wj && y0 + y < p.wl)
        out[out_rest + y0 + y + (x0 + x) * p.ostr_j] = conj_if(tile[y][x], conj != 0);
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void axpy_kernel(int64_t n, double a, const T* x, T* y);
template <>
__global__ void axpy_kernel<double>(int64_t n, double a, const double* x, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] += a * x[i];
}
template <>
__global__ void axpy_kernel<float>(int64_t n, double a, const float* x, float* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] += (float)a * x[i];
}
template <>
__global__ void axpy_kernel<double2>(int64_t n, double a, const double2* x, double2* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    double2 v = y[i];
    v.x += a * x[i].x;
    v.y += a * x[i].y;
    y[i] = v;
  }
}

// out = x + y (out-of-place cotangent accumulation)
template <typename T>
__global__ void add_kernel(int64_t n, const T* __restrict__ x, const T* __restrict__ y, T* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) out[i] = x[i] + y[i];
}
template <>
__global__ void add_kernel<double2>(int64_t n, const double2* __restrict__ x, const double2* __restrict__ y,
                                    double2* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    out[i] = make_double2(x[i].x + y[i].x, x[i].y + y[i].y);
}

template <typename T>
__device__ inline T scal(T v, double f) { return (T)(v * f); }
template <>
__device__ inline double2 scal<double2>(double2 v, double f) { return make_double2(v.x * f, v.y * f); }

template <typename T>
__global__ void scale_dev_kernel(int64_t n, const T* x, T* y, const double* s, int mode) {
  double f = mode == 0 ? 1.0 / s[0] : s[0];
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = scal(x[i], f);
}

__global__ void conj_kernel(int64_t n, const double2* x, double2* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = cconj(x[i]);
}

template <typename A, typename B>
__global__ void convert_kernel(int64_t n, const A* x, B* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = (B)x[i];
}

// real -> complex (zero imaginary part)
template <typename A>
__global__ void to_complex_kernel(int64_t n, const A* x, double2* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    y[i] = make_double2((double)x[i], 0.0);
}

template <typename T>
__global__ void scale_cols_kernel(int64_t rows, int64_t cols, const T* X, const double* s, T* Y, int rowwise) {
  int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    int64_t r = i % rows, c = i / rows;
    Y[i] = scal(X[i], rowwise ? s[r] : s[c]);
  }
}

__device__ inline float addv(float a, float b) { return a + b; }
__device__ inline double addv(double a, double b) { return a + b; }
__device__ inline double2 addv(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// per-column ratio a_c = num[c] / den[c] (0 where den[c] == 0)
__device__ inline double col_ratio(const double* num, const double* den, int64_t c) {
  const double d = den[c];
  return d != 0.0 ? num[c] / d : 0.0;
}
// X[:, c] += sign a_c P[:, c]
template <typename T>
__global__ void col_axpy_ratio_kernel(int64_t rows, int64_t cols, const double* num, const double* den, double sign,
                                      const T* P, T* X) {
  const int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    X[i] = addv(X[i], scal(P[i], sign * col_ratio(num, den, i / rows)));
}
// P[:, c] = R[:, c] + a_c P[:, c]
template <typename T>
__global__ void col_xpby_ratio_kernel(int64_t rows, int64_t cols, const double* num, const double* den, const T* R,
                                      T* P) {
  const int64_t n = rows * cols;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    P[i] = addv(R[i], scal(P[i], col_ratio(num, den, i / rows)));
}

// counter-based normal generator (splitmix64 + Box-Muller)
__device__ inline uint64_t splitmix(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ inline double u01(uint64_t x) { return ((x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
__device__ inline double normal_at(uint64_t seed, int64_t i) {
  uint64_t a = splitmix(seed ^ (uint64_t)(2 * i)), b = splitmix(seed ^ (uint64_t)(2 * i + 1) ^ 0x5bd1e995ull);
  double u1 = u01(a), u2 = u01(b);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
template <typename T>
__global__ void randn_kernel(int64_t n, uint64_t seed, T* out);
template <>
__global__ void randn_kernel<double>(int64_t n, uint64_t seed, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) out[i] = normal_at(seed, i);
}
template <>
__global__ void randn_kernel<float>(int64_t n, uint64_t seed, float* out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) out[i] = (float)normal_at(seed, i);
}
template <>
__global__ void randn_kernel<double2>(int64_t n, uint64_t seed, double2* out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    out[i] = make_double2(normal_at(seed, 2 * i), normal_at(seed, 2 * i + 1));
}

// seed of draw i of the stream keyed by base (the host's draw_seed)
__device__ inline uint64_t draw_seed_dev(uint64_t base, uint64_t i) {
  uint64_t z = base + (i + 1) * 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
template <typename T>
__global__ void randn_drawn_kernel(int64_t n, uint64_t base, const uint64_t* __restrict__ draws, uint64_t off,
                                   T* out) {
  const uint64_t seed = draw_seed_dev(base, draws[0] + off);
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    if constexpr (std::is_same<T, double2>::value) out[i] = make_double2(normal_at(seed, 2 * i), normal_at(seed, 2 * i + 1));
    else out[i] = (T)normal_at(seed, i);
  }
}
__global__ void counter_add_kernel(uint64_t* c, uint64_t add) {
  if (threadIdx.x == 0 && blockIdx.x == 0) c[0] += add;
}

// byte rows: 16-byte words when every address and length is a multiple of 16
__global__ void copy_rows16_kernel(uint4* __restrict__ dst, size_t dpitch, const uint4* __restrict__ src, size_t spitch,
                                   size_t wwords, size_t rows) {
  const size_t n = wwords * rows;
  for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
    const size_t r = i / wwords, w = i % wwords;
    dst[r * dpitch + w] = src[r * spitch + w];
  }
}
__global__ void copy_rows1_kernel(unsigned char* __restrict__ dst, size_t dpitch, const unsigned char* __restrict__ src,
                                  size_t spitch, size_t width, size_t rows) {
  const size_t n = width * rows;
  for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB) {
    const size_t r = i / width, w = i % width;
    dst[r * dpitch + w] = src[r * spitch + w];
  }
}
__global__ void zero_rows16_kernel(uint4* __restrict__ p, size_t pitch, size_t wwords, size_t rows) {
  const size_t n = wwords * rows;
  for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB)
    p[(i / wwords) * pitch + i % wwords] = make_uint4(0u, 0u, 0u, 0u);
}
__global__ void zero_rows1_kernel(unsigned char* __restrict__ p, size_t pitch, size_t width, size_t rows) {
  const size_t n = width * rows;
  for (size_t i = (size_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (size_t)gridDim.x * TPB)
    p[(i / width) * pitch + i % width] = 0;
}

// ---- reductions ----------------------------------------------------------
struct MaxIdx {
  double v;
  int64_t i;
};
__device__ inline MaxIdx better(MaxIdx a, MaxIdx b) {
  if (b.v > a.v || (b.v == a.v && b.i < a.i)) return b;
  return a;
}

template <typename T>
__global__ __launch_bounds__(TPB) void maxabs_partial(int64_t n, const T* x, double* pv, int64_t* pi) {
  __shared__ double sv[TPB];
  __shared__ int64_t si[TPB];
  MaxIdx m{-1.0, (int64_t)1 << 62};
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    m = better(m, MaxIdx{absval(x[i]), i});
  sv[threadIdx.x] = m.v;
  si[threadIdx.x] = m.i;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxIdx r = better(MaxIdx{sv[threadIdx.x], si[threadIdx.x]}, MaxIdx{sv[threadIdx.x + s], si[threadIdx.x + s]});
      sv[threadIdx.x] = r.v;
      si[threadIdx.x] = r.i;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

__global__ __launch_bounds__(TPB) void maxabs_final(int nb, const double* pv, const int64_t* pi, double* out,
                                                    int64_t* out_idx) {
  __shared__ double sv[TPB];
  __shared__ int64_t si[TPB];
  MaxIdx m{-1.0, (int64_t)1 << 62};
  for (int i = threadIdx.x; i < nb; i += TPB) m = better(m, MaxIdx{pv[i], pi[i]});
  sv[threadIdx.x] = m.v;
  si[threadIdx.x] = m.i;
  __syncthreads();
  for (int s = TPB / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      MaxIdx r = better(MaxIdx{sv[threadIdx.x], si[threadIdx.x]}, MaxIdx{sv[threadIdx.x + s], si[threadIdx.x + s]});
      sv[threadIdx.x] = r.v;
      si[threadIdx.x] = r.i;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sv[0];
    if (out_idx) out_idx[0] = si[0];
  }
}

__device__ inline double redot1(double a, double b) { return a * b; }
__device__ inline double redot1(float a, float b) { return (double)a * (double)b; }
__device__ inline double redot1(double2 a, double2 b) { return a.x * b.x + a.y * b.y; }

template <typename T>
__global__ __launch_bounds__(TPB) void redot_partial(int64_t n, const T* x, const T* y, double* part) {
  __shared__ double s[TPB];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) acc += redot1(x[i], y[i]);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = TPB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(TPB) void sum_final(int nb, const double* part, double* out) {
  __shared__ double s[TPB];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += TPB) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = TPB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = s[0];
}

template <typename T>
__global__ void renorm_back_kernel(int64_t n, const T* x, const T* gy, const double* mval, const int64_t* midx,
                                   const double* dotv, T* gx);
template <>
__global__ void renorm_back_kernel<double>(int64_t n, const double* x, const double* gy, const double* mval,
                                           const int64_t* midx, const double* dotv, double* gx) {
  double m = mval[0];
  int64_t im = midx[0];
  double coef = -dotv[0] / (m * m);
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    double v = gy[i] / m;
    if (i == im) v += coef * (x[i] / fabs(x[i]));
    gx[i] = v;
  }
}
template <>
__global__ void renorm_back_kernel<double2>(int64_t n, const double2* x, const double2* gy, const double* mval,
                                            const int64_t* midx, const double* dotv, double2* gx) {
  double m = mval[0];
  int64_t im = midx[0];
  double coef = -dotv[0] / (m * m);
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    double2 v = make_double2(gy[i].x / m, gy[i].y / m);
    if (i == im) {
      double a = hypot(x[i].x, x[i].y);
      v.x += coef * x[i].x / a;
      v.y += coef * x[i].y / a;
    }
    gx[i] = v;
  }
}
template <>
__global__ void renorm_back_kernel<float>(int64_t n, const float* x, const float* gy, const double* mval,
                                          const int64_t* midx, const double* dotv, float* gx) {
  double m = mval[0];
  int64_t im = midx[0];
  double coef = -dotv[0] / (m * m);
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) {
    double v = gy[i] / m;
    if (i == im) v += coef * (x[i] / fabs((double)x[i]));
    gx[i] = (float)v;
  }
}

// ---- SVD adjoint pieces -----------------------------------------------------
// F_ij = (s_j^2 - s_i^2) / ((s_j^2 - s_i^2)^2 + eta) with the squares rounded
// before the difference, as S2' .- S2 does (src/utility/autodiff.jl:46-47):
// a fused multiply-add here would leave a rounding residue instead of an exact
// 0 for equal singular values, which eta = 1e-12 then amplifies.
__device__ inline double Fij(const double* S, int64_t i, int64_t j, double eta) {
#pragma clang fp contract(off)
  double si = S[i] * S[i], sj = S[j] * S[j];
  double f = sj - si;
  return f / (f * f + eta);
}

__device__ inline double sinv_of(double s, double eta) {
#pragma clang fp contract(off)
  double s2 = s * s;
  return s / (s2 + eta);
}

// Gamma_ij = (J_ij + conj(J_ji)) s_j + s_i (K_ij + conj(K_ji)) + delta terms,
// J_ij = F_ij UdU_ij (j < n), K_ij = F_ij VdV_ij (j < n); UdU/VdV are ld x n.
template <typename T>
__global__ void svd_gamma_kernel(int64_t N, int64_t n, int64_t ld, const double* S, const T* UdU, const T* VdV,
                                 const double* dS, double eta, T* G);
template <>
__global__ void svd_gamma_kernel<double>(int64_t N, int64_t n, int64_t ld, const double* S, const double* UdU,
                                         const double* VdV, const double* dS, double eta, double* G) {
  int64_t tot = N * N;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < tot; t += (int64_t)gridDim.x * TPB) {
    int64_t i = t % N, j = t / N;
    double v = 0.0;
    double f = Fij(S, i, j, eta);
    double Jij = (j < n && UdU) ? f * UdU[i + j * ld] : 0.0;
    double Jji = (i < n && UdU) ? -f * UdU[j + i * ld] : 0.0;  // F_ji = -F_ij
    double Kij = (j < n && VdV) ? f * VdV[i + j * ld] : 0.0;
    double Kji = (i < n && VdV) ? -f * VdV[j + i * ld] : 0.0;
    v = (Jij + Jji) * S[j] + S[i] * (Kij + Kji);
    if (i == j && dS && i < n) v += dS[i];
    G[t] = v;
  }
}
template <>
__global__ void svd_gamma_kernel<double2>(int64_t N, int64_t n, int64_t ld, const double* S, const double2* UdU,
                                          const double2* VdV, const double* dS, double eta, double2* G) {
  int64_t tot = N * N;
  for (int64_t t = (int64_t)blockIdx.x * TPB + threadIdx.x; t < tot; t += (int64_t)gridDim.x * TPB) {
    int64_t i = t % N, j = t / N;
    double f = Fij(S, i, j, eta);
    double2 z = make_double2(0.0, 0.0);
    double2 Jij = (j < n && UdU) ? scal(UdU[i + j * ld], f) : z;
    double2 Jji = (i < n && UdU) ? scal(UdU[j + i * ld], -f) : z;
    double2 Kij = (j < n && VdV) ? scal(VdV[i + j * ld], f) : z;
    double2 Kji = (i < n && VdV) ? scal(VdV[j + i * ld], -f) : z;
    double2 a = make_double2(Jij.x + Jji.x, Jij.y - Jji.y);  // J + J^H
    double2 b = make_double2(Kij.x + Kji.x, Kij.y - Kji.y);  // K + K^H
    double2 v = make_double2(a.x * S[j] + S[i] * b.x, a.y * S[j] + S[i] * b.y);
    if (i == j && i < n) {
      if (dS) v.x += dS[i];
      if (UdU) {
        double sinv = sinv_of(S[i], eta);
        v.y += UdU[i + i * ld].y * sinv;
      }
    }
    G[t] = v;
  }
}

__global__ void sinv_kernel(int64_t n, const double* S, double eta, double* Sinv) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    Sinv[i] = sinv_of(S[i], eta);
}

__global__ void inv_sqrt_kernel(int64_t n, const double* s, double* out) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) out[i] = 1.0 / sqrt(s[i]);
}
__global__ void inv_sqrt_back_kernel(int64_t n, const double* s, const double* g, double* gs) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    gs[i] += -0.5 * g[i] / (s[i] * sqrt(s[i]));
}

// out[c] = Re sum_r conj(X[r,c]) Y[r,c]  (or over columns for rowwise)
template <typename T>
__global__ __launch_bounds__(TPB) void col_redot_kernel(int64_t rows, int64_t cols, const T* X, const T* Y,
                                                        double* out, int rowwise) {
  __shared__ double s[TPB];
  int64_t c = blockIdx.x;
  double acc = 0.0;
  if (!rowwise) {
    for (int64_t r = threadIdx.x; r < rows; r += TPB) acc += redot1(X[r + c * rows], Y[r + c * rows]);
  } else {
    for (int64_t q = threadIdx.x; q < cols; q += TPB) acc += redot1(X[c + q * rows], Y[c + q * rows]);
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int st = TPB / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = s[0];
}

__global__ void real_to_kernel_d(int64_t n, const double* x, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = x[i];
}
__global__ void real_to_kernel_c(int64_t n, const double* x, double2* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = make_double2(x[i], 0.0);
}
__global__ void real_of_kernel_c(int64_t n, const double2* x, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = x[i].x;
}

// ---- energy (tiny, one thread) -------------------------------------------------
__device__ inline double2 ld_el(const void* p, int64_t i, int cplx) {
  if (cplx) return ((const double2*)p)[i];
  return make_double2(((const double*)p)[i], 0.0);
}

__global__ void energy2_kernel(int64_t m, const void* rh, const void* rv, const double* H1, const double* H2,
                               double* out, int cplx) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double2* h1 = (const double2*)H1;
  const double2* h2 = (const double2*)H2;
  double2 ah = make_double2(0, 0), av = make_double2(0, 0), bh = make_double2(0, 0), bv = make_double2(0, 0);
  for (int64_t i = 0; i < m; ++i) {
    bh.x += ld_el(rh, i + i * m, cplx).x; bh.y += ld_el(rh, i + i * m, cplx).y;
    bv.x += ld_el(rv, i + i * m, cplx).x; bv.y += ld_el(rv, i + i * m, cplx).y;
    for (int64_t j = 0; j < m; ++j) {
      // tr(H rho) = sum_ij H_ij rho_ji
      double2 t1 = cmul(h1[i + j * m], ld_el(rh, j + i * m, cplx));
      double2 t2 = cmul(h2[i + j * m], ld_el(rv, j + i * m, cplx));
      ah.x += t1.x; ah.y += t1.y;
      av.x += t2.x; av.y += t2.y;
    }
  }
  double2 E1 = cdiv(ah, bh), E2 = cdiv(av, bv);
  out[0] = E1.x + E2.x;
  out[1] = E1.y + E2.y;
  out[2] = 0.5 * (bh.x + bv.x);
  out[3] = 0.5 * (bh.y + bv.y);
  out[4] = bh.x; out[5] = bh.y; out[6] = bv.x; out[7] = bv.y;
}

__global__ void energy2_back_kernel(int64_t m, const void* rh, const void* rv, const double* H1, const double* H2,
                                    void* gh, void* gv, int cplx) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * m * m) return;
  int which = t >= m * m;
  int64_t e = which ? t - m * m : t;
  const void* rho = which ? rv : rh;
  const double2* H = (const double2*)(which ? H2 : H1);
  // recompute a = tr(H rho), b = tr(rho) (m <= 16: cheap)
  double2 a = make_double2(0, 0), b = make_double2(0, 0);
  for (int64_t i = 0; i < m; ++i) {
    double2 d = ld_el(rho, i + i * m, cplx);
    b.x += d.x; b.y += d.y;
    for (int64_t j = 0; j < m; ++j) {
      double2 tt = cmul(H[i + j * m], ld_el(rho, j + i * m, cplx));
      a.x += tt.x; a.y += tt.y;
    }
  }
  // element rho_{jj', ii'} with column-major index e = r + c*m, r=row, c=col
  int64_t r = e % m, c = e / m;
  // dE/drho_rc = H_cr / b - delta_rc a / b^2
  double2 g = cdiv(H[c + r * m], b);
  if (r == c) {
    double2 ab2 = cdiv(cdiv(a, b), b);
    g.x -= ab2.x;
    g.y -= ab2.y;
  }
  // cotangent of rho for L = Re E: conj(dE/drho) (complex) or Re (real)
  if (cplx) ((double2*)(which ? gv : gh))[e] = cconj(g);
  else ((double*)(which ? gv : gh))[e] = g.x;
}

template <typename T>
__global__ void scale_cplx_kernel(int64_t n, double zr, double zi, const T* x, T* y);
template <>
__global__ void scale_cplx_kernel<double2>(int64_t n, double zr, double zi, const double2* x, double2* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    y[i] = cmul(make_double2(zr, zi), x[i]);
}
template <>
__global__ void scale_cplx_kernel<double>(int64_t n, double zr, double zi, const double* x, double* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB) y[i] = zr * x[i];
}
template <>
__global__ void scale_cplx_kernel<float>(int64_t n, double zr, double zi, const float* x, float* y) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n; i += (int64_t)gridDim.x * TPB)
    y[i] = (float)(zr * x[i]);
}

__global__ void es_energy_kernel(int64_t m, const double2* rh1, const double2* rv1, const double2* rh4,
                                 const double2* rv4, const double2* H1, const double2* H2, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double Nh = 0, Nv = 0;
  double2 ah = make_double2(0, 0), av = make_double2(0, 0);
  for (int64_t i = 0; i < m; ++i) {
    Nh += rh1[i + i * m].x;
    Nv += rv1[i + i * m].x;
    for (int64_t j = 0; j < m; ++j) {
      double2 t1 = cmul(H1[i + j * m], rh4[j + i * m]);
      double2 t2 = cmul(H2[i + j * m], rv4[j + i * m]);
      ah.x += t1.x; ah.y += t1.y;
      av.x += t2.x; av.y += t2.y;
    }
  }
  out[0] = ah.x / Nh + av.x / Nv;
  out[1] = ah.y / Nh + av.y / Nv;
  out[2] = Nh;
  out[3] = Nv;
}

__global__ void es_energy_back_kernel(int64_t m, const double2* H1, const double2* H2, const double* out, double2* gh,
                                      double2* gv) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m * m) return;
  int64_t r = t % m, c = t / m;
  // d E4 / d rho_rc = H_cr / N ; cotangent for L = Re E4 is its conjugate
  gh[t] = cconj(make_double2(H1[c + r * m].x / out[2], H1[c + r * m].y / out[2]));
  gv[t] = cconj(make_double2(H2[c + r * m].x / out[3], H2[c + r * m].y / out[3]));
}

// ---- projector-solver helpers ------------------------------------------------
__device__ inline double dev_from_eye(double g, bool d) { return fabs(d ? g - 1.0 : g); }
__device__ inline double dev_from_eye(float g, bool d) { return fabs(d ? (double)g - 1.0 : (double)g); }
__device__ inline double dev_from_eye(double2 g, bool d) {
  double re = d ? g.x - 1.0 : g.x;
  return sqrt(re * re + g.y * g.y);
}

// one workgroup; a NaN anywhere makes the result +inf
constexpr int EYE_TPB = 1024;
template <typename T>
__global__ __launch_bounds__(EYE_TPB) void eye_dev_kernel(int n, const T* G, double* out) {
  __shared__ double red[EYE_TPB];
  double m = 0.0;
  // flat over the n x n elements (n <= 46340, so n^2 fits 31 bits): thread t
  // covers i = t, t + 1024, ...; eight loads issued before their maxima
  const int nn = n * n;
  constexpr int U = 8;
  for (int base = 0; base < nn; base += U * EYE_TPB) {
    T g[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * EYE_TPB + (int)threadIdx.x;
      g[u] = G[i < nn ? i : 0];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int i = base + u * EYE_TPB + (int)threadIdx.x;
      if (i < nn) {
        const int col = i / n, r = i - col * n;
        const double v = dev_from_eye(g[u], r == col);
        m = isnan(v) ? INFINITY : fmax(m, v);
      }
    }
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = EYE_TPB / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

template <typename T>
__global__ void eig_to_svd_kernel(int64_t n, const double* lam, const T* W, double* s, T* Wr, double* sinv) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n * n; i += (int64_t)gridDim.x * TPB) {
    int64_t r = i % n, col = i / n;
    Wr[i] = W[r + (n - 1 - col) * n];
    if (r == 0) {
      double l = lam[n - 1 - col];
      double v = l > 0.0 ? sqrt(l) : 0.0;
      s[col] = v;
      sinv[col] = v > 0.0 ? 1.0 / v : 0.0;
    }
  }
}

__device__ inline void ns_el(double& g, bool d) { g = (d ? 1.5 : 0.0) - 0.5 * g; }
__device__ inline void ns_el(float& g, bool d) { g = (d ? 1.5f : 0.0f) - 0.5f * g; }
__device__ inline void ns_el(double2& g, bool d) { g = make_double2((d ? 1.5 : 0.0) - 0.5 * g.x, -0.5 * g.y); }

template <typename T>
__global__ void ns_coeff_kernel(int64_t n, T* G) {
  for (int64_t i = (int64_t)blockIdx.x * TPB + threadIdx.x; i < n * n; i += (int64_t)gridDim.x * TPB)
    ns_el(G[i], (i % n) == (i / n));
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void scale_cplx(Ctx& c, DType dt, int64_t n, double zr, double zi, const void* x, void* y) {
  if (n <= 0) return;
  IPEPS_CHECK(dt == DType::C128 || zi == 0.0, "scale_cplx: complex factor on real data");
  if (dt == DType::C128)
    hipLaunchKernelGGL(scale_cplx_kernel<double2>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, zr, zi,
                       (const double2*)x, (double2*)y);
  else if (dt == DType::F64)
    hipLaunchKernelGGL(scale_cplx_kernel<double>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, zr, zi,
                       (const double*)x, (double*)y);
  else
    hipLaunchKernelGGL(scale_cplx_kernel<float>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, zr, zi,
                       (const float*)x, (float*)y);
  HIP_CHECK(hipGetLastError());
}

void es_energy(Ctx& c, int64_t m, const void* rh1, const void* rv1, const void* rh4, const void* rv4, const double* H1,
               const double* H2, double* out) {
  hipLaunchKernelGGL(es_energy_kernel, dim3(1), dim3(64), 0, c.stream, m, (const double2*)rh1, (const double2*)rv1,
                     (const double2*)rh4, (const double2*)rv4, (const double2*)H1, (const double2*)H2, out);
  HIP_CHECK(hipGetLastError());
}

void es_energy_back(Ctx& c, int64_t m, const double* H1, const double* H2, const double* out, void* g_rh4,
                    void* g_rv4) {
  int64_t n = m * m;
  hipLaunchKernelGGL(es_energy_back_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, m,
                     (const double2*)H1, (const double2*)H2, out, (double2*)g_rh4, (double2*)g_rv4);
  HIP_CHECK(hipGetLastError());
}
static const char* dtype_name(DType dt) {
  return dt == DType::F64 ? "double" : dt == DType::C128 ? "HIP_vector_type<double, 2u>" : "float";
}

#define DISPATCH3(dt, KERNEL, ...)                                                                \
  do {                                                                                            \
    if ((dt) == DType::F64) { KERNEL(double, __VA_ARGS__); }                                      \
    else if ((dt) == DType::C128) { KERNEL(double2, __VA_ARGS__); }                               \
    else { KERNEL(float, __VA_ARGS__); }                                                          \
  } while (0)

void permute(Ctx& c, DType dt, const std::vector<int64_t>& in_dims, const std::vector<int>& perm, const void* in,
             void* out, bool conj) {
  int nd = (int)in_dims.size();
  IPEPS_CHECK(nd <= 8 && (int)perm.size() == nd, "permute: rank > 8 or bad perm");
  int64_t n = 1;
  for (auto d : in_dims) n *= d;
  if (n == 0) return;
  std::vector<int64_t> istride(nd);
  int64_t s = 1;
  for (int i = 0; i < nd; ++i) { istride[i] = s; s *= in_dims[i]; }
  // Drop size-1 axes and merge axes that stay adjacent (in order) in both.
  std::vector<int64_t> od, is;
  for (int i = 0; i < nd; ++i) {
    int64_t d = in_dims[perm[i]];
    if (d == 1) continue;
    if (!od.empty() && is.back() * od.back() == istride[perm[i]]) {
      od.back() *= d;
    } else {
      od.push_back(d);
      is.push_back(istride[perm[i]]);
    }
  }
  const int cj = conj && dt == DType::C128;
  if (od.size() <= 1) {  // plain copy
    if (cj) conj_copy(c, dt, n, in, out);
    else dev_copy(c, out, in, (size_t)n * dtype_size(dt));
    return;
  }
  if (n < ((int64_t)1 << 31)) {
    PermRowsP rp;
    PermTileP tp;
    const int esize = (int)dtype_size(dt);
    // A stride-1 fastest axis whose extent and the other input strides are
    // multiples of a 16-byte word: the same permutation of 16-byte elements
    // (bit copies; conjugation never applies to real data)
    bool words = false;
    std::vector<int64_t> odw = od, isw = is;
    if (!cj && is[0] == 1 && esize < 16 && ((uintptr_t)in & 15) == 0 && ((uintptr_t)out & 15) == 0) {
      const int64_t g = 16 / esize;
      words = od[0] % g == 0;
      for (size_t a = 1; a < is.size(); ++a) words = words && is[a] % g == 0;
      if (words) {
        odw[0] /= g;
        for (size_t a = 1; a < is.size(); ++a) isw[a] /= g;
      }
    }
    const int kind = words ? plan_permute((int)odw.size(), odw.data(), isw.data(), 16, rp, tp)
                           : plan_permute((int)od.size(), od.data(), is.data(), esize, rp, tp);
    if (kind >= 0) {
      std::string kname = std::string(kind == 0 ? "permute_rows_kernel<" : "permute_tiled_kernel<") + dtype_name(dt) + ">";
      if (c.timing_detail) {
        kname += " od=";
        for (size_t i = 0; i < od.size(); ++i) kname += std::to_string(od[i]) + (i + 1 < od.size() ? "," : "");
        kname += " is=";
        for (size_t i = 0; i < is.size(); ++i) kname += std::to_string(is[i]) + (i + 1 < is.size() ? "," : "");
      }
      TimedLaunch tl(c, kname, 0.0, 2.0 * n * dtype_size(dt));
      if (kind == 0) {
        const unsigned nb = (unsigned)(((int64_t)rp.nrows + rp.rb - 1) / rp.rb * rp.nchunk);
IPEPS_CHECK(rp.rb * rp.chunk <= (uint32_t)PERM_ROWS_MAX, "permute: rows block too large");
#define RK(T, ...) hipLaunchKernelGGL(permute_rows_kernel<T>, dim3(nb), dim3(TPB), 0, c.stream, rp, (const T*)in, (T*)out, cj)
        if (words) RK(float4, 0);
        else DISPATCH3(dt, RK, 0);
#undef RK
      } else {
        dim3 grid(tp.ntx * tp.nty, std::min<uint32_t>(tp.nrest_total, 65535u));
        if (esize >= 16)
          hipLaunchKernelGGL((permute_tiled_kernel<double2, 32>), grid, dim3(TPB), 0, c.stream, tp, (const double2*)in,
                             (double2*)out, cj);
        else if (esize == 8)
          hipLaunchKernelGGL((permute_tiled_kernel<double, 64>), grid, dim3(TPB), 0, c.stream, tp, (const double*)in,
                             (double*)out, cj);
        else
          hipLaunchKernelGGL((permute_tiled_kernel<float, 64>), grid, dim3(TPB), 0, c.stream, tp, (const float*)in,
                             (float*)out, cj);
      }
      HIP_CHECK(hipGetLastError());
      return;
    }
  }
  PermP p{};
  p.nd = (int)od.size();
  for (int i = 0; i < p.nd; ++i) { p.odims[i] = od[i]; p.istr[i] = is[i]; }
  p.n = n;
  std::string pname = std::string("permute_kernel<") + dtype_name(dt) + ">";
  if (c.timing_detail) {
    pname += " od=";
    for (int i = 0; i < p.nd; ++i) pname += std::to_string(od[i]) + (i + 1 < p.nd ? "," : "");
    pname += " is=";
    for (int i = 0; i < p.nd; ++i) pname += std::to_string(is[i]) + (i + 1 < p.nd ? "," : "");
  }
  TimedLaunch tl(c, pname, 0.0, 2.0 * n * dtype_size(dt));
#define PK(T, ...) hipLaunchKernelGGL(permute_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, p, (const T*)in, (T*)out, cj)
  DISPATCH3(dt, PK, 0);
#undef PK
  HIP_CHECK(hipGetLastError());
}

void fill_zero(Ctx& c, void* p, size_t bytes) { dev_zero(c, p, bytes); }

void axpy(Ctx& c, DType dt, int64_t n, double a, const void* x, void* y) {
  if (n <= 0) return;
#define AK(T, ...) hipLaunchKernelGGL(axpy_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, a, (const T*)x, (T*)y)
  DISPATCH3(dt, AK, 0);
#undef AK
  HIP_CHECK(hipGetLastError());
}

void add(Ctx& c, DType dt, int64_t n, const void* x, const void* y, void* out) {
  if (n <= 0) return;
#define AK(T, ...) hipLaunchKernelGGL(add_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const T*)x, (const T*)y, (T*)out)
  DISPATCH3(dt, AK, 0);
#undef AK
  HIP_CHECK(hipGetLastError());
}

void scale_by_dev(Ctx& c, DType dt, int64_t n, const void* x, void* y, const double* s, int mode) {
  if (n <= 0) return;
#define SK(T, ...) hipLaunchKernelGGL(scale_dev_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const T*)x, (T*)y, s, mode)
  DISPATCH3(dt, SK, 0);
#undef SK
  HIP_CHECK(hipGetLastError());
}

void conj_copy(Ctx& c, DType dt, int64_t n, const void* x, void* y) {
  if (n <= 0) return;
  if (dt != DType::C128) {
    if (x != y) HIP_CHECK(hipMemcpyAsync(y, x, (size_t)n * dtype_size(dt), hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  hipLaunchKernelGGL(conj_kernel, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const double2*)x, (double2*)y);
  HIP_CHECK(hipGetLastError());
}

void convert(Ctx& c, DType from, DType to, int64_t n, const void* x, void* y) {
  if (n <= 0) return;
  if (from == to) {
    HIP_CHECK(hipMemcpyAsync(y, x, (size_t)n * dtype_size(to), hipMemcpyDeviceToDevice, c.stream));
    return;
  }
  if (from == DType::F64 && to == DType::F32)
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const double*)x, (float*)y);
  else if (from == DType::F32 && to == DType::F64)
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const float*)x, (double*)y);
  else if (from == DType::F64 && to == DType::C128)
    hipLaunchKernelGGL((to_complex_kernel<double>), dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const double*)x, (double2*)y);
  else if (from == DType::F32 && to == DType::C128)
    hipLaunchKernelGGL((to_complex_kernel<float>), dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const float*)x, (double2*)y);
  else
    throw Error(IPEPS_ERR_ARG, "convert: unsupported dtype pair");
  HIP_CHECK(hipGetLastError());
}

void scale_cols(Ctx& c, DType dt, int64_t rows, int64_t cols, const void* X, const double* s, void* Y, bool rowwise) {
  int64_t n = rows * cols;
  if (n <= 0) return;
#define CK(T, ...) hipLaunchKernelGGL(scale_cols_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, rows, cols, (const T*)X, s, (T*)Y, (int)rowwise)
  DISPATCH3(dt, CK, 0);
#undef CK
  HIP_CHECK(hipGetLastError());
}

void col_axpy_ratio(Ctx& c, DType dt, int64_t rows, int64_t cols, const double* num, const double* den, double sign,
                    const void* P, void* X) {
  const int64_t n = rows * cols;
  if (n <= 0) return;
#define CK(T, ...) hipLaunchKernelGGL(col_axpy_ratio_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, rows, cols, num, den, sign, (const T*)P, (T*)X)
  DISPATCH3(dt, CK, 0);
#undef CK
  HIP_CHECK(hipGetLastError());
}

void col_xpby_ratio(Ctx& c, DType dt, int64_t rows, int64_t cols, const double* num, const double* den, const void* R,
                    void* P) {
  const int64_t n = rows * cols;
  if (n <= 0) return;
#define CK(T, ...) hipLaunchKernelGGL(col_xpby_ratio_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, rows, cols, num, den, (const T*)R, (T*)P)
  DISPATCH3(dt, CK, 0);
#undef CK
  HIP_CHECK(hipGetLastError());
}

void randn(Ctx& c, DType dt, int64_t n, uint64_t seed, void* out) {
  if (n <= 0) return;
#define RK(T, ...) hipLaunchKernelGGL(randn_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, seed, (T*)out)
  DISPATCH3(dt, RK, 0);
#undef RK
  HIP_CHECK(hipGetLastError());
}

void randn_drawn(Ctx& c, DType dt, int64_t n, uint64_t base, const uint64_t* draws, uint64_t off, void* out) {
  if (n <= 0) return;
#define RK(T, ...) hipLaunchKernelGGL(randn_drawn_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, base, draws, off, (T*)out)
  DISPATCH3(dt, RK, 0);
#undef RK
  HIP_CHECK(hipGetLastError());
}

void counter_add(Ctx& c, uint64_t* counter, uint64_t add) {
  hipLaunchKernelGGL(counter_add_kernel, dim3(1), dim3(64), 0, c.stream, counter, add);
  HIP_CHECK(hipGetLastError());
}

static bool al16z(const void* p) { return ((uintptr_t)p & 15) == 0; }

void copy_bytes_2d(Ctx& c, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t rows) {
  if (!width || !rows) return;
  const size_t n = width * rows;
  if (al16z(dst) && al16z(src) && dpitch % 16 == 0 && spitch % 16 == 0 && width % 16 == 0) {
    const size_t nw = n / 16;
    hipLaunchKernelGGL(copy_rows16_kernel, dim3((unsigned)std::min<size_t>((nw + TPB - 1) / TPB, 65536)), dim3(TPB), 0,
                       c.stream, (uint4*)dst, dpitch / 16, (const uint4*)src, spitch / 16, width / 16, rows);
  } else {
    hipLaunchKernelGGL(copy_rows1_kernel, dim3((unsigned)std::min<size_t>((n + TPB - 1) / TPB, 65536)), dim3(TPB), 0,
                       c.stream, (unsigned char*)dst, dpitch, (const unsigned char*)src, spitch, width, rows);
  }
  HIP_CHECK(hipGetLastError());
}

void copy_bytes(Ctx& c, void* dst, const void* src, size_t bytes) {
  if (!bytes) return;
  if (bytes % 16 == 0 && al16z(dst) && al16z(src)) {
    // as 1024-byte rows (any multiple of 16 that divides bytes)
    size_t w = 1024;
    while (bytes % w) w >>= 1;
    copy_bytes_2d(c, dst, w, src, w, w, bytes / w);
  } else {
    copy_bytes_2d(c, dst, bytes, src, bytes, bytes, 1);
  }
}

void zero_bytes_2d(Ctx& c, void* p, size_t pitch, size_t width, size_t rows) {
  if (!width || !rows) return;
  const size_t n = width * rows;
  if (al16z(p) && pitch % 16 == 0 && width % 16 == 0) {
    const size_t nw = n / 16;
    hipLaunchKernelGGL(zero_rows16_kernel, dim3((unsigned)std::min<size_t>((nw + TPB - 1) / TPB, 65536)), dim3(TPB), 0,
                       c.stream, (uint4*)p, pitch / 16, width / 16, rows);
  } else {
    hipLaunchKernelGGL(zero_rows1_kernel, dim3((unsigned)std::min<size_t>((n + TPB - 1) / TPB, 65536)), dim3(TPB), 0,
                       c.stream, (unsigned char*)p, pitch, width, rows);
  }
  HIP_CHECK(hipGetLastError());
}

void maxabs(Ctx& c, DType dt, int64_t n, const void* x, double* out, int64_t* out_idx) {
  IPEPS_CHECK(n > 0, "maxabs of an empty tensor");
  unsigned nb = std::min<unsigned>(blocks_for(n, 8), 1024u);
  std::shared_ptr<Buf> scratch(new Buf{c.pool->alloc(nb * 16), nb * 16, c.pool.get()});
  double* pv = (double*)scratch->p;
  int64_t* pi = (int64_t*)((char*)scratch->p + nb * 8);
#define MK(T, ...) hipLaunchKernelGGL(maxabs_partial<T>, dim3(nb), dim3(TPB), 0, c.stream, n, (const T*)x, pv, pi)
  DISPATCH3(dt, MK, 0);
#undef MK
  hipLaunchKernelGGL(maxabs_final, dim3(1), dim3(TPB), 0, c.stream, (int)nb, pv, pi, out, out_idx);
  HIP_CHECK(hipGetLastError());
}

void re_dot(Ctx& c, DType dt, int64_t n, const void* x, const void* y, double* out) {
  unsigned nb = std::min<unsigned>(blocks_for(n, 8), 1024u);
  std::shared_ptr<Buf> scratch(new Buf{c.pool->alloc(nb * 8), nb * 8, c.pool.get()});
  double* part = (double*)scratch->p;
  if (n <= 0) {
    fill_zero(c, out, 8);
    return;
  }
#define DK(T, ...) hipLaunchKernelGGL(redot_partial<T>, dim3(nb), dim3(TPB), 0, c.stream, n, (const T*)x, (const T*)y, part)
  DISPATCH3(dt, DK, 0);
#undef DK
  hipLaunchKernelGGL(sum_final, dim3(1), dim3(TPB), 0, c.stream, (int)nb, part, out);
  HIP_CHECK(hipGetLastError());
}

void renorm_back(Ctx& c, DType dt, int64_t n, const void* x, const void* gy, const double* mval, const int64_t* midx,
                 const double* dotv, void* gx) {
#define BK(T, ...) hipLaunchKernelGGL(renorm_back_kernel<T>, dim3(blocks_for(n, 4)), dim3(TPB), 0, c.stream, n, (const T*)x, (const T*)gy, mval, midx, dotv, (T*)gx)
  DISPATCH3(dt, BK, 0);
#undef BK
  HIP_CHECK(hipGetLastError());
}

void svd_gamma(Ctx& c, DType dt, int64_t N, int64_t n, const double* S, const void* UdU, const void* VdV,
               const double* dS, double eta, void* Gamma) {
  int64_t tot = N * N;
  if (dt == DType::F64)
    hipLaunchKernelGGL(svd_gamma_kernel<double>, dim3(blocks_for(tot, 4)), dim3(TPB), 0, c.stream, N, n, N, S,
                       (const double*)UdU, (const double*)VdV, dS, eta, (double*)Gamma);
  else if (dt == DType::C128)
    hipLaunchKernelGGL(svd_gamma_kernel<double2>, dim3(blocks_for(tot, 4)), dim3(TPB), 0, c.stream, N, n, N, S,
                       (const double2*)UdU, (const double2*)VdV, dS, eta, (double2*)Gamma);
  else
    throw Error(IPEPS_ERR_ARG, "svd_gamma: FP64 or C128 only");
  HIP_CHECK(hipGetLastError());
}

void svd_gamma_thin(Ctx& c, DType dt, int64_t n, const double* S, const void* UdU, const void* VdV, const double* dS,
                    double eta, void* Gamma, double* Sinv) {
  svd_gamma(c, dt, n, n, S, UdU, VdV, dS, eta, Gamma);
  hipLaunchKernelGGL(sinv_kernel, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, S, eta, Sinv);
  HIP_CHECK(hipGetLastError());
}

void svd_sinv(Ctx& c, int64_t n, const double* S, double eta, double* Sinv) {
  if (n <= 0) return;
  hipLaunchKernelGGL(sinv_kernel, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, S, eta, Sinv);
  HIP_CHECK(hipGetLastError());
}

void inv_sqrt(Ctx& c, int64_t n, const double* s, double* out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_sqrt_kernel, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, s, out);
  HIP_CHECK(hipGetLastError());
}

void inv_sqrt_back(Ctx& c, int64_t n, const double* s, const double* g, double* gs) {
  if (n <= 0) return;
  hipLaunchKernelGGL(inv_sqrt_back_kernel, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, s, g, gs);
  HIP_CHECK(hipGetLastError());
}

void col_re_dot(Ctx& c, DType dt, int64_t rows, int64_t cols, const void* X, const void* Y, double* out,
                bool rowwise) {
  int64_t nb = rowwise ? rows : cols;
  if (nb <= 0) return;
#define RD(T, ...) hipLaunchKernelGGL(col_redot_kernel<T>, dim3((unsigned)nb), dim3(TPB), 0, c.stream, rows, cols, (const T*)X, (const T*)Y, out, (int)rowwise)
  DISPATCH3(dt, RD, 0);
#undef RD
  HIP_CHECK(hipGetLastError());
}

void real_to(Ctx& c, DType dt, int64_t n, const double* x, void* y) {
  if (n <= 0) return;
  if (dt == DType::F64)
    hipLaunchKernelGGL(real_to_kernel_d, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, x, (double*)y);
  else if (dt == DType::C128)
    hipLaunchKernelGGL(real_to_kernel_c, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, x, (double2*)y);
  else
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, x, (float*)y);
  HIP_CHECK(hipGetLastError());
}

void real_of(Ctx& c, DType dt, int64_t n, const void* x, double* y) {
  if (n <= 0) return;
  if (dt == DType::F64)
    HIP_CHECK(hipMemcpyAsync(y, x, (size_t)n * 8, hipMemcpyDeviceToDevice, c.stream));
  else if (dt == DType::C128)
    hipLaunchKernelGGL(real_of_kernel_c, dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, (const double2*)x, y);
  else
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks_for(n)), dim3(TPB), 0, c.stream, n, (const float*)x, y);
  HIP_CHECK(hipGetLastError());
}

void energy2(Ctx& c, DType dt, int64_t m, const void* rho_h, const void* rho_v, const double* H1, const double* H2,
             double* out) {
  IPEPS_CHECK(dt != DType::F32, "energy2: FP64 reduction expects F64 or C128 density matrices");
  hipLaunchKernelGGL(energy2_kernel, dim3(1), dim3(64), 0, c.stream, m, rho_h, rho_v, H1, H2, out,
                     (int)(dt == DType::C128));
  HIP_CHECK(hipGetLastError());
}

void energy2_back(Ctx& c, DType dt, int64_t m, const void* rho_h, const void* rho_v, const double* H1,
                  const double* H2, void* grho_h, void* grho_v) {
  IPEPS_CHECK(dt != DType::F32, "energy2_back: F64 or C128 only");
  int64_t n = 2 * m * m;
  hipLaunchKernelGGL(energy2_back_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, m, rho_h, rho_v,
                     H1, H2, grho_h, grho_v, (int)(dt == DType::C128));
  HIP_CHECK(hipGetLastError());
}

void eye_dev(Ctx& c, DType dt, int64_t n, const void* G, double* out) {
  IPEPS_CHECK(n >= 1, "eye_dev: empty matrix");
  IPEPS_CHECK(n <= 46340, "eye_dev: n too large");
#define EK(T, ...) hipLaunchKernelGGL(eye_dev_kernel<T>, dim3(1), dim3(EYE_TPB), 0, c.stream, (int)n, (const T*)G, out)
  DISPATCH3(dt, EK, 0);
#undef EK
  HIP_CHECK(hipGetLastError());
}

void eig_to_svd(Ctx& c, DType dt, int64_t n, const double* lam, const void* W, double* s, void* Wr, double* sinv) {
  if (n <= 0) return;
#define VK(T, ...) hipLaunchKernelGGL(eig_to_svd_kernel<T>, dim3(blocks_for(n * n)), dim3(TPB), 0, c.stream, n, lam, (const T*)W, s, (T*)Wr, sinv)
  DISPATCH3(dt, VK, 0);
#undef VK
  HIP_CHECK(hipGetLastError());
}

void ns_coeff(Ctx& c, DType dt, int64_t n, void* G) {
  if (n <= 0) return;
#define NK(T, ...) hipLaunchKernelGGL(ns_coeff_kernel<T>, dim3(blocks_for(n * n)), dim3(TPB), 0, c.stream, n, (T*)G)
  DISPATCH3(dt, NK, 0);
#undef NK
  HIP_CHECK(hipGetLastError());
}

}  // namespace k
}  // namespace ipeps

// ---------------------------------------------------------------------------
// Small SPD Cholesky + triangular inverse in one workgroup (blockchol_impl.h)
// ---------------------------------------------------------------------------
namespace ipeps {
namespace k {

static_assert(BCH_MAXN == SCHOL_MAXN_HOST, "small_chol_inv size limit");

__device__ inline double readlane_d(double v, int l) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), l);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// 1 / sqrt(p), p > 0: v_rsq_f64 and two Newton steps
__device__ inline double rsqrt_nr(double p) {
  double r = __builtin_amdgcn_rsq(p);
  const double h = 0.5 * p;
  r = r * (1.5 - h * r * r);
  r = r * (1.5 - h * r * r);
  return r;
}

// (a) diagonal tile K by wave 0: lane i < 16 holds row i; the tile is replaced
// by the inverse of its Cholesky factor (zeros above the diagonal), its rows
// formed from the right with 1 / L[m][m] from the pivots' reciprocal square
// roots (no divisions) and L's rows read back from the tile as broadcasts.
// Returns the 1-based failing pivot within the tile, 0 on success (uniform).
__device__ inline int bch_diag(double* tile, int lane) {
  const int i = lane & (BCH_T - 1);
  double c[BCH_T], dinv[BCH_T];
#pragma unroll
  for (int j = 0; j < BCH_T; ++j) c[j] = tile[j * BCH_LD + i];
  int fail = 0;
#pragma unroll
  for (int k = 0; k < BCH_T; ++k) {
    const double pk = readlane_d(c[k], k);
    if (!(pk > 0.0) || !(pk < 1.7e308)) {
      if (!fail) fail = k + 1;
    }
    const double ap = fabs(pk) > 0.0 ? fabs(pk) : 1.0;
    const double rk = rsqrt_nr(ap), sk = ap * rk;
    dinv[k] = rk;
    c[k] = i == k ? sk : c[k] * rk;  // L[i][k] (rows below k; rows above hold garbage)
#pragma unroll
    for (int j = k + 1; j < BCH_T; ++j) c[j] -= c[k] * readlane_d(c[k], j);
  }
  // L to the tile (lane i: row i), then row i of L^-1 by columns from the
  // right: x[m] = (i == m ? 1 : -acc[m]) / L[m][m] (0 above the diagonal),
  // then acc[j] += x[m] L[m][j] for every j < m (independent updates)
  if (lane < BCH_T) {
#pragma unroll
    for (int j = 0; j < BCH_T; ++j) tile[j * BCH_LD + i] = c[j];
  }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  double acc[BCH_T];
#pragma unroll
  for (int j = 0; j < BCH_T; ++j) acc[j] = 0.0;
#pragma unroll
  for (int m = BCH_T - 1; m >= 0; --m) {
    const double xm = i == m ? dinv[m] : (i > m ? -acc[m] * dinv[m] : 0.0);
#pragma unroll
    for (int j = 0; j < m; ++j) acc[j] += xm * tile[j * BCH_LD + m];  // L[m][j], a broadcast read
    // column m of L is read no more (only columns j < m are, below): X[i][m] goes there
    __builtin_amdgcn_wave_barrier();
    if (lane < BCH_T) tile[m * BCH_LD + i] = xm;
    __builtin_amdgcn_sched_barrier(0);
  }
  return fail;
}

typedef double v4d __attribute__((ext_vector_type(4)));

// acc + A B over one 16 x 16 x 16 tile product (four 16x16x4 FP64 MFMAs);
// a(r, k) / b(k, c) give the operands: lane l supplies A[l % 16][4 s + l / 16]
// and B[4 s + l / 16][l % 16] at k step s.  Output lane l, element e:
// C[l / 16 + 4 e][l % 16].
template <class FA, class FB>
__device__ inline v4d bch_mm(FA a, FB b, v4d acc, int lane) {
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int s = 0; s < 4; ++s) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a(r, 4 * s + kq), b(4 * s + kq, r), acc, 0, 0, 0);
  return acc;
}

__device__ inline void bch_store(double* tile, v4d acc, int lane) {
#pragma unroll
  for (int e = 0; e < 4; ++e) tile[(lane & 15) * BCH_LD + (lane >> 4) + 4 * e] = acc[e];
}

__global__ __launch_bounds__(BCH_THREADS) void small_chol_inv_kernel(int n, const double* __restrict__ G, int64_t ldg,
                                                                     double* __restrict__ Linv, int64_t ldl,
                                                                     int* __restrict__ info) {
  constexpr int NW = BCH_THREADS / 64;
  __shared__ double T[BCH_NTILES * BCH_TE];
  __shared__ int fail;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int NB = (n + BCH_T - 1) / BCH_T;
  auto tile = [&](int I, int J) { return T + bch_tix(I, J) * BCH_TE; };
  if (tid == 0) fail = 0;
  // lower tiles of G, column j by wave, rows by lane (padding: identity)
  for (int j = wave; j < NB * BCH_T; j += NW) {
    const int J = j >> 4, cc = j & (BCH_T - 1);
    for (int i = (J << 4) + lane; i < NB * BCH_T; i += 64) {
      const double v = (i < n && j < n) ? G[i + (int64_t)j * ldg] : (i == j ? 1.0 : 0.0);
      T[bch_at(i >> 4, J, i & (BCH_T - 1), cc)] = v;
    }
  }
  __syncthreads();
  for (int K = 0; K < NB; ++K) {
    if (wave == 0) {
      const int f = bch_diag(tile(K, K), lane);
      if (lane == 0 && f) fail = K * BCH_T + f;
    }
    __syncthreads();
    if (fail) break;
    // (b) panel: L_IK = A_IK Dinv_K^T, a wave per tile (it reads only its own tile and Dinv_K)
    const double* D = tile(K, K);
    for (int I = K + 1 + wave; I < NB; I += NW) {
      double* t = tile(I, K);
      v4d acc = {0.0, 0.0, 0.0, 0.0};
      acc = bch_mm([&](int r, int k) { return t[k * BCH_LD + r]; }, [&](int k, int c) { return D[k * BCH_LD + c]; },
                   acc, lane);
      bch_store(t, acc, lane);
    }
    __syncthreads();
    // (c) trailing tiles (I, J), K < J <= I: A_IJ - L_IK L_JK^T, a wave per tile
    const int m = NB - K - 1, ntr = m * (m + 1) / 2;
    for (int q = wave; q < ntr; q += NW) {
      int a = 0;
      while ((a + 1) * (a + 2) / 2 <= q) ++a;
      const int I = K + 1 + a, J = K + 1 + (q - a * (a + 1) / 2);
      double* t = tile(I, J);
      const double *LI = tile(I, K), *LJ = tile(J, K);
      v4d acc;
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[e] = t[(lane & 15) * BCH_LD + (lane >> 4) + 4 * e];
      acc = bch_mm([&](int r, int k) { return -LI[k * BCH_LD + r]; }, [&](int k, int c) { return LJ[k * BCH_LD + c]; },
                   acc, lane);
      bch_store(t, acc, lane);
    }
    __syncthreads();
  }
  if (fail) {
    if (tid == 0) *info = fail <= n ? fail : n;
    return;
  }
  // inverse, tile columns J = NB-2 .. 0; wave w computes X_IJ for the rows
  // I = J + 1 + w + NW u (u < IPW), held until column J of L is read
  constexpr int IPW = (BCH_NBMAX - 1 + NW - 1) / NW;
  for (int J = NB - 2; J >= 0; --J) {
    v4d x[IPW];
#pragma unroll
    for (int u = 0; u < IPW; ++u) {
      const int I = J + 1 + wave + NW * u;
      x[u] = (v4d){0.0, 0.0, 0.0, 0.0};
      if (I >= NB) continue;  // uniform
      v4d wt = {0.0, 0.0, 0.0, 0.0};  // W^T = sum_M L_MJ^T X_IM^T
      for (int M = J + 1; M <= I; ++M) {
        const double *LM = tile(M, J), *XI = tile(I, M);
        wt = bch_mm([&](int r, int k) { return LM[r * BCH_LD + k]; }, [&](int k, int c) { return XI[k * BCH_LD + c]; },
                    wt, lane);
      }
      // X_IJ = -W X_JJ: the A operand at k step s is -wt[s]
      const double* XJ = tile(J, J);
      const int kq = lane >> 4, c = lane & 15;
#pragma unroll
      for (int s = 0; s < 4; ++s)
        x[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(-wt[s], XJ[c * BCH_LD + 4 * s + kq], x[u], 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < IPW; ++u) {
      const int I = J + 1 + wave + NW * u;
      if (I < NB) bch_store(tile(I, J), x[u], lane);
    }
    __syncthreads();
  }
  for (int j = wave; j < n; j += NW) {
    const int J = j >> 4, cc = j & (BCH_T - 1);
    for (int i = lane; i < n; i += 64)
      Linv[i + (int64_t)j * ldl] = i >= j ? T[bch_at(i >> 4, J, i & (BCH_T - 1), cc)] : 0.0;
  }
  if (tid == 0) *info = 0;
}

void small_chol_inv(Ctx& c, int n, const double* G, int64_t ldg, double* Linv, int64_t ldl, int* info) {
  IPEPS_CHECK(n >= 1 && n <= BCH_MAXN, "small_chol_inv: n out of range");
  TimedLaunch tl(c, "small_chol_inv_kernel", (double)n * n * n, 0.0);
  hipLaunchKernelGGL(small_chol_inv_kernel, dim3(1), dim3(BCH_THREADS), 0, c.stream, n, G, ldg, Linv, ldl, info);
  HIP_CHECK(hipGetLastError());
}

}  // namespace k
}  // namespace ipeps

// ---------------------------------------------------------------------------
// Symmetric tridiagonal reduction (one workgroup) and formation of Q
// (one wave per column), tridiag_impl.h
// ---------------------------------------------------------------------------
namespace ipeps {
namespace k {

__device__ inline double wave_sum_d(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// x moved across lanes by one DPP control (both 32-bit halves); rows outside
// ROWS receive 0.
template <int CTRL, int ROWS>
__device__ inline double dpp_move_d(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Sum over the 64 lanes of a wave, valid in lane 63 only: quad permutes and
// row rotates sum each 16-lane row, row_bcast:15 adds row 0 into row 1 and row
// 2 into row 3, row_bcast:31 adds row 1 into rows 2 and 3.  VALU moves only
// (no LDS permutes); every lane of the wave must be active.
__device__ inline double wave_sum_dpp63(double x) {
  x += dpp_move_d<0xB1, 0xF>(x);   // quad_perm [1, 0, 3, 2]
  x += dpp_move_d<0x4E, 0xF>(x);   // quad_perm [2, 3, 0, 1]
  x += dpp_move_d<0x124, 0xF>(x);  // row_ror:4
  x += dpp_move_d<0x128, 0xF>(x);  // row_ror:8
  x += dpp_move_d<0x142, 0xA>(x);  // row_bcast:15
  x += dpp_move_d<0x143, 0xC>(x);  // row_bcast:31
  return x;
}

// the value of lane l (wave-uniform l)
__device__ inline double lane_value_d(double x, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Householder tridiagonalisation in one workgroup (n <= TRI_BIG_MAXN); the
// scheme is described in tridiag_impl.h.
constexpr int SQ_NWAVE = TRI_THREADS / 64;
constexpr int SQ_RW = (TRI_BIG_MAXN + 63) / 64;  // 64-row blocks of a column (global phase)
constexpr int SQ_LM = TRI_SQ_LDS;                // LDS phase: trailing square at most SQ_LM a side
constexpr int SQ_U = 4;                          // columns per wave in flight

struct SqShared {
  double v[2][TRI_BIG_MAXN], w[2][TRI_BIG_MAXN];
  double part_l[SQ_NWAVE][SQ_LM];
  double red[SQ_NWAVE];
  double sc[2];
  double junk[64];
};

// reflector of step j from column j (rows j .. n-1) of the square, stored at
// col[row] for the column's base pointer; step j-1's update applied on the fly
template <typename P>
__device__ __attribute__((always_inline)) void sq_reflector(P col, int n, int j, bool upd, const double* vp,
                                                            const double* wp, double* vj, SqShared& sh,
                                                            double* __restrict__ d, double* __restrict__ e,
                                                            double* __restrict__ tau, double* __restrict__ V) {
  // col[i] = A(j + i, j), i < n - j
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nj = n - j;
  if (wave == 0) {
    const double vp0 = upd ? vp[0] : 0.0, wp0 = upd ? wp[0] : 0.0;
    double xs[SQ_RW];
    double x2 = 0.0;
#pragma unroll
    for (int q = 0; q < SQ_RW; ++q) {
      const int i = 2 + lane + 64 * q;
      double x = 0.0;
      if (i < nj) {
        x = col[i];
        if (upd) x = x - (vp[i] * wp0 + wp[i] * vp0);
        x2 += x * x;
      }
      xs[q] = x;
    }
    x2 = lane_value_d(wave_sum_dpp63(x2), 63);
    double scale = 0.0;
    if (lane == 0) {
      double a0 = col[0], alpha = col[1];
      if (upd) {
        a0 = a0 - (vp[0] * wp0 + wp[0] * vp0);
        alpha = alpha - (vp[1] * wp0 + wp[1] * vp0);
      }
      double t = 0.0, beta = alpha;
      if (x2 > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + x2), alpha);
        t = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
      }
      sh.sc[0] = t;
      d[j] = a0;
      e[j] = beta;
      tau[j] = t;
      vj[0] = 1.0;
      V[(j + 1) + (int64_t)j * n] = 1.0;
    }
    scale = lane_value_d(scale, 0);
#pragma unroll
    for (int q = 0; q < SQ_RW; ++q) {
      const int i = 2 + lane + 64 * q;
      if (i < nj) {
        const double vr = xs[q] * scale;
        vj[i - 1] = vr;
        V[(j + i) + (int64_t)j * n] = vr;
      }
    }
  }
  __syncthreads();
}

// pass of step j over the trailing square (local rows / columns 0 .. m-1,
// m = n-1-j; element (r, c) at base[r + c * ld]); QB = 64-row blocks per
// column (uniform), RW its compile-time bound.  Rows r >= m of the last block
// are padding: with JUNK they go to junk[lane] instead (the square's columns
// have no padding rows), else the storage has 64 padding rows per column.
template <int RW, bool JUNK, typename P>
__device__ __attribute__((always_inline)) void sq_pass(P base, int64_t ld, int m, bool upd, const double* vp,
                                                       const double* wp, const double* vj, double* wj,
                                                       double* part, int ldp, P junk, SqShared& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int QB = (m + 63) >> 6;
  const double t = sh.sc[0];
  double pp[RW], vpr[RW], wpr[RW];
  bool pad[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    pp[q] = 0.0;
    const int r = lane + 64 * q;
    pad[q] = r >= m;
    vpr[q] = (upd && q < QB) ? vp[min(r + 1, m)] : 0.0;
    wpr[q] = (upd && q < QB) ? wp[min(r + 1, m)] : 0.0;
  }
  for (int c0 = wave; c0 < m; c0 += SQ_U * SQ_NWAVE) {
    double a[SQ_U][RW];
#pragma unroll
    for (int u = 0; u < SQ_U; ++u) {
      const int c = min(c0 + u * SQ_NWAVE, m - 1);  // past the end: a repeat of the last column, not used
      P col = base + (int64_t)c * ld + lane;
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        P p = col + 64 * q;
        if (JUNK && pad[q]) p = junk;
        a[u][q] = q < QB ? *p : 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < SQ_U; ++u) {
      const int c = c0 + u * SQ_NWAVE;
      if (c >= m) break;  // uniform
      P col = base + (int64_t)c * ld + lane;
      if (upd) {
        const double vc = vp[c + 1], wc = wp[c + 1];
#pragma unroll
        for (int q = 0; q < RW; ++q)
          if (q < QB) {
            a[u][q] = a[u][q] - (vpr[q] * wc + wpr[q] * vc);
            P p = col + 64 * q;
            if (JUNK && pad[q]) p = junk;
            *p = a[u][q];
          }
      }
      const double vc = vj[c];
#pragma unroll
      for (int q = 0; q < RW; ++q) pp[q] += a[u][q] * vc;
    }
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const int r = lane + 64 * q;
    if (q < QB && r < m) part[wave * ldp + r] = pp[q];
  }
  __syncthreads();
  double prod = 0.0;
  for (int r = tid; r < m; r += TRI_THREADS) {
    double pr = 0.0;
#pragma unroll
    for (int w = 0; w < SQ_NWAVE; ++w) pr += part[w * ldp + r];
    pr *= t;
    wj[r] = pr;
    prod += pr * vj[r];
  }
  prod = wave_sum_dpp63(prod);
  if (lane == 63) sh.red[wave] = prod;
  __syncthreads();
  double dot = 0.0;
#pragma unroll
  for (int w = 0; w < SQ_NWAVE; ++w) dot += sh.red[w];
  for (int r = tid; r < m; r += TRI_THREADS) wj[r] += -0.5 * t * dot * vj[r];
  __syncthreads();
}

// pass of step j over the lower triangle of the trailing square kept in
// global memory with 64 padding rows per column (element (r, c) at
// base[r + c * ld], local r, c < m): rows r >= c of column c, the 64-row
// blocks wholly above the diagonal skipped (uniform), the diagonal block's
// rows above it read and written back unchanged-in-effect (their values are
// never used: only the lower triangle is kept) and masked out of the sums.
// p = tau A v takes the lower part per lane row and each column's transposed
// part by a wave sum.
template <int RW, typename P>
__device__ __attribute__((always_inline)) void sq_pass_lower(P base, int64_t ld, int m, bool upd, const double* vp,
                                                             const double* wp, const double* vj, double* wj,
                                                             double* part, int ldp, double* tp, SqShared& sh) {
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const double t = sh.sc[0];
  double pp[RW], vpr[RW], wpr[RW], vjr[RW];
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    pp[q] = 0.0;
    const int r = lane + 64 * q;
    vpr[q] = upd ? vp[min(r + 1, m)] : 0.0;
    wpr[q] = upd ? wp[min(r + 1, m)] : 0.0;
    vjr[q] = r < m ? vj[r] : 0.0;
  }
  const int QB = (m + 63) >> 6;
  for (int c0 = wave; c0 < m; c0 += SQ_U * SQ_NWAVE) {
    double a[SQ_U][RW];
#pragma unroll
    for (int u = 0; u < SQ_U; ++u) {
      const int c = min(c0 + u * SQ_NWAVE, m - 1);
      P col = base + (int64_t)c * ld + lane;
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        a[u][q] = 0.0;
        if (q < QB && 64 * q + 63 >= c) a[u][q] = col[64 * q];  // uniform
      }
    }
#pragma unroll
    for (int u = 0; u < SQ_U; ++u) {
      const int c = c0 + u * SQ_NWAVE;
      if (c >= m) break;  // uniform
      P col = base + (int64_t)c * ld + lane;
      const double vc = vj[c];
      double s = 0.0;
      const double vpc = upd ? vp[c + 1] : 0.0, wpc = upd ? wp[c + 1] : 0.0;
#pragma unroll
      for (int q = 0; q < RW; ++q) {
        if (q >= QB || 64 * q + 63 < c) continue;  // uniform
        const int r = lane + 64 * q;
        if (upd) {
          a[u][q] = a[u][q] - (vpr[q] * wpc + wpr[q] * vpc);
          col[64 * q] = a[u][q];
        }
        const double x = (r >= c) ? a[u][q] : 0.0;
        pp[q] += x * vc;
        s += (r > c ? x : 0.0) * vjr[q];
      }
      s = wave_sum_dpp63(s);
      if (lane == 63) tp[c] = s;
    }
  }
#pragma unroll
  for (int q = 0; q < RW; ++q) {
    const int r = lane + 64 * q;
    if (q < QB && r < m) part[wave * ldp + r] = pp[q];
  }
  __syncthreads();
  double prod = 0.0;
  for (int r = tid; r < m; r += TRI_THREADS) {
    double pr = tp[r];
#pragma unroll
    for (int w = 0; w < SQ_NWAVE; ++w) pr += part[w * ldp + r];
    pr *= t;
    wj[r] = pr;
    prod += pr * vj[r];
  }
  prod = wave_sum_dpp63(prod);
  if (lane == 63) sh.red[wave] = prod;
  __syncthreads();
  double dot = 0.0;
#pragma unroll
  for (int w = 0; w < SQ_NWAVE; ++w) dot += sh.red[w];
  for (int r = tid; r < m; r += TRI_THREADS) wj[r] += -0.5 * t * dot * vj[r];
  __syncthreads();
}

// G (lower triangle read) -> d, e, tau, V (reflectors, column-major n x n).
// W: for n > TRI_SQ_LDS, (n + 64) x n doubles, the square with ld = n + 64
// (rows n .. n+63 padding); unused otherwise.
__global__ __launch_bounds__(TRI_THREADS) void sytrd_kernel(int n, const double* __restrict__ G, int64_t ldg,
                                                          double* __restrict__ d, double* __restrict__ e,
                                                          double* __restrict__ tau, double* __restrict__ V,
                                                          double* __restrict__ W) {
  __shared__ double SQ[SQ_LM * SQ_LM];
  __shared__ SqShared sh;
  const int tid = threadIdx.x;
  const int64_t ldw = n + 64;
  if (n > SQ_LM)
    for (int64_t idx = tid; idx < (int64_t)n * n; idx += TRI_THREADS) {
      const int r = (int)(idx % n), c = (int)(idx / n);
      W[r + c * ldw] = r >= c ? G[r + c * ldg] : G[c + r * ldg];
    }
  for (int idx = tid; idx < n * n; idx += TRI_THREADS) V[idx] = 0.0;
  __syncthreads();
  int j = 0;
  for (; j + 2 < n && n - j > SQ_LM; ++j) {
    const int b = j & 1;
    const bool upd = j > 0;
    sq_reflector(W + j + (int64_t)j * ldw, n, j, upd, sh.v[b ^ 1], sh.w[b ^ 1], sh.v[b], sh, d, e, tau, V);
    sq_pass_lower<SQ_RW>(W + (j + 1) + (int64_t)(j + 1) * ldw, ldw, n - 1 - j, upd, sh.v[b ^ 1], sh.w[b ^ 1],
                         sh.v[b], sh.w[b], SQ, TRI_BIG_MAXN, SQ + SQ_NWAVE * TRI_BIG_MAXN, sh);
  }
  // the trailing square A(j.., j..) into LDS, local (r, c) at SQ[r + c * SQ_LM]
  const int nl = n - j;
  for (int idx = tid; idx < nl * nl; idx += TRI_THREADS) {
    const int r = idx % nl, c = idx / nl;
    if (n <= SQ_LM)  // j = 0: straight from G
      SQ[r + c * SQ_LM] = r >= c ? G[r + c * ldg] : G[c + r * ldg];
    else
      SQ[r + c * SQ_LM] = r >= c ? W[(j + r) + (int64_t)(j + c) * ldw] : W[(j + c) + (int64_t)(j + r) * ldw];
  }
  __syncthreads();
  for (; j + 2 < n; ++j) {
    const int b = j & 1;
    const bool upd = j > 0;
    const int o = j - (n - nl);  // local index of row / column j
    sq_reflector(SQ + o + o * SQ_LM, n, j, upd, sh.v[b ^ 1], sh.w[b ^ 1], sh.v[b], sh, d, e, tau, V);
    sq_pass<2, true>(SQ + (o + 1) + (o + 1) * SQ_LM, SQ_LM, n - 1 - j, upd, sh.v[b ^ 1], sh.w[b ^ 1], sh.v[b],
                     sh.w[b], &sh.part_l[0][0], SQ_LM, sh.junk + (threadIdx.x & 63), sh);
  }
  if (tid == 0) {
    if (n >= 2) {
      const int o = (n - 2) - (n - nl);
      double a00 = SQ[o + o * SQ_LM], a10 = SQ[(o + 1) + o * SQ_LM], a11 = SQ[(o + 1) + (o + 1) * SQ_LM];
      if (j > 0) {
        const double* vp = sh.v[(j - 1) & 1];
        const double* wp = sh.w[(j - 1) & 1];
        a00 = a00 - (vp[0] * wp[0] + wp[0] * vp[0]);
        a10 = a10 - (vp[1] * wp[0] + wp[1] * vp[0]);
        a11 = a11 - (vp[1] * wp[1] + wp[1] * vp[1]);
      }
      d[n - 2] = a00;
      e[n - 2] = a10;
      tau[n - 2] = 0.0;
      d[n - 1] = a11;
    } else {
      d[0] = SQ[0];
    }
  }
}

// column c of Q = H(0) ... H(c-1) e_c (one wave; rows lane + 64 t, t < R)
template <int R>
__global__ __launch_bounds__(256) void orgtr_kernel(int n, const double* __restrict__ V, const double* __restrict__ tau,
                                                    double* __restrict__ Q, int64_t ldq) {
  const int c = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (c >= n) return;
  double q[R];
#pragma unroll
  for (int t = 0; t < R; ++t) q[t] = (lane + 64 * t == c) ? 1.0 : 0.0;
  for (int i = min(c - 1, n - 3); i >= 0; --i) {
    const double* vi = V + (int64_t)i * n;
    double vr[R];
    double dot = 0.0;
#pragma unroll
    for (int t = 0; t < R; ++t) {
      const int r = lane + 64 * t;
      vr[t] = r < n ? vi[r] : 0.0;
      dot += vr[t] * q[t];
    }
    dot = wave_sum_d(dot) * tau[i];
#pragma unroll
    for (int t = 0; t < R; ++t) q[t] -= vr[t] * dot;
  }
#pragma unroll
  for (int t = 0; t < R; ++t) {
    const int r = lane + 64 * t;
    if (r < n) Q[r + (int64_t)c * ldq] = q[t];
  }
}

// ---------------------------------------------------------------------------
// Cyclic two-sided Jacobi for a symmetric n x n FP64 matrix, n <= 112, in one
// workgroup with everything in LDS: the lower triangle of A packed by columns
// (tri_off) and V = the product of the rotations.  A sweep is m - 1 rounds
// (m = n rounded up to even; index n is a dummy when n is odd) of the
// round-robin ordering: in round r the players at positions i and m - 1 - i
// form pair i, position 0 fixed and the others rotating, so the m / 2 pairs of
// a round are disjoint and their rotations J(p, q) commute.  A round computes
// each pair's rotation (Golub & Van Loan, sym.schur2: tau = (a_qq - a_pp) /
// 2 a_pq, t = sign(tau) / (|tau| + sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2),
// s = t c), skipped when |a_pq| <= tol sqrt(|a_pp a_qq|), then updates every
// 2 x 2 block of A below the diagonal (rows of pair a, columns of pair b,
// a > b; the block above is its transpose) as B <- J_a^T B J_b, the diagonal
// blocks by the closed forms a_pp - t a_pq, a_qq + t a_pq and an exact zero,
// and V <- V J.  The sweeps stop after one with no rotation.
// ---------------------------------------------------------------------------
constexpr int JAC_MAXN = 112;
constexpr int JAC_THREADS = 1024;
constexpr int JAC_MAXBLK = (JAC_MAXN / 2) * (JAC_MAXN / 2 - 1) / 2;
constexpr int JAC_MAXSWEEP = 40;
constexpr double JAC_TOL = 1e-14;

__device__ inline int jac_player(int pos, int r, int m) { return pos == 0 ? 0 : 1 + (pos - 1 + r) % (m - 1); }
// packed lower-triangle index of A(i, j), either order
__device__ inline int jac_at(int n, int i, int j) { return i >= j ? tri_off(n, j) + i - j : tri_off(n, i) + j - i; }

__global__ __launch_bounds__(JAC_THREADS) void jacobi_eig_kernel(int n, const double* __restrict__ G, int64_t ldg,
                                                                 double* __restrict__ lam, double* __restrict__ Q,
                                                                 int64_t ldq, int* __restrict__ info) {
  __shared__ double A[JAC_MAXN * (JAC_MAXN + 1) / 2];
  __shared__ double V[JAC_MAXN * JAC_MAXN];  // V(i, j) at V[i + j n]
  __shared__ double cs[JAC_MAXN / 2], sn[JAC_MAXN / 2], tt[JAC_MAXN / 2];
  __shared__ int pp[JAC_MAXN / 2], pq[JAC_MAXN / 2];
  __shared__ int rot_any;
  __shared__ unsigned char blk_a[JAC_MAXBLK], blk_b[JAC_MAXBLK];  // block it -> pairs (a > b)
  const int tid = threadIdx.x;
  const int m = n + (n & 1), np = m / 2;
  for (int j = 0; j < n; ++j)
    for (int i = j + tid; i < n; i += JAC_THREADS) A[tri_off(n, j) + i - j] = G[i + (int64_t)j * ldg];
  for (int idx = tid; idx < n * n; idx += JAC_THREADS) V[idx] = (idx % n) == (idx / n) ? 1.0 : 0.0;
  const int nblk = np * (np - 1) / 2;  // blocks a > b
  for (int a = 1 + tid; a < np; a += JAC_THREADS)
    for (int b = 0; b < a; ++b) {
      blk_a[a * (a - 1) / 2 + b] = (unsigned char)a;
      blk_b[a * (a - 1) / 2 + b] = (unsigned char)b;
    }
  __syncthreads();
  bool converged = n <= 1;
  for (int sweep = 0; sweep < JAC_MAXSWEEP && !converged; ++sweep) {
    if (tid == 0) rot_any = 0;
    __syncthreads();
    for (int r = 0; r < m - 1; ++r) {
      if (tid < np) {
        const int a = jac_player(tid, r, m), b = jac_player(m - 1 - tid, r, m);
        const int p = min(a, b), q = max(a, b);
        double c = 1.0, s = 0.0, t = 0.0;
        if (q < n) {
          const double apq = A[jac_at(n, q, p)], app = A[tri_off(n, p)], aqq = A[tri_off(n, q)];
          if (apq != 0.0 && fabs(apq) > JAC_TOL * sqrt(fabs(app * aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
            rot_any = 1;
          }
        }
        cs[tid] = c;
        sn[tid] = s;
        tt[tid] = t;
        pp[tid] = p;
        pq[tid] = q;
      }
      __syncthreads();
      // blocks below the diagonal: item -> (ia > ib)
      for (int it = tid; it < nblk + np; it += JAC_THREADS) {
        if (it >= nblk) {  // diagonal block of pair it - nblk
          const int ia = it - nblk;
          const int p0 = pp[ia], q0 = pq[ia];
          if (q0 < n && tt[ia] != 0.0) {
            const int ipq = jac_at(n, q0, p0);
            const double apq = A[ipq];
            A[tri_off(n, p0)] -= tt[ia] * apq;
            A[tri_off(n, q0)] += tt[ia] * apq;
            A[ipq] = 0.0;
          }
          continue;
        }
        const int ia = blk_a[it], ib = blk_b[it];
        const double ca = cs[ia], sa = sn[ia], cb = cs[ib], sb = sn[ib];
        if (sa == 0.0 && sb == 0.0) continue;
        const int p0 = pp[ia], q0 = pq[ia], p1 = pp[ib], q1 = pq[ib];
        const bool rq = q0 < n, cq = q1 < n;  // rows / columns with a real q
        const int i00 = jac_at(n, p0, p1);
        const int i01 = cq ? jac_at(n, p0, q1) : 0;
        const int i10 = rq ? jac_at(n, q0, p1) : 0;
        const int i11 = rq && cq ? jac_at(n, q0, q1) : 0;
        const double b00 = A[i00];
        const double b01 = cq ? A[i01] : 0.0;
        const double b10 = rq ? A[i10] : 0.0;
        const double b11 = rq && cq ? A[i11] : 0.0;
        // X = B J_b, then J_a^T X
        const double x00 = cb * b00 - sb * b01, x01 = sb * b00 + cb * b01;
        const double x10 = cb * b10 - sb * b11, x11 = sb * b10 + cb * b11;
        A[i00] = ca * x00 - sa * x10;
        if (cq) A[i01] = ca * x01 - sa * x11;
        if (rq) A[i10] = sa * x00 + ca * x10;
        if (rq && cq) A[i11] = sa * x01 + ca * x11;
      }
      // V <- V J: row i, pair b
      for (int it = tid; it < n * np; it += JAC_THREADS) {
        const int i = it % n, ib = it / n;
        const int p1 = pp[ib], q1 = pq[ib];
        if (q1 >= n || tt[ib] == 0.0) continue;
        const double cb = cs[ib], sb = sn[ib];
        const double vp = V[i + p1 * n], vq = V[i + q1 * n];
        V[i + p1 * n] = cb * vp - sb * vq;
        V[i + q1 * n] = sb * vp + cb * vq;
      }
      __syncthreads();
    }
    converged = rot_any == 0;
    __syncthreads();
  }
  // ascending order: rank of d_i among the diagonal (ties by index)
  for (int i = tid; i < n; i += JAC_THREADS) {
    const double di = A[tri_off(n, i)];
    int rk = 0;
    for (int j = 0; j < n; ++j) {
      const double dj = A[tri_off(n, j)];
      rk += (dj < di || (dj == di && j < i)) ? 1 : 0;
    }
    lam[rk] = di;
    for (int rr = 0; rr < n; ++rr) Q[rr + (int64_t)rk * ldq] = V[rr + i * n];
  }
  if (tid == 0) *info = converged ? 0 : 1;
}

void sytrd_q(Ctx& c, int n, const double* G, int64_t ldg, double* d, double* e, double* Q, int64_t ldq) {
  static_assert(TRI_BIG_MAXN == TRI_MAXN_HOST, "sytrd_q size limit");
  IPEPS_CHECK(n >= 1 && n <= TRI_BIG_MAXN, "sytrd_q: n out of range");
  const bool big = n > TRI_MAXN;
  const size_t nws = (size_t)n * n + n + (n > TRI_SQ_LDS ? (size_t)(n + 64) * n : 0);
  std::shared_ptr<Buf> ws(new Buf{c.pool->alloc(nws * 8), nws * 8, c.pool.get()});
  double* V = (double*)ws->p;
  double* tau = V + (size_t)n * n;
  double* W = n > TRI_SQ_LDS ? tau + n : nullptr;
  {
    TimedLaunch tl(c, "sytrd_kernel", 4.0 / 3.0 * n * n * n, 0.0);
    hipLaunchKernelGGL(sytrd_kernel, dim3(1), dim3(TRI_THREADS), 0, c.stream, n, G, ldg, d, e, tau, V, W);
  }
  HIP_CHECK(hipGetLastError());
  {
    TimedLaunch tl(c, "orgtr_kernel", 4.0 / 3.0 * n * n * n, 0.0);
    if (big)
      hipLaunchKernelGGL(orgtr_kernel<(TRI_BIG_MAXN + 63) / 64>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.stream,
                         n, V, tau, Q, ldq);
    else
      hipLaunchKernelGGL(orgtr_kernel<(TRI_MAXN + 63) / 64>, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, c.stream, n,
                         V, tau, Q, ldq);
  }
  HIP_CHECK(hipGetLastError());
}

// ---- tridiagonal eigenpairs: bisection + inverse iteration (trieig_impl.h) ----
// One wave per eigenvalue: each multisection step counts at 64 TRIEIG_PTS
// interior points of the current interval (TRIEIG_PTS per lane, their
// recurrences interleaved) and keeps the two points around the first one
// whose count exceeds j.  Every wave computes the Gershgorin bounds
// itself from the LDS copies of d and e.  Stops at the rounding of the end
// points (relative accuracy, tri_converged).
constexpr int TRIB_WAVES = 4;
__global__ __launch_bounds__(64 * TRIB_WAVES) void tri_bisect_kernel(int n, const double* __restrict__ d,
                                                                     const double* __restrict__ e,
                                                                     double* __restrict__ lam) {
  __shared__ double sd[TRIEIG_MAXN], se[TRIEIG_MAXN], se2[TRIEIG_MAXN];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sd[i] = d[i];
    const double v = i < n - 1 ? e[i] : 0.0;
    se[i] = v;
    se2[i] = v * v;
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * TRIB_WAVES + (int)(threadIdx.x >> 6);
  if (j >= n) return;  // whole waves
  double lo, hi, tnorm, pivmin;
  tri_bounds(n, sd, se, lo, hi, tnorm, pivmin);
  const double atol = tri_atol(pivmin);
  constexpr int NP = 64 * TRIEIG_PTS;
  for (int it = 0; it < TRIEIG_MAXIT && !tri_converged(lo, hi, atol); ++it) {
    double s[TRIEIG_PTS];
    int cnt[TRIEIG_PTS];
#pragma unroll
    for (int k = 0; k < TRIEIG_PTS; ++k) s[k] = tri_point(lo, hi, k * 64 + lane, NP);
    tri_sturm<TRIEIG_PTS>(n, sd, se2, s, pivmin, cnt);
    int f = NP;  // first point (k * 64 + lane order) whose count exceeds j; wave-uniform
#pragma unroll
    for (int k = TRIEIG_PTS - 1; k >= 0; --k) {
      const unsigned long long above = __ballot(cnt[k] > j);
      if (above) f = k * 64 + __builtin_ctzll(above);
    }
    const double nlo = f > 0 ? tri_point(lo, hi, f - 1, NP) : lo;
    const double nhi = f < NP ? tri_point(lo, hi, f, NP) : hi;
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) lam[j] = 0.5 * (lo + hi);
}

// One thread per eigenvalue: TRIEIG_SOLVES inverse-iteration solves from a
// pseudo-random start, each normalised (the last with the sign that makes the
// vector's product with the start positive), the factor rows in w (4 n x n, row i
// of eigenvalue j at w[(4 i + f) n + j]: a wave's accesses are contiguous);
// the unit eigenvector j goes to row j of Zt (Zt[j + i n] = z_j[i]).  Small
// blocks spread the latency-bound sweeps over more CUs.
constexpr int TRIV_TPB = 32;
__global__ __launch_bounds__(TRIV_TPB) void tri_invit_kernel(int n, const double* __restrict__ d,
                                                             const double* __restrict__ e,
                                                             const double* __restrict__ lam, double* Zt, double* w) {
  __shared__ double sd[TRIEIG_MAXN], se[TRIEIG_MAXN];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    sd[i] = d[i];
    se[i] = i < n - 1 ? e[i] : 0.0;
  }
  __syncthreads();
  const int j = blockIdx.x * TRIV_TPB + (int)threadIdx.x;
  if (j >= n) return;
  double gl, gu, tnorm, pivmin;
  tri_bounds(n, sd, se, gl, gu, tnorm, pivmin);
  const double pert = fmax(2.220446049250313e-16 * tnorm, 1e-300);
  const double l = lam[j];
  double* x = Zt + j;
  double* wj = w + j;
  double ss = tri_invit_solve(n, j, sd, se, l, pert, 1.0, nullptr, n, x, n, wj, n);
  for (int it = 1; it < TRIEIG_SOLVES; ++it)
    ss = tri_invit_solve(n, j, sd, se, l, pert, 1.0 / sqrt(ss), x, n, x, n, wj, n);
  const double sc = tri_sign_scale(n, j, x, n, ss);
  for (int i0 = 0; i0 < n; i0 += TRIEIG_BLK) {
    double v[TRIEIG_BLK];
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t) v[t] = i0 + t < n ? x[(int64_t)(i0 + t) * n] : 0.0;
#pragma unroll
    for (int t = 0; t < TRIEIG_BLK; ++t)
      if (i0 + t < n) x[(int64_t)(i0 + t) * n] = v[t] * sc;
  }
}

void tri_eig(Ctx& c, int n, const double* d, const double* e, double* lam, double* Zt) {
  static_assert(TRIEIG_MAXN == TRIEIG_MAXN_HOST, "tri_eig size limit");
  IPEPS_CHECK(n >= 1 && n <= TRIEIG_MAXN, "tri_eig: n out of range");
  const size_t nws = (size_t)4 * n * n;
  std::shared_ptr<Buf> ws(new Buf{c.pool->alloc(nws * 8), nws * 8, c.pool.get()});
  {
    TimedLaunch tl(c, "tri_bisect_kernel", 0.0, 0.0);
    hipLaunchKernelGGL(tri_bisect_kernel, dim3((unsigned)((n + TRIB_WAVES - 1) / TRIB_WAVES)), dim3(64 * TRIB_WAVES),
                       0, c.stream, n, d, e, lam);
  }
  HIP_CHECK(hipGetLastError());
  {
    TimedLaunch tl(c, "tri_invit_kernel", 0.0, 0.0);
    hipLaunchKernelGGL(tri_invit_kernel, dim3((unsigned)((n + TRIV_TPB - 1) / TRIV_TPB)), dim3(TRIV_TPB), 0, c.stream,
                       n, d, e, lam, Zt, (double*)ws->p);
  }
  HIP_CHECK(hipGetLastError());
}

void jacobi_eig(Ctx& c, int n, const double* G, int64_t ldg, double* lam, double* Q, int64_t ldq, int* info) {
  static_assert(JAC_MAXN == JAC_MAXN_HOST, "jacobi_eig size limit");
  IPEPS_CHECK(n >= 1 && n <= JAC_MAXN, "jacobi_eig: n out of range");
  TimedLaunch tl(c, "jacobi_eig_kernel", 0.0, 0.0);
  hipLaunchKernelGGL(jacobi_eig_kernel, dim3(1), dim3(JAC_THREADS), 0, c.stream, n, G, ldg, lam, Q, ldq, info);
  HIP_CHECK(hipGetLastError());
}

}  // namespace k
}  // namespace ipeps
