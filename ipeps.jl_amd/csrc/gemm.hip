// MFMA GEMM kernels for gfx950 (CDNA4), column-major, batched.
//
//  * FP64  : v_mfma_f64_16x16x4_f64, 64x64 block tile, K tile 16.
//  * C128  : the same tile with separate re/im LDS planes; each K step issues
//            four FP64 MFMAs (re*re, -im*im, re*im, im*re).
//  * F32   : FP32 storage, operands split into NSPLIT bf16 terms when staged
//            into LDS (hi = bf16(x), mid = bf16(x - hi), ...); products with
//            i + j < NSPLIT go through v_mfma_f32_16x16x32_bf16 with FP32
//            accumulation.
//
// Block: 256 threads = 4 waves of 64 in a 2 x 2 grid, each wave owns a
// 32 x 32 sub-tile = 2 x 2 MFMA 16 x 16 output tiles.
//
// MFMA register layouts used (16x16 family):
//   A operand: lane l holds A[l % 16][kgroup = l / 16] (f64: 1 value,
//              bf16 16x16x32: 8 consecutive k = 8*(l/16) + e)
//   B operand: lane l holds B[kgroup = l / 16][l % 16]
//   C/D      : f64 16x16x4: lane l, element r holds C[l/16 + 4*r][l % 16]
//              bf16 16x16x32 (f32 out): C[4*(l/16) + r][l % 16]
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ipeps {
namespace k {

typedef double v4f64 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf16 __attribute__((ext_vector_type(8)));
typedef unsigned short v8u16 __attribute__((ext_vector_type(8)));

struct GemmP {
  int64_t M, N, K;
  double alpha, beta;
  const void* A;
  int64_t lda, sA;
  const void* B;
  int64_t ldb, sB;
  void* C;
  int64_t ldc, sC;
  int conjA, conjB;
  int64_t kchunk;  // > 0: split-K, blockIdx.z indexes the K chunk (batch must be 1)
};

constexpr int BM = 64, BN = 64;

// ---------------------------------------------------------------------------
// FP64 real
// ---------------------------------------------------------------------------
template <bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_f64_kernel(GemmP p) {
  constexpr int BK = 16;
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const double* A = (const double*)p.A + (int64_t)blockIdx.z * p.sA;
  const double* B = (const double*)p.B + (int64_t)blockIdx.z * p.sB;
  double* C = (double*)p.C + (int64_t)blockIdx.z * p.sC;

  v4f64 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4f64){0.0, 0.0, 0.0, 0.0};

  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - (int64_t)blockIdx.z * p.kchunk) : p.K;
  for (int64_t k0 = 0; k0 < Keff; k0 += BK) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int m, kk;
      if (!AT) { m = tid & 63; kk = (tid >> 6) + 4 * r; }
      else { kk = tid & 15; m = (tid >> 4) + 16 * r; }
      int64_t gm = m0 + m, gk = k0 + kk;
      double v = 0.0;
      if (gm < p.M && gk < Keff) v = AT ? A[gk + gm * p.lda] : A[gm + gk * p.lda];
      As[kk][m] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int n, kk;
      if (!BT) { kk = tid & 15; n = (tid >> 4) + 16 * r; }
      else { n = tid & 63; kk = (tid >> 6) + 4 * r; }
      int64_t gn = n0 + n, gk = k0 + kk;
      double v = 0.0;
      if (gn < p.N && gk < Keff) v = BT ? B[gn + gk * p.ldb] : B[gk + gn * p.ldb];
      Bs[kk][n] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) {
          double v = p.alpha * acc[i][j][r];
          double* dst = C + row + col * p.ldc;
          if (p.beta != 0.0) v += p.beta * *dst;
          *dst = v;
        }
      }
}

// ---------------------------------------------------------------------------
// Complex FP64
// ---------------------------------------------------------------------------
template <bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_c128_kernel(GemmP p) {
  constexpr int BK = 8;
  __shared__ double Ar[BK][BM + 1], Ai[BK][BM + 1];
  __shared__ double Br[BK][BN + 1], Bi[BK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const double2* A = (const double2*)p.A + (int64_t)blockIdx.z * p.sA;
  const double2* B = (const double2*)p.B + (int64_t)blockIdx.z * p.sB;
  double2* C = (double2*)p.C + (int64_t)blockIdx.z * p.sC;
  const double sa = p.conjA ? -1.0 : 1.0, sb = p.conjB ? -1.0 : 1.0;

  v4f64 cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = (v4f64){0.0, 0.0, 0.0, 0.0};
      ci[i][j] = (v4f64){0.0, 0.0, 0.0, 0.0};
    }

  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - (int64_t)blockIdx.z * p.kchunk) : p.K;
  for (int64_t k0 = 0; k0 < Keff; k0 += BK) {
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int m, kk;
      if (!AT) { m = tid & 63; kk = (tid >> 6) + 4 * r; }
      else { kk = tid & 7; m = (tid >> 3) + 32 * r; }
      int64_t gm = m0 + m, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gm < p.M && gk < Keff) v = AT ? A[gk + gm * p.lda] : A[gm + gk * p.lda];
      Ar[kk][m] = v.x;
      Ai[kk][m] = sa * v.y;
    }
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      int n, kk;
      if (!BT) { kk = tid & 7; n = (tid >> 3) + 32 * r; }
      else { n = tid & 63; kk = (tid >> 6) + 4 * r; }
      int64_t gn = n0 + n, gk = k0 + kk;
      double2 v = make_double2(0.0, 0.0);
      if (gn < p.N && gk < Keff) v = BT ? B[gn + gk * p.ldb] : B[gk + gn * p.ldb];
      Br[kk][n] = v.x;
      Bi[kk][n] = sb * v.y;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        int mi = wm * 32 + i * 16 + (lane & 15);
        ar[i] = Ar[kk + (lane >> 4)][mi];
        ai[i] = Ai[kk + (lane >> 4)][mi];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        int nj = wn * 32 + j * 16 + (lane & 15);
        br[j] = Br[kk + (lane >> 4)][nj];
        bi[j] = Bi[kk + (lane >> 4)][nj];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cr[i][j], 0, 0, 0);
          cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], cr[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], ci[i][j], 0, 0, 0);
          ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], ci[i][j], 0, 0, 0);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * 32 + i * 16 + (lane >> 4) + 4 * r;
        int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) {
          double vr = p.alpha * cr[i][j][r], vi = p.alpha * ci[i][j][r];
          double2* dst = C + row + col * p.ldc;
          if (p.beta != 0.0) {
            double2 o = *dst;
            vr += p.beta * o.x;
            vi += p.beta * o.y;
          }
          *dst = make_double2(vr, vi);
        }
      }
}

// ---------------------------------------------------------------------------
// FP32 storage, split-bf16 MFMA products
// ---------------------------------------------------------------------------
__device__ inline unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ inline float bf16_to_f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

template <bool AT, bool BT, int NSPLIT>
__global__ __launch_bounds__(256) void gemm_f32_kernel(GemmP p) {
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;  // bf16 elements per LDS row (k contiguous), padded
  __shared__ unsigned short As[NSPLIT][BM][LDK];
  __shared__ unsigned short Bs[NSPLIT][BN][LDK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM, n0 = (int64_t)blockIdx.y * BN;
  const float* A = (const float*)p.A + (int64_t)blockIdx.z * p.sA;
  const float* B = (const float*)p.B + (int64_t)blockIdx.z * p.sB;
  float* C = (float*)p.C + (int64_t)blockIdx.z * p.sC;

  v4f32 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (v4f32){0.f, 0.f, 0.f, 0.f};

  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - (int64_t)blockIdx.z * p.kchunk) : p.K;
  for (int64_t k0 = 0; k0 < Keff; k0 += BK) {
    // 64 x 32 tile = 2048 values, 8 per thread.
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int m, kk;
      if (!AT) { m = tid & 63; kk = (tid >> 6) + 4 * r; }
      else { kk = tid & 31; m = (tid >> 5) + 8 * r; }
      int64_t gm = m0 + m, gk = k0 + kk;
      float v = 0.f;
      if (gm < p.M && gk < Keff) v = AT ? A[gk + gm * p.lda] : A[gm + gk * p.lda];
#pragma unroll
      for (int s = 0; s < NSPLIT; ++s) {
        unsigned short h = bf16_rne(v);
        As[s][m][kk] = h;
        v -= bf16_to_f(h);
      }
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      int n, kk;
      if (!BT) { kk = tid & 31; n = (tid >> 5) + 8 * r; }
      else { n = tid & 63; kk = (tid >> 6) + 4 * r; }
      int64_t gn = n0 + n, gk = k0 + kk;
      float v = 0.f;
      if (gn < p.N && gk < Keff) v = BT ? B[gn + gk * p.ldb] : B[gk + gn * p.ldb];
#pragma unroll
      for (int s = 0; s < NSPLIT; ++s) {
        unsigned short h = bf16_rne(v);
        Bs[s][n][kk] = h;
        v -= bf16_to_f(h);
      }
    }
    __syncthreads();
    v8bf16 a[NSPLIT][2], b[NSPLIT][2];
    const int kq = 8 * (lane >> 4);
#pragma unroll
    for (int s = 0; s < NSPLIT; ++s) {
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        v8u16 t = *reinterpret_cast<const v8u16*>(&As[s][wm * 32 + i * 16 + (lane & 15)][kq]);
        a[s][i] = __builtin_bit_cast(v8bf16, t);
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        v8u16 t = *reinterpret_cast<const v8u16*>(&Bs[s][wn * 32 + j * 16 + (lane & 15)][kq]);
        b[s][j] = __builtin_bit_cast(v8bf16, t);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        // smallest products first
#pragma unroll
        for (int t = NSPLIT - 1; t >= 0; --t)
#pragma unroll
          for (int sa = 0; sa <= t; ++sa)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[t - sa][j], acc[i][j], 0, 0, 0);
      }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
        int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) {
          float v = (float)p.alpha * acc[i][j][r];
          float* dst = C + row + col * p.ldc;
          if (p.beta != 0.0) v += (float)p.beta * *dst;
          *dst = v;
        }
      }
}

// ---------------------------------------------------------------------------
// split-K reduction: C = alpha * sum_s W_s + beta * C
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T zero_of() { return T(0); }
template <>
__device__ inline double2 zero_of<double2>() { return make_double2(0.0, 0.0); }
__device__ inline double2 add_(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double add_(double a, double b) { return a + b; }
__device__ inline float add_(float a, float b) { return a + b; }
__device__ inline double2 axpby_(double al, double2 s, double be, double2 c) {
  return make_double2(al * s.x + be * c.x, al * s.y + be * c.y);
}
__device__ inline double axpby_(double al, double s, double be, double c) { return al * s + be * c; }
__device__ inline float axpby_(double al, float s, double be, float c) { return (float)(al * s + be * c); }

template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce(int64_t M, int64_t N, int64_t S, const T* __restrict__ W,
                                                     double alpha, double beta, T* __restrict__ C, int64_t ldc) {
  const int64_t MN = M * N;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < MN; t += (int64_t)gridDim.x * 256) {
    T acc = zero_of<T>();
    for (int64_t s = 0; s < S; ++s) acc = add_(acc, W[s * MN + t]);
    int64_t m = t % M, n = t / M;
    T* dst = C + m + n * ldc;
    *dst = beta != 0.0 ? axpby_(alpha, acc, beta, *dst) : axpby_(alpha, acc, 0.0, zero_of<T>());
  }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
static void launch(Ctx& c, const GemmArgs& g, const GemmP& p, dim3 grid) {
  dim3 block(256);
  const bool at = g.transA, bt = g.transB;
  if (g.dt == DType::F64) {
    if (!at && !bt) hipLaunchKernelGGL((gemm_f64_kernel<false, false>), grid, block, 0, c.stream, p);
    else if (at && !bt) hipLaunchKernelGGL((gemm_f64_kernel<true, false>), grid, block, 0, c.stream, p);
    else if (!at && bt) hipLaunchKernelGGL((gemm_f64_kernel<false, true>), grid, block, 0, c.stream, p);
    else hipLaunchKernelGGL((gemm_f64_kernel<true, true>), grid, block, 0, c.stream, p);
  } else if (g.dt == DType::C128) {
    if (!at && !bt) hipLaunchKernelGGL((gemm_c128_kernel<false, false>), grid, block, 0, c.stream, p);
    else if (at && !bt) hipLaunchKernelGGL((gemm_c128_kernel<true, false>), grid, block, 0, c.stream, p);
    else if (!at && bt) hipLaunchKernelGGL((gemm_c128_kernel<false, true>), grid, block, 0, c.stream, p);
    else hipLaunchKernelGGL((gemm_c128_kernel<true, true>), grid, block, 0, c.stream, p);
  } else {
#define IPEPS_F32_LAUNCH(NS)                                                                         \
  do {                                                                                               \
    if (!at && !bt) hipLaunchKernelGGL((gemm_f32_kernel<false, false, NS>), grid, block, 0, c.stream, p); \
    else if (at && !bt) hipLaunchKernelGGL((gemm_f32_kernel<true, false, NS>), grid, block, 0, c.stream, p); \
    else if (!at && bt) hipLaunchKernelGGL((gemm_f32_kernel<false, true, NS>), grid, block, 0, c.stream, p); \
    else hipLaunchKernelGGL((gemm_f32_kernel<true, true, NS>), grid, block, 0, c.stream, p);          \
  } while (0)
    switch (g.mixed) {
      case Mixed::BF16: IPEPS_F32_LAUNCH(1); break;
      case Mixed::BF16X3: IPEPS_F32_LAUNCH(2); break;
      case Mixed::BF16X6: IPEPS_F32_LAUNCH(3); break;
    }
#undef IPEPS_F32_LAUNCH
  }
  HIP_CHECK(hipGetLastError());
}

void gemm(Ctx& c, const GemmArgs& g) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  if (g.K <= 0) {
    IPEPS_CHECK(g.beta == 0.0 || g.beta == 1.0, "gemm: K=0 with beta not in {0,1}");
    if (g.beta == 0.0) {
      size_t es = dtype_size(g.dt);
      for (int64_t b = 0; b < g.batch; ++b)
        for (int64_t j = 0; j < g.N; ++j)
          fill_zero(c, (char*)g.C + (b * g.sC + j * g.ldc) * es, g.M * es);
    }
    return;
  }
  IPEPS_CHECK(g.batch <= 65535, "gemm: batch too large");
  GemmP p;
  p.M = g.M; p.N = g.N; p.K = g.K; p.alpha = g.alpha; p.beta = g.beta;
  p.A = g.A; p.lda = g.lda; p.sA = g.sA;
  p.B = g.B; p.ldb = g.ldb; p.sB = g.sB;
  p.C = g.C; p.ldc = g.ldc; p.sC = g.sC;
  p.conjA = g.conjA; p.conjB = g.conjB;
  p.kchunk = 0;
  dim3 grid((unsigned)((g.M + BM - 1) / BM), (unsigned)((g.N + BN - 1) / BN), (unsigned)g.batch);
  IPEPS_CHECK(grid.y <= 65535, "gemm: N too large for the grid");
  const bool at = g.transA, bt = g.transB;
  const char* tf[2] = {"false", "true"};
  std::string name;
  double flops = 2.0 * (double)g.M * (double)g.N * (double)g.K * (double)g.batch;
  if (g.dt == DType::F64) name = std::string("gemm_f64_kernel<") + tf[at] + ", " + tf[bt] + ">";
  else if (g.dt == DType::C128) { name = std::string("gemm_c128_kernel<") + tf[at] + ", " + tf[bt] + ">"; flops *= 4; }
  else name = std::string("gemm_f32_kernel<") + tf[at] + ", " + tf[bt] + ", " + std::to_string(g.mixed == Mixed::BF16 ? 1 : g.mixed == Mixed::BF16X3 ? 2 : 3) + ">";
  const double bytes = ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N * (g.beta != 0.0 ? 2 : 1)) *
                       (double)dtype_size(g.dt) * (double)g.batch;
  if (c.timing_detail)
    name += " M=" + std::to_string(g.M) + " N=" + std::to_string(g.N) + " K=" + std::to_string(g.K) +
            " b=" + std::to_string(g.batch);
  // Split-K when the output tiles cannot fill the 256 CUs and K is long:
  // each z-slice of the grid reduces one K chunk into a workspace, then a
  // reduction kernel applies alpha / beta.
  const int64_t tiles = (int64_t)grid.x * grid.y * g.batch;
  int64_t S = 1, kc = 0;
  if (g.batch == 1 && tiles < 600 && g.K >= 1024) {
    S = std::min<int64_t>((1024 + tiles - 1) / tiles, g.K / 256);
    if (S >= 2) {
      kc = ((g.K + S - 1) / S + 63) / 64 * 64;
      S = (g.K + kc - 1) / kc;
    }
    if (S > 65535) S = 65535;
  }
  if (S >= 2) {
    const size_t es = dtype_size(g.dt);
    const size_t wbytes = (size_t)S * g.M * g.N * es;
    std::shared_ptr<Buf> ws(new Buf{c.pool->alloc(wbytes), wbytes, c.pool.get()});
    GemmP q = p;
    q.alpha = 1.0;
    q.beta = 0.0;
    q.C = ws->p;
    q.ldc = g.M;
    q.sC = g.M * g.N;
    q.kchunk = kc;
    q.sA = at ? kc : kc * g.lda;
    q.sB = bt ? kc * g.ldb : kc;
    grid.z = (unsigned)S;
    TimedLaunch tl(c, name, flops, bytes);
    launch(c, g, q, grid);
    const int64_t MN = g.M * g.N;
    unsigned rb = (unsigned)std::min<int64_t>((MN + 255) / 256, 4096);
    if (g.dt == DType::F64)
      hipLaunchKernelGGL(splitk_reduce<double>, dim3(rb), dim3(256), 0, c.stream, g.M, g.N, S, (const double*)ws->p,
                         g.alpha, g.beta, (double*)g.C, g.ldc);
    else if (g.dt == DType::C128)
      hipLaunchKernelGGL(splitk_reduce<double2>, dim3(rb), dim3(256), 0, c.stream, g.M, g.N, S,
                         (const double2*)ws->p, g.alpha, g.beta, (double2*)g.C, g.ldc);
    else
      hipLaunchKernelGGL(splitk_reduce<float>, dim3(rb), dim3(256), 0, c.stream, g.M, g.N, S, (const float*)ws->p,
                         g.alpha, g.beta, (float*)g.C, g.ldc);
    HIP_CHECK(hipGetLastError());
    return;
  }
  TimedLaunch tl(c, name, flops, bytes);
  launch(c, g, p, grid);
}

}  // namespace k
}  // namespace ipeps
