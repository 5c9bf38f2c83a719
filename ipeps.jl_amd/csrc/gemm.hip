// MFMA GEMM kernels for gfx950 (CDNA4), column-major, batched.
//
//  * FP64  : v_mfma_f64_16x16x4_f64, 64x64 block tile, K tile 16.
//  * C128  : the same tile with separate re/im LDS planes; each K step issues
//            four FP64 MFMAs (re*re, -im*im, re*im, im*re).
//  * F32   : FP32 storage, operands split into NSPLIT bf16 terms when staged
//            into LDS (hi = bf16(x), mid = bf16(x - hi), ...); products with
//            i + j < NSPLIT go through v_mfma_f32_16x16x32_bf16 with FP32
//            accumulation.  At full FP32 accuracy (the "6" setting) the
//            operands are staged as FP32 and multiplied by
//            v_mfma_f32_16x16x4f32 (gemm_f32x_kernel).
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

#include <algorithm>

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
  // FP32 kernel: 16-B vector access allowed on the k-contiguous A / B
  // operand and on C (pointers, leading dims and batch strides aligned)
  int vecA, vecB, vecC;
  // FP32 kernel, pre-split operands: elements per bf16 plane
  int64_t planeA, planeB;
  // output offsets (batch 1, no split-K): element (m, n) goes to
  // C[offM[m] + offN[n]] instead of C[m + n ldc] (a permuted result layout)
  const int* offM;
  const int* offN;
};

constexpr int BM = 64, BN = 64;

// XCD-aware block order.  Workgroups are dealt to the 8 XCDs round-robin in
// dispatch order, each XCD with its own L2.  The hardware block number lin is
// remapped so that XCD j works through the contiguous logical range
// [j T/8, (j+1) T/8) with the M tile fastest: the blocks an XCD runs together
// then share their N tile and K chunk, i.e. the same B panel stays in that
// XCD's L2 while the A panels stream.  Used when T is a multiple of 8.
struct BlockXYZ {
  int64_t x, y, z;
};
__device__ inline BlockXYZ xcd_block() {
  const int64_t gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
  const int64_t T = gx * gy * gz;
  int64_t lin = blockIdx.x + gx * (blockIdx.y + gy * (int64_t)blockIdx.z);
  if (T % 8 == 0) lin = (lin % 8) * (T / 8) + lin / 8;
  BlockXYZ b;
  b.x = lin % gx;
  b.y = (lin / gx) % gy;
  b.z = lin / (gx * gy);
  return b;
}

// ---------------------------------------------------------------------------
// FP64 real and complex FP64: 64 x 64 block tile, 4 waves in a 2 x 2 grid,
// each wave a 32 x 32 block of 16 x 16 f64 MFMA outputs.  The operand tiles
// are fetched into registers one K step ahead with unconditional loads (row
// indices clamped: rows past the end only feed outputs that are never
// written; in the last, partial K tile k is clamped too and the values past
// Keff are zeroed at staging), then written to the other of two LDS buffers
// while the MFMAs of the current one run: one barrier per K step.
// ---------------------------------------------------------------------------
template <typename T>
__device__ inline T zero_t();
template <>
__device__ inline double zero_t<double>() { return 0.0; }
template <>
__device__ inline double2 zero_t<double2>() { return make_double2(0.0, 0.0); }

// Staging coordinates of one operand tile of TR rows (m or n) x BK k's:
// item r of a thread is (row[r], kk[r]); `kfast` when k is the contiguous
// index of the operand in memory.
template <typename T, int TR, int BK>
struct TileLoader {
  static constexpr int NI = TR * BK / 256;
  int row[NI], kk[NI];
  const T* base[NI];  // element (clamped row, k = 0)
  int64_t ks;         // element stride along k
  T v[NI];
  __device__ inline void init(const T* X, int64_t ld, bool kfast, int64_t r0, int64_t R, int tid) {
    ks = kfast ? 1 : ld;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      if (kfast) { kk[i] = tid % BK; row[i] = tid / BK + (256 / BK) * i; }
      else { row[i] = tid % TR; kk[i] = tid / TR + (256 / TR) * i; }
      const int64_t r = min(r0 + row[i], R - 1);
      base[i] = X + (kfast ? r * ld : r);
    }
  }
  __device__ inline void fetch(int64_t k0) {
#pragma unroll
    for (int i = 0; i < NI; ++i) v[i] = base[i][(k0 + kk[i]) * ks];
  }
  __device__ inline void fetch_tail(int64_t k0, int64_t Keff) {
#pragma unroll
    for (int i = 0; i < NI; ++i) v[i] = base[i][min(k0 + kk[i], Keff - 1) * ks];
  }
  __device__ inline void zero_tail(int64_t k0, int64_t Keff) {
#pragma unroll
    for (int i = 0; i < NI; ++i)
      if (k0 + kk[i] >= Keff) v[i] = zero_t<T>();
  }
};

template <typename T>
struct RealMfma;
template <>
struct RealMfma<double> {
  typedef v4f64 acc_t;
  static __device__ inline acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <typename T, bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_real_kernel(GemmP p) {
  typedef RealMfma<T> MF;
  constexpr int BK = 16;
  __shared__ T As[2][BK][BM + 1];
  __shared__ T Bs[2][BK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const BlockXYZ bk = xcd_block();
  const int64_t m0 = bk.x * BM, n0 = bk.y * BN;
  const T* A = (const T*)p.A + bk.z * p.sA;
  const T* B = (const T*)p.B + bk.z * p.sB;
  T* C = (T*)p.C + bk.z * p.sC;
  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - bk.z * p.kchunk) : p.K;

  TileLoader<T, BM, BK> la;
  TileLoader<T, BN, BK> lb;
  la.init(A, p.lda, AT, m0, p.M, tid);
  lb.init(B, p.ldb, !BT, n0, p.N, tid);
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < la.NI; ++i) As[buf][la.kk[i]][la.row[i]] = la.v[i];
#pragma unroll
    for (int i = 0; i < lb.NI; ++i) Bs[buf][lb.kk[i]][lb.row[i]] = lb.v[i];
  };

  typename MF::acc_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = (typename MF::acc_t){0, 0, 0, 0};

  const int64_t nk = (Keff + BK - 1) / BK, nfull = Keff / BK;
  if (nfull > 0) {
    la.fetch(0);
    lb.fetch(0);
  } else {
    la.fetch_tail(0, Keff);
    lb.fetch_tail(0, Keff);
    la.zero_tail(0, Keff);
    lb.zero_tail(0, Keff);
  }
  stage(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt & 1);
    const int64_t kn = (kt + 1) * BK;
    const bool more = kt + 1 < nk, full = kt + 1 < nfull;
    if (full) {
      la.fetch(kn);
      lb.fetch(kn);
    } else if (more) {
      la.fetch_tail(kn, Keff);
      lb.fetch_tail(kn, Keff);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[cur][kk + (lane >> 4)][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[cur][kk + (lane >> 4)][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = MF::op(a[i], b[j], acc[i][j]);
    }
    if (more) {
      if (!full) {
        la.zero_tail(kn, Keff);
        lb.zero_tail(kn, Keff);
      }
      stage(cur ^ 1);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t row = m0 + wm * 32 + i * 16 + MF::row(lane, r);
        int64_t col = n0 + wn * 32 + j * 16 + (lane & 15);
        if (row < p.M && col < p.N) {
          double v = p.alpha * (double)acc[i][j][r];
          T* dst = p.offM ? C + p.offM[row] + p.offN[col] : C + row + col * p.ldc;
          if (p.beta != 0.0) v += p.beta * (double)*dst;
          *dst = (T)v;
        }
      }
}

template <bool AT, bool BT>
__global__ __launch_bounds__(256) void gemm_c128_kernel(GemmP p) {
  constexpr int BK = 8;
  __shared__ double Ar[2][BK][BM + 1], Ai[2][BK][BM + 1];
  __shared__ double Br[2][BK][BN + 1], Bi[2][BK][BN + 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const BlockXYZ bk = xcd_block();
  const int64_t m0 = bk.x * BM, n0 = bk.y * BN;
  const double2* A = (const double2*)p.A + bk.z * p.sA;
  const double2* B = (const double2*)p.B + bk.z * p.sB;
  double2* C = (double2*)p.C + bk.z * p.sC;
  const double sa = p.conjA ? -1.0 : 1.0, sb = p.conjB ? -1.0 : 1.0;
  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - bk.z * p.kchunk) : p.K;

  TileLoader<double2, BM, BK> la;
  TileLoader<double2, BN, BK> lb;
  la.init(A, p.lda, AT, m0, p.M, tid);
  lb.init(B, p.ldb, !BT, n0, p.N, tid);
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < la.NI; ++i) {
      Ar[buf][la.kk[i]][la.row[i]] = la.v[i].x;
      Ai[buf][la.kk[i]][la.row[i]] = sa * la.v[i].y;
    }
#pragma unroll
    for (int i = 0; i < lb.NI; ++i) {
      Br[buf][lb.kk[i]][lb.row[i]] = lb.v[i].x;
      Bi[buf][lb.kk[i]][lb.row[i]] = sb * lb.v[i].y;
    }
  };

  v4f64 cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = (v4f64){0.0, 0.0, 0.0, 0.0};
      ci[i][j] = (v4f64){0.0, 0.0, 0.0, 0.0};
    }

  const int64_t nk = (Keff + BK - 1) / BK, nfull = Keff / BK;
  if (nfull > 0) {
    la.fetch(0);
    lb.fetch(0);
  } else {
    la.fetch_tail(0, Keff);
    lb.fetch_tail(0, Keff);
    la.zero_tail(0, Keff);
    lb.zero_tail(0, Keff);
  }
  stage(0);
  __syncthreads();
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt & 1);
    const int64_t kn = (kt + 1) * BK;
    const bool more = kt + 1 < nk, full = kt + 1 < nfull;
    if (full) {
      la.fetch(kn);
      lb.fetch(kn);
    } else if (more) {
      la.fetch_tail(kn, Keff);
      lb.fetch_tail(kn, Keff);
    }
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        int mi = wm * 32 + i * 16 + (lane & 15);
        ar[i] = Ar[cur][kk + (lane >> 4)][mi];
        ai[i] = Ai[cur][kk + (lane >> 4)][mi];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        int nj = wn * 32 + j * 16 + (lane & 15);
        br[j] = Br[cur][kk + (lane >> 4)][nj];
        bi[j] = Bi[cur][kk + (lane >> 4)][nj];
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
    if (more) {
      if (!full) {
        la.zero_tail(kn, Keff);
        lb.zero_tail(kn, Keff);
      }
      stage(cur ^ 1);
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

// Staging of a TR x 32 slab of op(X) (rows r = m or n, 32 consecutive k).
// Item (r, g) is the 8 values k = 8g .. 8g+7 of row r.
//   k-contiguous operand (X[k + r ld]): item q -> g = q & 3, r = q >> 2, so
//     four lanes read the 128 B of one row (two float4 each when aligned);
//   row-contiguous operand (X[r + k ld]): item q -> r = q % TR, g = q / TR,
//     each lane reads 8 scalars down k and consecutive lanes hit consecutive r.
// The loads are unconditional, so a K step's loads stay in flight together:
// the row index is clamped to the last row by the caller (rows past the end
// only feed output rows / columns that are never written), and in the last,
// partial K tile of a block k is clamped too and the values past the end are
// zeroed at staging (zero8).
template <bool KV>
__device__ inline void load8(const float* __restrict__ X, int64_t ld, int64_t rc, int64_t kb, bool vec,
                             float (&v)[8]) {
  if (KV) {
    const float* src = X + kb + rc * ld;
    if (vec) {
      const float4 x = *reinterpret_cast<const float4*>(src);
      const float4 y = *reinterpret_cast<const float4*>(src + 4);
      v[0] = x.x; v[1] = x.y; v[2] = x.z; v[3] = x.w;
      v[4] = y.x; v[5] = y.y; v[6] = y.z; v[7] = y.w;
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = src[e];
    }
  } else {
    const float* src = X + rc + kb * ld;
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = src[e * ld];
  }
}
template <bool KV>
__device__ inline void load8_tail(const float* __restrict__ X, int64_t ld, int64_t rc, int64_t kb, int64_t Keff,
                                  float (&v)[8]) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int64_t k = kb + e < Keff ? kb + e : Keff - 1;
    v[e] = KV ? X[k + rc * ld] : X[rc + k * ld];
  }
}
__device__ inline void zero8(float (&v)[8], int64_t kb, int64_t Keff) {
#pragma unroll
  for (int e = 0; e < 8; ++e)
    if (kb + e >= Keff) v[e] = 0.f;
}

// hi / mid / lo bf16 planes of 8 values, 16 B per plane (plane stride in elements)
template <int NSPLIT>
__device__ inline void split_store(unsigned short* dst, int plane, const float (&v)[8]) {
  float r[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) r[e] = v[e];
#pragma unroll
  for (int s = 0; s < NSPLIT; ++s) {
    v8u16 h;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const unsigned short b = bf16_rne(r[e]);
      h[e] = b;
      r[e] -= bf16_to_f(b);
    }
    *reinterpret_cast<v8u16*>(dst + s * plane) = h;
  }
}

// One operand of the FP32 kernel: TR rows (m or n) x BK k's per K step,
// staged as NSPLIT bf16 planes.  MODE 0: FP32, row-contiguous (X[r + k ld]);
// 1: FP32, k-contiguous (X[k + r ld]); 2: pre-split, NSPLIT bf16 planes
// P_s[k + r ld] (k-contiguous, zero-padded in k to a multiple of 32, planes
// `plane` elements apart; split_operand below).  Item (r, g) of a thread is
// the 8 values k = 8g .. 8g+7 of row r.
template <int MODE, int NSPLIT, int TR>
struct F32Operand {
  static constexpr int NI = TR / 64;
  int r[NI], g[NI];
  int64_t rc[NI];  // clamped rows
  float v[MODE == 2 ? 1 : NI][8];
  v8u16 h[MODE == 2 ? NI : 1][NSPLIT];
  __device__ inline void init(int tid, int64_t r0, int64_t R) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int q = tid + 256 * i;
      if (MODE == 0) { r[i] = q % TR; g[i] = q / TR; }
      else { g[i] = q & 3; r[i] = q >> 2; }
      rc[i] = min(r0 + r[i], R - 1);
    }
  }
  __device__ inline void fetch(const void* X, int64_t ld, int64_t plane, int64_t k0, bool vec) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      if constexpr (MODE == 2) {
        const unsigned short* src = (const unsigned short*)X + rc[i] * ld + k0 + 8 * g[i];
#pragma unroll
        for (int s = 0; s < NSPLIT; ++s) h[i][s] = *reinterpret_cast<const v8u16*>(src + s * plane);
      } else {
        load8<MODE == 1>((const float*)X, ld, rc[i], k0 + 8 * g[i], vec, v[i]);
      }
    }
  }
  // last, partial K tile of a block (pre-split planes are zero-padded)
  __device__ inline void fetch_tail(const void* X, int64_t ld, int64_t plane, int64_t k0, int64_t Keff) {
    if constexpr (MODE == 2) {
      fetch(X, ld, plane, k0, true);
    } else {
#pragma unroll
      for (int i = 0; i < NI; ++i) load8_tail<MODE == 1>((const float*)X, ld, rc[i], k0 + 8 * g[i], Keff, v[i]);
    }
  }
  __device__ inline void zero_tail(int64_t k0, int64_t Keff) {
    if constexpr (MODE != 2) {
#pragma unroll
      for (int i = 0; i < NI; ++i) zero8(v[i], k0 + 8 * g[i], Keff);
    }
  }
  // dst: plane 0 of one LDS buffer, rows LDK apart, planes `lplane` apart
  template <int LDK>
  __device__ inline void stage(unsigned short* dst, int lplane) const {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      unsigned short* d = dst + r[i] * LDK + 8 * g[i];
      if constexpr (MODE == 2) {
#pragma unroll
        for (int s = 0; s < NSPLIT; ++s) *reinterpret_cast<v8u16*>(d + s * lplane) = h[i][s];
      } else {
        split_store<NSPLIT>(d, lplane, v[i]);
      }
    }
  }
};

// FP32 storage, split-bf16 MFMA.  Block tile TBM x TBN (128 x 128,
// 128 x 64 or 64 x 192), K tile 32, 256 threads = 4 waves in a 2 x 2 grid,
// each wave owning a (TBM/2) x (TBN/2) block of 16 x 16 MFMA outputs.  The
// next K tile is fetched into registers while the MFMAs of the current one
// run, then split into NSPLIT bf16 planes (or copied, for a pre-split
// operand) and written to the other of two LDS buffers: one barrier per K
// step.  LDS rows are k-contiguous and 80 B apart, so the 16-B fragment reads
// of 16 consecutive rows hit distinct banks.
template <int MA, int MB, int NSPLIT, int TBM, int TBN>
__global__ __launch_bounds__(256) void gemm_f32_kernel(GemmP p) {
  constexpr int BK = 32, LDK = BK + 8;
  constexpr int WTM = TBM / 2, WTN = TBN / 2;
  constexpr int TI = WTM / 16, TJ = WTN / 16;
  constexpr int PA = TBM * LDK, PB = TBN * LDK;
  __shared__ __attribute__((aligned(16))) unsigned short As[2][NSPLIT][TBM][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short Bs[2][NSPLIT][TBN][LDK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const BlockXYZ bk = xcd_block();
  const int64_t m0 = bk.x * TBM, n0 = bk.y * TBN;
  const void* A = MA == 2 ? (const void*)((const unsigned short*)p.A + bk.z * p.sA) : (const void*)((const float*)p.A + bk.z * p.sA);
  const void* B = MB == 2 ? (const void*)((const unsigned short*)p.B + bk.z * p.sB) : (const void*)((const float*)p.B + bk.z * p.sB);
  float* C = (float*)p.C + bk.z * p.sC;
  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - bk.z * p.kchunk) : p.K;
  const bool vecA = p.vecA != 0, vecB = p.vecB != 0;

  F32Operand<MA, NSPLIT, TBM> oa;
  F32Operand<MB, NSPLIT, TBN> ob;
  oa.init(tid, m0, p.M);
  ob.init(tid, n0, p.N);
  auto fetch = [&](int64_t k0) {
    oa.fetch(A, p.lda, p.planeA, k0, vecA);
    ob.fetch(B, p.ldb, p.planeB, k0, vecB);
  };
  auto fetch_tail = [&](int64_t k0) {
    oa.fetch_tail(A, p.lda, p.planeA, k0, Keff);
    ob.fetch_tail(B, p.ldb, p.planeB, k0, Keff);
  };
  auto stage = [&](int buf) {
    oa.template stage<LDK>(&As[buf][0][0][0], PA);
    ob.template stage<LDK>(&Bs[buf][0][0][0], PB);
  };
  auto stage_tail = [&](int buf, int64_t k0) {
    oa.zero_tail(k0, Keff);
    ob.zero_tail(k0, Keff);
    stage(buf);
  };

  v4f32 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = (v4f32){0.f, 0.f, 0.f, 0.f};

  // K tiles [0, nfull) lie inside the block's K range; the last tile is
  // partial when Keff is not a multiple of BK.
  const int64_t nk = (Keff + BK - 1) / BK, nfull = Keff / BK;
  if (nfull > 0) {
    fetch(0);
    stage(0);
  } else {
    fetch_tail(0);
    stage_tail(0, 0);
  }
  __syncthreads();
  const int kq = 8 * (lane >> 4), lr = lane & 15;
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt & 1);
    const int64_t kn = (kt + 1) * BK;
    const bool more = kt + 1 < nk, full = kt + 1 < nfull;
    if (full) fetch(kn);
    else if (more) fetch_tail(kn);
    v8bf16 a[NSPLIT][TI], b[NSPLIT][TJ];
#pragma unroll
    for (int s = 0; s < NSPLIT; ++s) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
        a[s][i] = __builtin_bit_cast(v8bf16, *reinterpret_cast<const v8u16*>(&As[cur][s][wm * WTM + 16 * i + lr][kq]));
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        b[s][j] = __builtin_bit_cast(v8bf16, *reinterpret_cast<const v8u16*>(&Bs[cur][s][wn * WTN + 16 * j + lr][kq]));
    }
    // smallest products first
#pragma unroll
    for (int t = NSPLIT - 1; t >= 0; --t)
#pragma unroll
      for (int sa = 0; sa <= t; ++sa)
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
          for (int j = 0; j < TJ; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[t - sa][j], acc[i][j], 0, 0, 0);
    if (full) stage(cur ^ 1);
    else if (more) stage_tail(cur ^ 1, kn);
    __syncthreads();
  }

  const float al = (float)p.alpha, be = (float)p.beta;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int64_t col = n0 + wn * WTN + 16 * j + lr;
      const int64_t row0 = m0 + wm * WTM + 16 * i + 4 * (lane >> 4);
      if (col >= p.N || row0 >= p.M) continue;
      if (p.offM) {
        const int on = p.offN[col];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row0 + r < p.M) {
            float* d = C + p.offM[row0 + r] + on;
            float v = al * acc[i][j][r];
            if (p.beta != 0.0) v += be * *d;
            *d = v;
          }
        continue;
      }
      float* dst = C + row0 + col * p.ldc;
      if (p.vecC && row0 + 3 < p.M) {
        float4 v = make_float4(al * acc[i][j][0], al * acc[i][j][1], al * acc[i][j][2], al * acc[i][j][3]);
        if (p.beta != 0.0) {
          const float4 o = *reinterpret_cast<const float4*>(dst);
          v.x += be * o.x; v.y += be * o.y; v.z += be * o.z; v.w += be * o.w;
        }
        *reinterpret_cast<float4*>(dst) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row0 + r < p.M) {
            float v = al * acc[i][j][r];
            if (p.beta != 0.0) v += be * dst[r];
            dst[r] = v;
          }
      }
    }
}

// ---------------------------------------------------------------------------
// FP32 storage, FP32 MFMA products (the full-accuracy "6" setting): the block
// structure of gemm_f32_kernel with the operand tiles staged in LDS as FP32,
// k-contiguous rows of 32 (+4 pad: the 16-B reads of 16 consecutive rows hit
// distinct banks).  Each v_mfma_f32_16x16x4f32 sums four k; lane group
// g = lane / 16 takes physical k = 16 h + 4 g + kk at MFMA kk of half h, the
// same for A and B, so one 16-B LDS read per operand fragment feeds four
// MFMAs.  Block tile TBM x TBN (128 x 64, 128 x 96, 128 x 128 or 64 x 192),
// 256 threads = 4 waves in a 2 x 2 grid.
// ---------------------------------------------------------------------------
// One operand: TR rows x 32 k per K step, item (r, g) = 8 values k = 8g .. 8g+7
// of row r; MODE 0 row-contiguous (X[r + k ld]), 1 k-contiguous (X[k + r ld]).
// Items past TR x 4 (TR not a multiple of 64) load a valid row and are not staged.
template <int MODE, int TR>
struct F32xOperand {
  static constexpr int NI = (TR * 4 + 255) / 256;
  int r[NI], g[NI];
  bool live[NI];
  int64_t rc[NI];
  float v[NI][8];
  __device__ inline void init(int tid, int64_t r0, int64_t R) {
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int q = tid + 256 * i;
      live[i] = q < TR * 4;
      const int qq = live[i] ? q : 0;
      if (MODE == 0) { r[i] = qq % TR; g[i] = qq / TR; }
      else { g[i] = qq & 3; r[i] = qq >> 2; }
      rc[i] = min(r0 + r[i], R - 1);
    }
  }
  __device__ inline void fetch(const float* X, int64_t ld, int64_t k0, bool vec) {
#pragma unroll
    for (int i = 0; i < NI; ++i) load8<MODE == 1>(X, ld, rc[i], k0 + 8 * g[i], vec, v[i]);
  }
  __device__ inline void fetch_tail(const float* X, int64_t ld, int64_t k0, int64_t Keff) {
#pragma unroll
    for (int i = 0; i < NI; ++i) load8_tail<MODE == 1>(X, ld, rc[i], k0 + 8 * g[i], Keff, v[i]);
  }
  __device__ inline void zero_tail(int64_t k0, int64_t Keff) {
#pragma unroll
    for (int i = 0; i < NI; ++i) zero8(v[i], k0 + 8 * g[i], Keff);
  }
  template <int LDK>
  __device__ inline void stage(float* dst) const {
#pragma unroll
    for (int i = 0; i < NI; ++i)
      if (live[i]) {
        float* d = dst + r[i] * LDK + 8 * g[i];
        *reinterpret_cast<float4*>(d) = make_float4(v[i][0], v[i][1], v[i][2], v[i][3]);
        *reinterpret_cast<float4*>(d + 4) = make_float4(v[i][4], v[i][5], v[i][6], v[i][7]);
      }
  }
};

template <int MA, int MB, int TBM, int TBN>
__global__ __launch_bounds__(256) void gemm_f32x_kernel(GemmP p) {
  constexpr int BK = 32, LDK = BK + 4;
  constexpr int WTM = TBM / 2, WTN = TBN / 2;
  constexpr int TI = WTM / 16, TJ = WTN / 16;
  __shared__ __attribute__((aligned(16))) float As[2][TBM][LDK];
  __shared__ __attribute__((aligned(16))) float Bs[2][TBN][LDK];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const BlockXYZ bk = xcd_block();
  const int64_t m0 = bk.x * TBM, n0 = bk.y * TBN;
  const float* A = (const float*)p.A + bk.z * p.sA;
  const float* B = (const float*)p.B + bk.z * p.sB;
  float* C = (float*)p.C + bk.z * p.sC;
  // split-K: block z covers k in [z*kchunk, min(K, (z+1)*kchunk)) (A/B offsets via sA/sB)
  const int64_t Keff = p.kchunk ? min(p.kchunk, p.K - bk.z * p.kchunk) : p.K;
  const bool vecA = p.vecA != 0, vecB = p.vecB != 0;

  F32xOperand<MA, TBM> oa;
  F32xOperand<MB, TBN> ob;
  oa.init(tid, m0, p.M);
  ob.init(tid, n0, p.N);
  auto stage = [&](int buf) {
    oa.template stage<LDK>(&As[buf][0][0]);
    ob.template stage<LDK>(&Bs[buf][0][0]);
  };

  v4f32 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = (v4f32){0.f, 0.f, 0.f, 0.f};

  const int64_t nk = (Keff + BK - 1) / BK, nfull = Keff / BK;
  if (nfull > 0) {
    oa.fetch(A, p.lda, 0, vecA);
    ob.fetch(B, p.ldb, 0, vecB);
  } else {
    oa.fetch_tail(A, p.lda, 0, Keff);
    ob.fetch_tail(B, p.ldb, 0, Keff);
    oa.zero_tail(0, Keff);
    ob.zero_tail(0, Keff);
  }
  stage(0);
  __syncthreads();
  const int lr = lane & 15, kq = 4 * (lane >> 4);
  for (int64_t kt = 0; kt < nk; ++kt) {
    const int cur = (int)(kt & 1);
    const int64_t kn = (kt + 1) * BK;
    const bool more = kt + 1 < nk, full = kt + 1 < nfull;
    if (full) {
      oa.fetch(A, p.lda, kn, vecA);
      ob.fetch(B, p.ldb, kn, vecB);
    } else if (more) {
      oa.fetch_tail(A, p.lda, kn, Keff);
      ob.fetch_tail(B, p.ldb, kn, Keff);
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      float4 a[TI], b[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) a[i] = *reinterpret_cast<const float4*>(&As[cur][wm * WTM + 16 * i + lr][16 * h + kq]);
#pragma unroll
      for (int j = 0; j < TJ; ++j) b[j] = *reinterpret_cast<const float4*>(&Bs[cur][wn * WTN + 16 * j + lr][16 * h + kq]);
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].x, b[j].x, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].y, b[j].y, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].z, b[j].z, acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i].w, b[j].w, acc[i][j], 0, 0, 0);
        }
    }
    if (more) {
      if (!full) {
        oa.zero_tail(kn, Keff);
        ob.zero_tail(kn, Keff);
      }
      stage(cur ^ 1);
    }
    __syncthreads();
  }

  const float al = (float)p.alpha, be = (float)p.beta;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int64_t col = n0 + wn * WTN + 16 * j + lr;
      const int64_t row0 = m0 + wm * WTM + 16 * i + 4 * (lane >> 4);
      if (col >= p.N || row0 >= p.M) continue;
      if (p.offM) {
        const int on = p.offN[col];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row0 + r < p.M) {
            float* d = C + p.offM[row0 + r] + on;
            float v = al * acc[i][j][r];
            if (p.beta != 0.0) v += be * *d;
            *d = v;
          }
        continue;
      }
      float* dst = C + row0 + col * p.ldc;
      if (p.vecC && row0 + 3 < p.M) {
        float4 v = make_float4(al * acc[i][j][0], al * acc[i][j][1], al * acc[i][j][2], al * acc[i][j][3]);
        if (p.beta != 0.0) {
          const float4 o = *reinterpret_cast<const float4*>(dst);
          v.x += be * o.x; v.y += be * o.y; v.z += be * o.z; v.w += be * o.w;
        }
        *reinterpret_cast<float4*>(dst) = v;
      } else {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (row0 + r < p.M) {
            float v = al * acc[i][j][r];
            if (p.beta != 0.0) v += be * dst[r];
            dst[r] = v;
          }
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

// One output element per thread (MN < 2^32: 32-bit row / column split); the
// S slice loads are independent and issued together.
template <typename T>
__global__ __launch_bounds__(256) void splitk_reduce(unsigned M, int64_t MN, int S, const T* __restrict__ W,
                                                     double alpha, double beta, T* __restrict__ C, int64_t ldc,
                                                     const int* __restrict__ offM, const int* __restrict__ offN) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= MN) return;
  const unsigned n = (unsigned)t / M, m = (unsigned)t - n * M;
  T acc = zero_of<T>();
  int s = 0;
  for (; s + 4 <= S; s += 4) {
    const T w0 = W[s * MN + t], w1 = W[(s + 1) * MN + t], w2 = W[(s + 2) * MN + t], w3 = W[(s + 3) * MN + t];
    acc = add_(add_(add_(add_(acc, w0), w1), w2), w3);
  }
  for (; s < S; ++s) acc = add_(acc, W[s * MN + t]);
  T* dst = offM ? C + offM[m] + offN[n] : C + m + (int64_t)n * ldc;
  *dst = beta != 0.0 ? axpby_(alpha, acc, beta, *dst) : axpby_(alpha, acc, 0.0, zero_of<T>());
}

// W2[y][e] = sum of W[s][e] over s in [y F, min(S, (y+1) F)): folds many
// split-K slices so that the final reduction has few terms per element and
// the fold itself runs on S/F x MN/256 blocks
template <typename T>
__global__ __launch_bounds__(256) void splitk_fold(int64_t MN, int64_t S, int64_t F, const T* __restrict__ W,
                                                   T* __restrict__ W2) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= MN) return;
  const int64_t s0 = (int64_t)blockIdx.y * F, s1 = min(S, s0 + F);
  T acc = zero_of<T>();
  for (int64_t s = s0; s < s1; ++s) acc = add_(acc, W[s * MN + e]);
  W2[(int64_t)blockIdx.y * MN + e] = acc;
}

// ---------------------------------------------------------------------------
// Operand pre-splitting: op(X) (R rows x K) -> NSPLIT bf16 planes
// P_s[k + r Kp], k-contiguous, zero for K <= k < Kp (Kp a multiple of 32).
// Item (r, g) = the 8 values k = 8g .. 8g+7 of row r; consecutive threads take
// consecutive g for a k-contiguous X, consecutive r otherwise (coalesced reads).
// ---------------------------------------------------------------------------
// k-contiguous X: item (r, g) per thread, consecutive threads along k.
template <int NSPLIT>
__global__ __launch_bounds__(256) void split_operand_kv_kernel(const float* __restrict__ X, int64_t ld, int64_t R,
                                                               int64_t K, int64_t Kp, int vec,
                                                               unsigned short* __restrict__ P) {
  const int64_t ng = Kp / 8, items = R * ng;
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= items) return;
  const int64_t r = q / ng, kb = 8 * (q % ng);
  float v[8];
  if (kb + 8 <= K) {
    load8<true>(X, ld, r, kb, vec != 0, v);
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = kb + e < K ? X[kb + e + r * ld] : 0.f;
  }
  split_store<NSPLIT>(P + r * Kp + kb, (int)(R * Kp), v);
}

// row-contiguous X: 64 rows x 64 k per block, read along r into LDS, written
// along k (8 threads cover the 64 k of one row).
template <int NSPLIT>
__global__ __launch_bounds__(256) void split_operand_rows_kernel(const float* __restrict__ X, int64_t ld, int64_t R,
                                                                 int64_t K, int64_t Kp,
                                                                 unsigned short* __restrict__ P) {
  __shared__ float tile[64][65];  // [k][r]
  const int tid = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int idx = tid + 256 * i, rl = idx & 63, kl = idx >> 6;
    const int64_t r = min(r0 + rl, R - 1), k = k0 + kl;
    tile[kl][rl] = k < K ? X[r + k * ld] : 0.f;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int item = tid + 256 * j, g = item & 7, rl = item >> 3;
    const int64_t r = r0 + rl, kb = k0 + 8 * g;
    if (r >= R || kb >= Kp) continue;
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = tile[8 * g + e][rl];
    split_store<NSPLIT>(P + r * Kp + kb, (int)(R * Kp), v);
  }
}

// ---------------------------------------------------------------------------
// launcher
// ---------------------------------------------------------------------------
// FP32 block tile: 128 x 64 for outputs up to 64 columns; 64 x 192 for
// 129..192 columns (one tile row spans the output, so the M x K operand is
// streamed from HBM once; a 256-row tile measured slower on the skinny
// chi-sized products: fewer, single-wave-per-SIMD blocks); else 128 x 128.
static int f32_tile_n(int64_t N) { return N <= 64 ? 64 : (N > 128 && N <= 192) ? 192 : 128; }
// FP32-MFMA kernel: the column tile (64, 96, 128, 192) that pads N least, the
// wider one on a tie
static int f32x_tile_n(int64_t N) {
  int best = 64;
  int64_t waste = (N + 63) / 64 * 64 - N;
  for (int tn : {96, 128, 192}) {
    const int64_t w = (N + tn - 1) / tn * tn - N;
    if (w <= waste) { best = tn; waste = w; }
  }
  return best;
}
static int f32_tile_m(int64_t N, Mixed) { return f32_tile_n(N) == 192 ? 64 : 128; }

static bool al16(const void* ptr) { return ((uintptr_t)ptr & 15u) == 0; }

// FP32 data at full FP32 accuracy (the bf16x6 setting) runs on FP32 MFMA
// (gemm_f32x_kernel); the other FP32 settings on the split-bf16 kernel.
static bool real_f32(const GemmArgs& g) { return g.dt == DType::F32 && g.mixed == Mixed::BF16X6; }
static bool split_f32(const GemmArgs& g) { return g.dt == DType::F32 && g.mixed != Mixed::BF16X6; }

// 16-B access flags of the FP32 kernel for operands as given in p
static void set_vec_flags(const GemmArgs& g, GemmP& p) {
  const bool one = g.batch == 1;
  p.vecA = g.transA && al16(p.A) && p.lda % 4 == 0 && (one || p.sA % 4 == 0);
  p.vecB = !g.transB && al16(p.B) && p.ldb % 4 == 0 && (one || p.sB % 4 == 0);
  p.vecC = al16(p.C) && p.ldc % 4 == 0 && (p.sC % 4 == 0 || (one && p.kchunk == 0));
}

// ma / mb: operand modes of the FP32 kernel (0 row-contiguous, 1 k-contiguous, 2 pre-split)
static void launch(Ctx& c, const GemmArgs& g, const GemmP& p, dim3 grid, int ma, int mb, int tn) {
  dim3 block(256);
  const bool at = g.transA, bt = g.transB;
  if (g.dt == DType::F64) {
    if (!at && !bt) hipLaunchKernelGGL((gemm_real_kernel<double, false, false>), grid, block, 0, c.stream, p);
    else if (at && !bt) hipLaunchKernelGGL((gemm_real_kernel<double, true, false>), grid, block, 0, c.stream, p);
    else if (!at && bt) hipLaunchKernelGGL((gemm_real_kernel<double, false, true>), grid, block, 0, c.stream, p);
    else hipLaunchKernelGGL((gemm_real_kernel<double, true, true>), grid, block, 0, c.stream, p);
  } else if (real_f32(g)) {
#define IPEPS_F32X_MB(MA, TM, TN)                                                                       \
  do {                                                                                                  \
    if (mb == 0) hipLaunchKernelGGL((gemm_f32x_kernel<MA, 0, TM, TN>), grid, block, 0, c.stream, p);     \
    else hipLaunchKernelGGL((gemm_f32x_kernel<MA, 1, TM, TN>), grid, block, 0, c.stream, p);             \
  } while (0)
#define IPEPS_F32X(TM, TN)                      \
  do {                                          \
    if (ma == 0) IPEPS_F32X_MB(0, TM, TN);      \
    else IPEPS_F32X_MB(1, TM, TN);              \
  } while (0)
    if (tn == 64) IPEPS_F32X(128, 64);
    else if (tn == 96) IPEPS_F32X(128, 96);
    else if (tn == 192) IPEPS_F32X(64, 192);
    else IPEPS_F32X(128, 128);
#undef IPEPS_F32X
#undef IPEPS_F32X_MB
  } else if (g.dt == DType::C128) {
    if (!at && !bt) hipLaunchKernelGGL((gemm_c128_kernel<false, false>), grid, block, 0, c.stream, p);
    else if (at && !bt) hipLaunchKernelGGL((gemm_c128_kernel<true, false>), grid, block, 0, c.stream, p);
    else if (!at && bt) hipLaunchKernelGGL((gemm_c128_kernel<false, true>), grid, block, 0, c.stream, p);
    else hipLaunchKernelGGL((gemm_c128_kernel<true, true>), grid, block, 0, c.stream, p);
  } else {
#define IPEPS_F32_LAUNCH_MB(MA, NS, TM, TN)                                                                 \
  do {                                                                                                      \
    if (mb == 0) hipLaunchKernelGGL((gemm_f32_kernel<MA, 0, NS, TM, TN>), grid, block, 0, c.stream, p);      \
    else if (mb == 1) hipLaunchKernelGGL((gemm_f32_kernel<MA, 1, NS, TM, TN>), grid, block, 0, c.stream, p); \
    else hipLaunchKernelGGL((gemm_f32_kernel<MA, 2, NS, TM, TN>), grid, block, 0, c.stream, p);              \
  } while (0)
#define IPEPS_F32_LAUNCH_T(NS, TM, TN)            \
  do {                                            \
    if (ma == 0) IPEPS_F32_LAUNCH_MB(0, NS, TM, TN); \
    else if (ma == 1) IPEPS_F32_LAUNCH_MB(1, NS, TM, TN); \
    else IPEPS_F32_LAUNCH_MB(2, NS, TM, TN);      \
  } while (0)
#define IPEPS_F32_LAUNCH(NS, TM192)                                   \
  do {                                                                \
    if (tn == 64) IPEPS_F32_LAUNCH_T(NS, 128, 64);                    \
    else if (tn == 192) IPEPS_F32_LAUNCH_T(NS, TM192, 192);           \
    else IPEPS_F32_LAUNCH_T(NS, 128, 128);                            \
  } while (0)
    switch (g.mixed) {
      case Mixed::BF16: IPEPS_F32_LAUNCH(1, 64); break;
      case Mixed::BF16X3: IPEPS_F32_LAUNCH(2, 64); break;
      case Mixed::BF16X6: throw Error(IPEPS_ERR_INTERNAL, "gemm: FP32 products run on gemm_f32x_kernel");
    }
#undef IPEPS_F32_LAUNCH
#undef IPEPS_F32_LAUNCH_T
#undef IPEPS_F32_LAUNCH_MB
  }
  HIP_CHECK(hipGetLastError());
}

// bf16 planes of the split kernel (the "6" setting runs on FP32 MFMA instead)
static int nsplit_of(Mixed m) { return m == Mixed::BF16 ? 1 : 2; }

// Pre-split op(X) (R rows x K; kv: X is k-contiguous, X[k + r ld]) into
// NSPLIT bf16 planes of R x Kp, Kp = K rounded up to 32; returns the buffer.
static std::shared_ptr<Buf> split_operand_new(Ctx& c, const float* X, int64_t ld, bool kv, int64_t R, int64_t K,
                                              int ns, int64_t* Kp_out);
// The same through the context's cache of registered operands (SplitCacheScope).
static std::shared_ptr<Buf> split_operand(Ctx& c, const float* X, int64_t ld, bool kv, int64_t R, int64_t K,
                                          int ns, int64_t* Kp_out) {
  const char* px = (const char*)X;
  bool reg = false;
  for (const auto& r : c.split_src)
    if (px >= r.first && px < r.first + r.second) reg = true;
  if (!reg) return split_operand_new(c, X, ld, kv, R, K, ns, Kp_out);
  for (const auto& e : c.split_cache)
    if (e.p == X && e.ld == ld && e.kv == kv && e.R == R && e.K == K && e.ns == ns) {
      *Kp_out = e.kp;
      return e.planes;
    }
  auto planes = split_operand_new(c, X, ld, kv, R, K, ns, Kp_out);
  c.split_cache.push_back(Ctx::SplitEntry{X, ld, R, K, *Kp_out, kv, ns, planes});
  return planes;
}

static std::shared_ptr<Buf> split_operand_new(Ctx& c, const float* X, int64_t ld, bool kv, int64_t R, int64_t K,
                                              int ns, int64_t* Kp_out) {
  const int64_t Kp = (K + 31) / 32 * 32;
  const size_t bytes = (size_t)ns * R * Kp * 2;
  std::shared_ptr<Buf> buf(new Buf{c.pool->alloc(bytes), bytes, c.pool.get()});
  const int64_t items = R * (Kp / 8);
  TimedLaunch tl(c, "split_operand", 0.0, (double)R * K * 4 + (double)bytes);
  const dim3 grid((unsigned)((items + 255) / 256));
  unsigned short* P = (unsigned short*)buf->p;
  const int vec = al16(X) && ld % 4 == 0;
  const dim3 tgrid((unsigned)((R + 63) / 64), (unsigned)((Kp + 63) / 64));
#define IPEPS_SPLIT(NS)                                                                                              \
  do {                                                                                                               \
    if (kv) hipLaunchKernelGGL((split_operand_kv_kernel<NS>), grid, dim3(256), 0, c.stream, X, ld, R, K, Kp, vec, P); \
    else hipLaunchKernelGGL((split_operand_rows_kernel<NS>), tgrid, dim3(256), 0, c.stream, X, ld, R, K, Kp, P);     \
  } while (0)
  if (ns == 1) IPEPS_SPLIT(1);
  else IPEPS_SPLIT(2);
#undef IPEPS_SPLIT
  HIP_CHECK(hipGetLastError());
  *Kp_out = Kp;
  return buf;
}

// ---------------------------------------------------------------------------
// Low-K FP32 products: C = alpha op(A) op(B) + beta C with K <= 96 and
// N <= 96 over many rows (the density-matrix absorptions of the energy:
// M = chi^2 D^2 d, N, K in {D^2, D^2 d}).  Such a product moves 4 (M K + M N)
// bytes for 2 M N K flops, well below the ridge point of any MFMA type, so it
// is HBM-bound: FP32 MFMA (v_mfma_f32_16x16x4f32, FP32 products) costs nothing
// extra and needs no bf16 planes.  A streaming kernel: op(B) sits in LDS for
// the whole launch (zero beyond K and N); each wave of the persistent blocks
// walks 32-row groups of op(A) and reads them straight into MFMA operand
// registers, the next group's loads in flight during the current group's
// MFMAs; the accumulators go straight to C (four consecutive rows per lane,
// 16-B stores where C allows).  The K axis is taken in chunks of 16: lane group
// g of MFMA s in chunk h holds k = 16 h + 4 g + s, so a k-contiguous op(A)
// gives each lane one 16-B load per chunk and row; op(B) is read with the same
// k mapping.
// ---------------------------------------------------------------------------
constexpr int LK_KMAX = 96, LK_NMAX = 96, LK_KC = LK_KMAX / 16, LK_NT = LK_NMAX / 16;
constexpr int LK_TPB = 256, LK_RG = 32;  // threads per block, rows per wave group

__host__ __device__ inline int lowk_lds_floats(int K, int N) {
  return (K + 15) / 16 * 16 * ((N + 15) / 16 * 16 + 1);
}

// The 32-row group's A fragments: a[t][h][s] = op(A)[row 16 t + lr][k = 16 h + 4 lg + s]
// of rows m0 .. m0+31 (rows past M read row M - 1; k past K read a valid
// element, which meets a zero row of op(B)).
template <int TA>
__device__ inline void lowk_fetch(const float* __restrict__ A, int lda, int M, int K, int KC, int m0, int lr, int lg,
                                  bool vec, float (&a)[2][LK_KC][4]) {
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int m = min(m0 + 16 * t + lr, M - 1);
#pragma unroll
    for (int h = 0; h < LK_KC; ++h) {
      if (h >= KC) break;
      const int k0 = 16 * h + 4 * lg;
      if (TA) {  // k contiguous: A[k + m lda]
        if (vec && k0 + 3 < K) {
          const float4 v = *reinterpret_cast<const float4*>(A + k0 + (int64_t)m * lda);
          a[t][h][0] = v.x; a[t][h][1] = v.y; a[t][h][2] = v.z; a[t][h][3] = v.w;
        } else {
#pragma unroll
          for (int s = 0; s < 4; ++s) a[t][h][s] = A[min(k0 + s, K - 1) + (int64_t)m * lda];
        }
      } else {  // m contiguous: A[m + k lda]
#pragma unroll
        for (int s = 0; s < 4; ++s) a[t][h][s] = A[m + (int64_t)min(k0 + s, K - 1) * lda];
      }
    }
  }
}

template <int TA>
__global__ __launch_bounds__(LK_TPB) void gemm_lowk_f32_kernel(GemmP p, int transB) {
  extern __shared__ float lk_smem[];
  const int M = (int)p.M, N = (int)p.N, K = (int)p.K;
  const int KC = (K + 15) / 16, NT = (N + 15) / 16, LDB = 16 * NT + 1;
  float* Bs = lk_smem;  // Bs[k LDB + n], k < 16 KC
  const float* A = (const float*)p.A;
  const float* B = (const float*)p.B;
  float* C = (float*)p.C;
  const int lda = (int)p.lda, ldb = (int)p.ldb, ldc = (int)p.ldc;
  const float alpha = (float)p.alpha, beta = (float)p.beta;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lg = lane >> 4;
  for (int e = tid; e < 16 * KC * LDB; e += LK_TPB) {
    const int k = e / LDB, n = e - k * LDB;
    float v = 0.f;
    if (k < K && n < N) v = transB ? B[n + (int64_t)k * ldb] : B[k + (int64_t)n * ldb];
    Bs[e] = v;
  }
  __syncthreads();
  const bool vecA = TA && ((uintptr_t)A & 15) == 0 && lda % 4 == 0;
  const bool vecC = ((uintptr_t)C & 15) == 0 && ldc % 4 == 0;
  const int ngroups = (M + LK_RG - 1) / LK_RG;
  const int gstride = (int)gridDim.x * (LK_TPB / 64);
  int g = (int)blockIdx.x * (LK_TPB / 64) + wave;
  float a[2][LK_KC][4], an[2][LK_KC][4];
  if (g < ngroups) lowk_fetch<TA>(A, lda, M, K, KC, g * LK_RG, lr, lg, vecA, a);
  for (; g < ngroups; g += gstride) {
    const int gn = g + gstride;
    if (gn < ngroups) lowk_fetch<TA>(A, lda, M, K, KC, gn * LK_RG, lr, lg, vecA, an);
    v4f32 acc[2][LK_NT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int j = 0; j < LK_NT; ++j) acc[t][j] = (v4f32){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int h = 0; h < LK_KC; ++h) {
      if (h >= KC) break;
#pragma unroll
      for (int sidx = 0; sidx < 4; ++sidx) {
        const float* brow = Bs + (16 * h + 4 * lg + sidx) * LDB + lr;
#pragma unroll
        for (int j = 0; j < LK_NT; ++j) {
          if (j >= NT) break;
          const float b = brow[16 * j];
#pragma unroll
          for (int t = 0; t < 2; ++t)
            acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[t][h][sidx], b, acc[t][j], 0, 0, 0);
        }
      }
    }
    // accumulator element i of lane l: row 16 t + 4 lg + i, column 16 j + lr
    const int m0 = g * LK_RG;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int row0 = m0 + 16 * t + 4 * lg;
      if (row0 >= M) continue;
#pragma unroll
      for (int j = 0; j < LK_NT; ++j) {
        const int col = 16 * j + lr;
        if (j >= NT || col >= N) continue;
        if (p.offM) {
          const int on = p.offN[col];
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (row0 + i < M) {
              float* d = C + p.offM[row0 + i] + on;
              const float o = alpha * acc[t][j][i];
              *d = beta != 0.f ? o + beta * *d : o;
            }
          continue;
        }
        float* dst = C + row0 + (int64_t)col * ldc;
        if (vecC && row0 + 3 < M) {
          float4 v = make_float4(alpha * acc[t][j][0], alpha * acc[t][j][1], alpha * acc[t][j][2], alpha * acc[t][j][3]);
          if (beta != 0.f) {
            const float4 o = *reinterpret_cast<const float4*>(dst);
            v.x += beta * o.x; v.y += beta * o.y; v.z += beta * o.z; v.w += beta * o.w;
          }
          *reinterpret_cast<float4*>(dst) = v;
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (row0 + i < M) {
              const float o = alpha * acc[t][j][i];
              dst[i] = beta != 0.f ? o + beta * dst[i] : o;
            }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int h = 0; h < LK_KC; ++h)
#pragma unroll
        for (int sidx = 0; sidx < 4; ++sidx) a[t][h][sidx] = an[t][h][sidx];
  }
}

static bool lowk_eligible(const GemmArgs& g) {
  return g.dt == DType::F32 && g.batch == 1 && g.K <= LK_KMAX && g.N <= LK_NMAX && g.M >= 4096 &&
         g.M < ((int64_t)1 << 31) && g.lda < ((int64_t)1 << 31) && g.ldb < ((int64_t)1 << 31) &&
         g.ldc < ((int64_t)1 << 31);
}

static void gemm_lowk(Ctx& c, const GemmArgs& g) {
  GemmP p{};
  p.M = g.M; p.N = g.N; p.K = g.K; p.alpha = g.alpha; p.beta = g.beta;
  p.A = g.A; p.lda = g.lda; p.B = g.B; p.ldb = g.ldb; p.C = g.C; p.ldc = g.ldc;
  p.offM = g.offM; p.offN = g.offN;
  const int64_t ngroups = (g.M + LK_RG - 1) / LK_RG;
  const int lds = lowk_lds_floats((int)g.K, (int)g.N) * 4;
  // persistent blocks, LK_TPB / 64 row groups each per round: up to two per CU
  const int per_cu = std::max(1, std::min(2, (160 * 1024) / lds));
  const unsigned grid =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((ngroups + LK_TPB / 64 - 1) / (LK_TPB / 64), 256 * per_cu));
  std::string name = "gemm_lowk_f32_kernel";
  if (c.timing_detail)
    name += " M=" + std::to_string(g.M) + " N=" + std::to_string(g.N) + " K=" + std::to_string(g.K);
  TimedLaunch tl(c, name, 2.0 * (double)g.M * g.N * g.K,
                 ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N * (g.beta != 0.0 ? 2 : 1)) * 4.0);
  if (g.transA)
    hipLaunchKernelGGL(gemm_lowk_f32_kernel<1>, dim3(grid), dim3(LK_TPB), lds, c.stream, p, (int)g.transB);
  else
    hipLaunchKernelGGL(gemm_lowk_f32_kernel<0>, dim3(grid), dim3(LK_TPB), lds, c.stream, p, (int)g.transB);
  HIP_CHECK(hipGetLastError());
}

// A launch plan: the FP32 block tile (by its column count; 64 x 192 tiles for
// tn = 192, else 128 x tn) and the number of split-K slices.
struct Plan {
  int tn = 0;
  int64_t S = 1;
};

static bool tiled_f32(const GemmArgs& g) { return g.dt == DType::F32; }
static int64_t tile_m(const GemmArgs& g, const Plan& pl) { return tiled_f32(g) ? (pl.tn == 192 ? 64 : 128) : BM; }
static int64_t tile_n(const GemmArgs& g, const Plan& pl) { return tiled_f32(g) ? pl.tn : BN; }
static int64_t out_tiles(const GemmArgs& g, const Plan& pl) {
  return ((g.M + tile_m(g, pl) - 1) / tile_m(g, pl)) * ((g.N + tile_n(g, pl) - 1) / tile_n(g, pl)) * g.batch;
}
// Split-K applies to one long-K product (batch 1, K >= 1024) whose output
// tiles cannot fill the 256 CUs; slices are at least kmin long.
static bool splitk_allowed(const GemmArgs& g, int64_t tiles) { return g.batch == 1 && tiles < 600 && g.K >= 1024; }
static int64_t splitk_kmin(const GemmArgs& g) { return tiled_f32(g) ? 256 : 64; }
// blocks the split-K rule aims at: the FP32-MFMA kernel's 128-row tiles at two
// per CU; the split-bf16 kernel ~1024; FP64 / C128 (64 x 64 tiles) ~2048
static int64_t splitk_target(const GemmArgs& g) { return real_f32(g) ? 512 : split_f32(g) ? 1024 : 2048; }
// A stored plan is applied only where it is valid for the shape: FP32 tiles of
// 64 / 128 / 192 columns, and split-K (S >= 2) only where splitk_allowed holds
// (execute's grid.z indexes the K chunks then, so a batched product would
// lose every batch but the first).
static bool plan_valid(const GemmArgs& g, const Plan& pl) {
  if (split_f32(g) && !(pl.tn == 64 || pl.tn == 128 || pl.tn == 192)) return false;
  if (real_f32(g) && !(pl.tn == 64 || pl.tn == 96 || pl.tn == 128 || pl.tn == 192)) return false;
  return pl.S == 1 || (pl.S >= 2 && splitk_allowed(g, out_tiles(g, pl)));
}

// Shape rules: tile by the output column count; split-K to about 1024 blocks
// (FP32, chunks >= 8 K tiles) or 2048 blocks (FP64 / C128, latency-bound per
// K step, chunks >= 64).
static Plan rule_plan(const GemmArgs& g) {
  Plan pl;
  pl.tn = split_f32(g) ? f32_tile_n(g.N) : real_f32(g) ? f32x_tile_n(g.N) : BN;
  const int64_t tiles = out_tiles(g, pl);
  if (splitk_allowed(g, tiles)) {
    const int64_t target = splitk_target(g);
    pl.S = std::max<int64_t>(1, std::min<int64_t>((target + tiles - 1) / tiles, g.K / splitk_kmin(g)));
  }
  return pl;
}

// Runs op(A) op(B) with plan `pl` (timed: bracketed for the kernel statistics).
static void execute(Ctx& c, const GemmArgs& g, const Plan& pl, bool timed) {
  GemmP p;
  p.M = g.M; p.N = g.N; p.K = g.K; p.alpha = g.alpha; p.beta = g.beta;
  p.A = g.A; p.lda = g.lda; p.sA = g.sA;
  p.B = g.B; p.ldb = g.ldb; p.sB = g.sB;
  p.C = g.C; p.ldc = g.ldc; p.sC = g.sC;
  p.conjA = g.conjA; p.conjB = g.conjB;
  p.kchunk = 0;
  p.offM = g.offM;
  p.offN = g.offN;
  const int64_t tbm = tile_m(g, pl), tbn = tile_n(g, pl);
  set_vec_flags(g, p);
  p.planeA = p.planeB = 0;
  dim3 grid((unsigned)((g.M + tbm - 1) / tbm), (unsigned)((g.N + tbn - 1) / tbn), (unsigned)g.batch);
  IPEPS_CHECK(grid.y <= 65535, "gemm: N too large for the grid");
  // FP32 operand modes: an operand that at least 3 blocks along the other
  // output dimension re-read over a long K (>= 512: the staging conversion
  // then outweighs the extra pass) is split into bf16 planes once
  // (split_operand) instead of in every block's staging.
  int ma = g.transA ? 1 : 0, mb = g.transB ? 0 : 1;
  std::shared_ptr<Buf> splitA, splitB;
  if (split_f32(g) && g.batch == 1 && g.K >= 512) {
    const int ns = nsplit_of(g.mixed);
    const int64_t Kp = (g.K + 31) / 32 * 32;
    const bool fits = Kp <= (int64_t)64 * 65535;
    if (grid.y >= 3 && fits && g.M * Kp < ((int64_t)1 << 31)) {
      int64_t kp;
      splitA = split_operand(c, (const float*)g.A, g.lda, g.transA, g.M, g.K, ns, &kp);
      p.A = splitA->p;
      p.lda = kp;
      p.planeA = g.M * kp;
      ma = 2;
    }
    if (grid.x >= 3 && fits && g.N * Kp < ((int64_t)1 << 31)) {
      int64_t kp;
      splitB = split_operand(c, (const float*)g.B, g.ldb, !g.transB, g.N, g.K, ns, &kp);
      p.B = splitB->p;
      p.ldb = kp;
      p.planeB = g.N * kp;
      mb = 2;
    }
  }
  const bool at = g.transA, bt = g.transB;
  const char* tf[2] = {"false", "true"};
  std::string name;
  double flops = 2.0 * (double)g.M * (double)g.N * (double)g.K * (double)g.batch;
  if (g.dt == DType::F64) name = std::string("gemm_f64_kernel<") + tf[at] + ", " + tf[bt] + ">";
  else if (g.dt == DType::C128) { name = std::string("gemm_c128_kernel<") + tf[at] + ", " + tf[bt] + ">"; flops *= 4; }
  else if (real_f32(g)) name = std::string("gemm_f32mfma_kernel<") + tf[at] + ", " + tf[bt] + ">";
  // FP32: the kernel's template head <operand mode A, operand mode B, bf16 planes>
  else name = "gemm_f32_kernel<" + std::to_string(ma) + ", " + std::to_string(mb) + ", " + std::to_string(nsplit_of(g.mixed)) + ">";
  const double bytes = ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N * (g.beta != 0.0 ? 2 : 1)) *
                       (double)dtype_size(g.dt) * (double)g.batch;
  if (c.timing_detail)
    name += " M=" + std::to_string(g.M) + " N=" + std::to_string(g.N) + " K=" + std::to_string(g.K) +
            " b=" + std::to_string(g.batch);
  if (!timed) name.clear();
  // Split-K: each z-slice of the grid reduces one K chunk (a multiple of 64)
  // into a workspace, then a reduction kernel applies alpha / beta.
  int64_t S = 1, kc = 0;
  if (pl.S >= 2) {
    kc = ((g.K + pl.S - 1) / pl.S + 63) / 64 * 64;
    S = (g.K + kc - 1) / kc;
    IPEPS_CHECK(S <= 65535, "gemm: too many split-K slices");
  }
  if (S >= 2) {
    const size_t es = dtype_size(g.dt);
    const size_t wbytes = (size_t)S * g.M * g.N * es;
    std::shared_ptr<Buf> ws(new Buf{c.pool->alloc(wbytes), wbytes, c.pool.get()});
    GemmP q = p;
    q.alpha = 1.0;
    q.beta = 0.0;
    q.offM = q.offN = nullptr;  // the reduction applies the output offsets
    q.C = ws->p;
    q.ldc = g.M;
    q.sC = g.M * g.N;
    q.kchunk = kc;
    // element offset of K chunk 1 (k-contiguous and pre-split operands: kc)
    q.sA = (split_f32(g) ? ma != 0 : at) ? kc : kc * g.lda;
    q.sB = (split_f32(g) ? mb != 0 : !bt) ? kc : kc * g.ldb;
    set_vec_flags(g, q);
    grid.z = (unsigned)S;
    {
      TimedLaunch tl(c, name, flops, bytes);
      launch(c, g, q, grid, ma, mb, pl.tn);
    }
    const int64_t MN = g.M * g.N;
    const char* rname = g.dt == DType::F64 ? "splitk_reduce<double>" : g.dt == DType::C128 ? "splitk_reduce<double2>"
                                                                                          : "splitk_reduce<float>";
    TimedLaunch tr(c, timed ? rname : "", 0.0, (double)(S + 1) * MN * es);
    IPEPS_CHECK(MN < ((int64_t)1 << 32), "gemm: split-K output too large");
    const unsigned rb = (unsigned)((MN + 255) / 256);
    std::shared_ptr<Buf> ws2;
    if (S > 32) {
      const int64_t F = 16, G = (S + F - 1) / F;
      const size_t w2bytes = (size_t)G * MN * es;
      ws2.reset(new Buf{c.pool->alloc(w2bytes), w2bytes, c.pool.get()});
      dim3 fgrid((unsigned)((MN + 255) / 256), (unsigned)G);
      if (g.dt == DType::F64)
        hipLaunchKernelGGL(splitk_fold<double>, fgrid, dim3(256), 0, c.stream, MN, S, F, (const double*)ws->p,
                           (double*)ws2->p);
      else if (g.dt == DType::C128)
        hipLaunchKernelGGL(splitk_fold<double2>, fgrid, dim3(256), 0, c.stream, MN, S, F, (const double2*)ws->p,
                           (double2*)ws2->p);
      else
        hipLaunchKernelGGL(splitk_fold<float>, fgrid, dim3(256), 0, c.stream, MN, S, F, (const float*)ws->p,
                           (float*)ws2->p);
      HIP_CHECK(hipGetLastError());
      ws = ws2;
      S = G;
    }
    if (g.dt == DType::F64)
      hipLaunchKernelGGL(splitk_reduce<double>, dim3(rb), dim3(256), 0, c.stream, (unsigned)g.M, MN, (int)S,
                         (const double*)ws->p, g.alpha, g.beta, (double*)g.C, g.ldc, g.offM, g.offN);
    else if (g.dt == DType::C128)
      hipLaunchKernelGGL(splitk_reduce<double2>, dim3(rb), dim3(256), 0, c.stream, (unsigned)g.M, MN, (int)S,
                         (const double2*)ws->p, g.alpha, g.beta, (double2*)g.C, g.ldc, g.offM, g.offN);
    else
      hipLaunchKernelGGL(splitk_reduce<float>, dim3(rb), dim3(256), 0, c.stream, (unsigned)g.M, MN, (int)S,
                         (const float*)ws->p, g.alpha, g.beta, (float*)g.C, g.ldc, g.offM, g.offN);
    HIP_CHECK(hipGetLastError());
    return;
  }
  TimedLaunch tl(c, name, flops, bytes);
  launch(c, g, p, grid, ma, mb, pl.tn);
}

// ---------------------------------------------------------------------------
// Autotuner: with tuning on, the first product of each shape (at least
// `min_flops`) times every candidate plan (FP32 tiles 64 / 128 / 192 columns
// x split-K slices {rule / 2, rule, 2 rule}, and 1 where the rule asks for at
// most 4; FP64 / C128 the slices) on the
// caller's stream into a scratch output and keeps the fastest.  The table is
// process-wide (both lanes of a context, every context) and can be saved and
// loaded, so a run can replay a table and reproduce its arithmetic exactly.
// ---------------------------------------------------------------------------
namespace {
struct Tuner {
  std::mutex mu;
  int mode = 0;  // 0 off (shape rules), 1 tune unseen shapes, 2 table only
  double min_flops = 1e8;
  std::map<std::string, Plan> table;
  long tuned = 0, hits = 0;
};
Tuner& tuner() {
  static Tuner t;
  return t;
}

std::string tune_key(const GemmArgs& g) {
  GemmP p;
  p.A = g.A; p.lda = g.lda; p.sA = g.sA; p.B = g.B; p.ldb = g.ldb; p.sB = g.sB;
  p.C = g.C; p.ldc = g.ldc; p.sC = g.sC; p.kchunk = 0;
  set_vec_flags(g, p);
  char b[200];
  snprintf(b, sizeof b, "%d.%d.%d%d.%lld.%lld.%lld.%lld.%d.%d%d%d", (int)g.dt, (int)g.mixed, (int)g.transA,
           (int)g.transB, (long long)g.M, (long long)g.N, (long long)g.K, (long long)g.batch, (int)(g.beta != 0.0),
           (int)p.vecA, (int)p.vecB, (int)p.vecC);
  return b;
}

std::vector<Plan> candidates(const GemmArgs& g) {
  std::vector<int> tns;
  if (split_f32(g)) {
    tns = {64, 128};
    if (g.N > 128) tns.push_back(192);
  } else if (real_f32(g)) {
    tns = {64, 96, 128};
    if (g.N > 128) tns.push_back(192);
  } else {
    tns = {(int)BN};
  }
  std::vector<Plan> out;
  for (int tn : tns) {
    Plan pl;
    pl.tn = tn;
    const int64_t tiles = out_tiles(g, pl);
    // slices around the rule's count; no split only where the rule asks for
    // at most 4 slices (an unsplit long-K product on a handful of blocks can
    // run 100x longer than the rule's plan, and timing it costs as much)
    std::vector<int64_t> ss;
    if (splitk_allowed(g, tiles)) {
      const int64_t smax = g.K / splitk_kmin(g);
      const int64_t target = splitk_target(g);
      const int64_t r = std::max<int64_t>(1, std::min<int64_t>((target + tiles - 1) / tiles, smax));
      if (r <= 4) ss.push_back(1);
      for (int64_t s : {r / 2, r, 2 * r})
        if (s >= 2 && s <= smax && std::find(ss.begin(), ss.end(), s) == ss.end()) ss.push_back(s);
      if (ss.empty()) ss.push_back(1);
    } else {
      ss.push_back(1);
    }
    for (int64_t s : ss) {
      pl.S = s;
      out.push_back(pl);
    }
  }
  return out;
}

Plan tune(Ctx& c, const GemmArgs& g) {
  const std::vector<Plan> cand = candidates(g);
  if (cand.size() == 1) return cand[0];
  // scratch output with the caller's strides (same vector-access flags)
  const size_t es = dtype_size(g.dt);
  const int64_t span = (g.batch - 1) * g.sC + (g.N - 1) * g.ldc + g.M;
  const size_t bytes = (size_t)span * es;
  Buf scratch{c.pool->alloc(bytes), bytes, c.pool.get()};
  GemmArgs t = g;
  t.C = scratch.p;
  // the candidates' launches stay out of the kernel statistics
  struct TimingOff {
    Ctx& c;
    bool was;
    explicit TimingOff(Ctx& cc) : c(cc), was(cc.timing) { c.timing = false; }
    ~TimingOff() { c.timing = was; }
  } off(c);
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  Plan best = cand[0];
  float best_ms = 1e30f;
  for (const Plan& pl : cand) {
    execute(c, t, pl, false);  // warm-up (also the first touch of the pool blocks)
    float ms = 1e30f;
    for (int r = 0; r < 2; ++r) {
      HIP_CHECK(hipEventRecord(e0, c.stream));
      execute(c, t, pl, false);
      HIP_CHECK(hipEventRecord(e1, c.stream));
      HIP_CHECK(hipEventSynchronize(e1));
      float m = 0;
      HIP_CHECK(hipEventElapsedTime(&m, e0, e1));
      ms = std::min(ms, m);
    }
    if (ms < best_ms) {
      best_ms = ms;
      best = pl;
    }
  }
  HIP_CHECK(hipEventDestroy(e0));
  HIP_CHECK(hipEventDestroy(e1));
  return best;
}
}  // namespace

void tune_config(int mode, double min_flops) {
  IPEPS_CHECK(mode >= 0 && mode <= 2, "tune mode must be 0, 1 or 2");
  std::lock_guard<std::mutex> lk(tuner().mu);
  tuner().mode = mode;
  tuner().min_flops = min_flops;
}

void tune_stats(long out[3]) {
  std::lock_guard<std::mutex> lk(tuner().mu);
  out[0] = (long)tuner().table.size();
  out[1] = tuner().tuned;
  out[2] = tuner().hits;
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(tuner().mu);
  tuner().table.clear();
  tuner().tuned = tuner().hits = 0;
}

void tune_save(const std::string& path) {
  std::lock_guard<std::mutex> lk(tuner().mu);
  FILE* f = fopen(path.c_str(), "w");
  IPEPS_CHECK(f, "tune_save: cannot open " + path);
  fprintf(f, "# ipeps gemm tuning table: key tile_n splitk\n");
  for (auto& kv : tuner().table) fprintf(f, "%s %d %lld\n", kv.first.c_str(), kv.second.tn, (long long)kv.second.S);
  fclose(f);
}

void tune_load(const std::string& path) {
  FILE* f = fopen(path.c_str(), "r");
  IPEPS_CHECK(f, "tune_load: cannot open " + path);
  std::map<std::string, Plan> t;
  char line[512];
  int ln = 0;
  while (fgets(line, sizeof line, f)) {
    ++ln;
    if (line[0] == '#' || line[0] == '\n') continue;
    char key[256];
    int tn;
    long long S;
    if (sscanf(line, "%255s %d %lld", key, &tn, &S) != 3 || !(tn == 64 || tn == 96 || tn == 128 || tn == 192) || S < 1 ||
        S > 65535) {
      fclose(f);
      throw Error(IPEPS_ERR_ARG, "tune_load: malformed line " + std::to_string(ln) + " of " + path);
    }
    // key = dtype.mixed.tAtB.M.N.K.batch.beta.vec (tune_key): split-K plans
    // exist for batch-1 products with K >= 1024 only
    int kdt, kmx, ktt, kbeta;
    long long kM, kN, kK, kbatch;
    if (sscanf(key, "%d.%d.%d.%lld.%lld.%lld.%lld.%d.", &kdt, &kmx, &ktt, &kM, &kN, &kK, &kbatch, &kbeta) != 8 ||
        kM < 1 || kN < 1 || kK < 1 || kbatch < 1) {
      fclose(f);
      throw Error(IPEPS_ERR_ARG, "tune_load: malformed key on line " + std::to_string(ln) + " of " + path);
    }
    // column tiles: FP64 / C128 kernels 64; split-bf16 FP32 64 / 128 / 192;
    // FP32-MFMA also 96
    const bool tile_ok = kdt == (int)DType::F32 ? (tn != 96 || kmx == (int)Mixed::BF16X6) : tn == (int)BN;
    if (!tile_ok) {
      fclose(f);
      throw Error(IPEPS_ERR_ARG, "tune_load: line " + std::to_string(ln) + " of " + path + " gives tile " +
                                     std::to_string(tn) + " to a kernel without it");
    }
    if (S >= 2 && (kbatch != 1 || kK < 1024)) {
      fclose(f);
      throw Error(IPEPS_ERR_ARG, "tune_load: line " + std::to_string(ln) + " of " + path +
                                     " gives split-K slices to a product with batch " + std::to_string(kbatch) +
                                     " and K " + std::to_string(kK) + " (split-K needs batch 1, K >= 1024)");
    }
    Plan pl;
    pl.tn = tn;
    pl.S = S;
    t[key] = pl;
  }
  fclose(f);
  std::lock_guard<std::mutex> lk(tuner().mu);
  for (auto& kv : t) tuner().table[kv.first] = kv.second;
}

namespace {
struct MixedRadixP {
  int nd;
  int64_t ext[16], stride[16];
};
__global__ __launch_bounds__(256) void mixed_radix_offsets_kernel(MixedRadixP p, int64_t count, int* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (int64_t)gridDim.x * 256) {
    int64_t r = i, off = 0;
    for (int a = 0; a < p.nd; ++a) {
      const int64_t q = r / p.ext[a];
      off += (r - q * p.ext[a]) * p.stride[a];
      r = q;
    }
    out[i] = (int)off;
  }
}
}  // namespace

void mixed_radix_offsets(Ctx& c, int nd, const int64_t* ext, const int64_t* stride, int64_t count, int* out) {
  IPEPS_CHECK(nd >= 0 && nd <= 16, "mixed_radix_offsets: at most 16 axes");
  if (count <= 0) return;
  MixedRadixP p{};
  p.nd = nd;
  for (int a = 0; a < nd; ++a) {
    p.ext[a] = ext[a];
    p.stride[a] = stride[a];
  }
  const unsigned grid = (unsigned)std::min<int64_t>((count + 255) / 256, 65535);
  TimedLaunch tl(c, "mixed_radix_offsets_kernel", 0.0, 4.0 * (double)count);
  hipLaunchKernelGGL(mixed_radix_offsets_kernel, dim3(grid), dim3(256), 0, c.stream, p, count, out);
  HIP_CHECK(hipGetLastError());
}

void gemm(Ctx& c, const GemmArgs& g) {
  if (g.M <= 0 || g.N <= 0 || g.batch <= 0) return;
  IPEPS_CHECK(!g.offM || ((g.dt == DType::F32 || g.dt == DType::F64) && g.batch == 1 && g.K > 0 && g.offN),
              "gemm: output offsets need F32 / F64, batch 1 and K > 0");
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
  if (lowk_eligible(g)) {
    gemm_lowk(c, g);
    return;
  }
  Plan pl = rule_plan(g);
  Tuner& T = tuner();
  int mode;
  double min_flops;
  {
    std::lock_guard<std::mutex> lk(T.mu);
    mode = T.mode;
    min_flops = T.min_flops;
  }
  if (mode != 0) {
    const std::string key = tune_key(g);
    bool found = false;
    {
      std::lock_guard<std::mutex> lk(T.mu);
      auto it = T.table.find(key);
      if (it == T.table.end() && g.beta != 0.0) {
        // an accumulating product (beta != 0) takes the plan of the same
        // product without accumulation when only that one is in the table
        GemmArgs g0 = g;
        g0.beta = 0.0;
        it = T.table.find(tune_key(g0));
      }
      if (it != T.table.end() && plan_valid(g, it->second)) {
        pl = it->second;
        found = true;
        ++T.hits;
      }
    }
    const double fl = 2.0 * (double)g.M * (double)g.N * (double)g.K * (double)g.batch;
    // (never while a graph is captured: tuning times the candidates)
    if (!found && mode == 1 && fl >= min_flops && !c.graph_safe) {
      Plan best = tune(c, g);
      std::lock_guard<std::mutex> lk(T.mu);
      // two lanes that tuned one shape together keep the first entry
      auto ins = T.table.emplace(key, best);
      pl = ins.first->second;
      if (ins.second) ++T.tuned;
    }
  }
  execute(c, g, pl, true);
}

}  // namespace k
}  // namespace ipeps
