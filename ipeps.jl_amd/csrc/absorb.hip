// Fused absorption of a site into an enlarged corner (FP32 data, FP32 MFMA):
//
//   Y[r, a, b] = sum_{u, v, p} X[r, u, v] A[u, a, p] Ad[v, b, p]
//
// r runs over the two row legs of X (the environment legs), u / v over the
// two ket / bra legs that meet A / Ad (u, v, a, b < U = D^2), p over the
// physical leg.  Per row this is Y_r = sum_p A_p^T X_r Ad_p: T_p = X_r Ad_p
// is accumulated by v_mfma_f32_16x16x4f32 and fed, from the accumulator
// registers, as the A operand of the second product Y_r^T += T_p^T A_p.  The
// accumulator of a 16 x 16 tile holds D[4 (l/16) + i][l % 16] in lane l,
// element i; taking the k index of the second product's sub-step s as
// u = 16 tu + 4 (l/16) + s makes element s of lane l exactly the operand lane
// l must supply, so the intermediate never leaves registers.
//
// The same kernel gives the cotangent of X (A and Ad enter transposed), and a
// second kernel the cotangents of A and Ad:
//   Abar[u, a, p]  = sum_r (X_r Ad_p)     Ybar_r^T   [u, a]
//   Adbar[v, b, p] = sum_r (X_r^T A_p)    Ybar_r     [v, b]
// accumulated per wave over the rows a persistent block visits, then summed
// over waves (LDS) and blocks (a reduction kernel): no atomics.
//
// X and Y are addressed through strides: X[r1, r2, u1, u2, v1, v2] with
// u = u1 + D u2 and v = v1 + D v2, each leg with its own stride; rows of a
// block are consecutive r1 (the stride-1 leg when there is one) for one r2.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ipeps {
namespace k {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int ABS_UMAX = 36;  // U = D^2 <= 36 (D <= 6)
constexpr int ABS_PMAX = 2;   // physical extent
constexpr int ABS_RB = 8;     // rows per chunk (forward)
constexpr int ABS_RBW = 8;    // rows per chunk (weight gradient)
// Waves per block (one block per CU, LDS-bound): the forward kernel runs two
// per SIMD (one wave's LDS waits overlap the other's MFMAs; 238 VGPRs at D=6),
// the weight gradient one per SIMD (its two accumulator sets need ~250 VGPRs
// with the prefetch of two row chunks; at two waves per SIMD it would spill).
constexpr int ABS_NW = 8;
constexpr int ABS_THREADS = 64 * ABS_NW;
constexpr int ABS_NWW = 4;
constexpr int ABS_THREADS_W = 64 * ABS_NWW;

__device__ inline v4f mfma4(float a, float b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Element offsets fit in 32 bits (the planner admits tensors of < 2^31
// elements), and D is a compile-time constant, so the index math is a few
// multiply-shifts.
template <int D>
__device__ inline int off2(int i, int s1, int s2) { return (i % D) * s1 + (i / D) * s2; }

struct Legs32 {
  int n1, n2, s1, s2, sx1, sx2, sy1, sy2;
};
__device__ inline Legs32 legs32(const AbsorbLegs& g) {
  return Legs32{g.n1, g.n2, (int)g.s1, (int)g.s2, (int)g.sx1, (int)g.sx2, (int)g.sy1, (int)g.sy2};
}

// A / Ad gathered into LDS as [u + U (a + U p)] from their strided layout.
template <int D>
__device__ inline void gather_site(const float* __restrict__ src, const int (&s)[5], float* dst, int P, int nth) {
  constexpr int U = D * D;
  for (int e = threadIdx.x; e < U * U * P; e += nth) {
    const int u = e % U, r = e / U, a = r % U, pp = r / U;
    dst[e] = src[(u % D) * s[0] + (u / D) * s[1] + (a % D) * s[2] + (a / D) * s[3] + pp * s[4]];
  }
}

// RB rows (r1 = r10 .. r10+RB-1 of one r2) of a strided (r, x, y) tensor,
// held in registers between the fetch and the LDS store dst[rr][x][y] (row
// stride U * (U + 1)).  Element e of the chunk is (rr = e % RB, x, y); lanes
// with consecutive e read consecutive rows (the stride-1 leg), and every load
// is unconditional: rows past n1 read row n1 - 1 (their results are never
// written) and the elements past the chunk end re-read its last element.
template <int D, int RB, int NTH>
struct RowChunk {
  static constexpr int U = D * D, LD = U + 1, TOT = U * U * RB;
  static constexpr int NE = (TOT + NTH - 1) / NTH;
  float v[NE];
  __device__ inline void fetch(const float* __restrict__ src, const Legs32& g, int r10, int r2) {
    const float* base = src + r2 * g.s2;
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = min((int)threadIdx.x + j * NTH, TOT - 1);
      const int rr = e % RB, xy = e / RB;
      const int r1 = min(r10 + rr, g.n1 - 1);
      v[j] = base[r1 * g.s1 + off2<D>(xy % U, g.sx1, g.sx2) + off2<D>(xy / U, g.sy1, g.sy2)];
    }
  }
  __device__ inline void store(float* __restrict__ dst) const {
#pragma unroll
    for (int j = 0; j < NE; ++j) {
      const int e = (int)threadIdx.x + j * NTH;
      if (e < TOT) {
        const int rr = e % RB, xy = e / RB;
        dst[rr * U * LD + (xy % U) * LD + xy / U] = v[j];
      }
    }
  }
};

// Chunk schedule of the persistent kernels (gridDim.x a multiple of 32).
// Chunks 4g .. 4g+3 hold consecutive rows (32 B each), i.e. the four
// quarters of the same 128-B lines.  They go to four blocks on the same XCD
// (blocks are dealt to the 8 XCDs round-robin) that step through their line
// groups in lockstep, so one line fill in that XCD's L2 serves all four.
struct ChunkSched {
  int j, g, gstep;  // member of the quad, first line group, group stride
  __device__ ChunkSched() {
    const int x = blockIdx.x % 8, y = blockIdx.x / 8;
    j = y % 4;
    g = x + 8 * (y / 4);
    gstep = gridDim.x / 4;
  }
  __device__ int chunk(int t) const { return 4 * (g + t * gstep) + j; }
};

// Persistent blocks walk the row chunks (RB rows of one r2) with stride
// gridDim.x; the next chunk's loads are in flight while the MFMAs of the
// current one run.
template <int D>
__global__ __launch_bounds__(ABS_THREADS) void absorb_fwd_kernel(AbsorbP p) {
  constexpr int U = D * D, LD = U + 1, NT = (U + 15) / 16;
  constexpr int YLD = ABS_RB + 1;  // staged output ys[(a + U b) YLD + row]
  __shared__ float xs[ABS_RB * U * LD];
  __shared__ float ys[U * U * YLD];
  __shared__ float am[U * U * ABS_PMAX], adm[U * U * ABS_PMAX];
  const int P = p.P;
  const Legs32 gx = legs32(p.x), gy = legs32(p.y);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int nb1 = (gx.n1 + ABS_RB - 1) / ABS_RB, nchunk = nb1 * gx.n2;
  gather_site<D>(p.Am, p.sa, am, P, ABS_THREADS);
  gather_site<D>(p.Adm, p.sd, adm, P, ABS_THREADS);
  const ChunkSched sch;
  RowChunk<D, ABS_RB, ABS_THREADS> rc;
  int ch = sch.chunk(0);
  if (ch < nchunk) rc.fetch(p.X, gx, (ch % nb1) * ABS_RB, ch / nb1);
  for (int t = 0; ch < nchunk; ch = sch.chunk(++t)) {
    const int r2 = ch / nb1, r10 = (ch % nb1) * ABS_RB;
    __syncthreads();  // the previous chunk's rows are consumed and its output written
    rc.store(xs);
    __syncthreads();
    const int nx = sch.chunk(t + 1);
    if (nx < nchunk) rc.fetch(p.X, gx, (nx % nb1) * ABS_RB, nx / nb1);
    for (int rr = wave; rr < ABS_RB; rr += ABS_NW) {
      const int r1 = r10 + rr;
      if (r1 >= gx.n1) break;
      const float* xr = xs + rr * U * LD;
      v4f yacc[NT][NT];  // [tb][ta]: Y^T[b][a]
#pragma unroll
      for (int i = 0; i < NT; ++i)
#pragma unroll
        for (int j = 0; j < NT; ++j) yacc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      for (int pp = 0; pp < P; ++pp) {
        v4f tacc[NT][NT];  // [tu][tb]: T[u][b] = (X_r Ad_p)[u][b]
#pragma unroll
        for (int i = 0; i < NT; ++i)
#pragma unroll
          for (int j = 0; j < NT; ++j) tacc[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < (U + 3) / 4; ++ks) {
          const int v = 4 * ks + lg;
          float ao[NT], bo[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int u = 16 * t + lr, b = 16 * t + lr;
            ao[t] = (u < U && v < U) ? xr[u * LD + v] : 0.f;
            bo[t] = (v < U && b < U) ? adm[v + U * (b + U * pp)] : 0.f;
          }
#pragma unroll
          for (int tu = 0; tu < NT; ++tu)
#pragma unroll
            for (int tb = 0; tb < NT; ++tb) tacc[tu][tb] = mfma4(ao[tu], bo[tb], tacc[tu][tb]);
        }
#pragma unroll
        for (int tu = 0; tu < NT; ++tu)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int u = 16 * tu + 4 * lg + s;
            float bo[NT];
#pragma unroll
            for (int ta = 0; ta < NT; ++ta) {
              const int a = 16 * ta + lr;
              bo[ta] = (u < U && a < U) ? am[u + U * (a + U * pp)] : 0.f;
            }
#pragma unroll
            for (int tb = 0; tb < NT; ++tb)
#pragma unroll
              for (int ta = 0; ta < NT; ++ta) yacc[tb][ta] = mfma4(tacc[tu][tb][s], bo[ta], yacc[tb][ta]);
          }
      }
#pragma unroll
      for (int tb = 0; tb < NT; ++tb)
#pragma unroll
        for (int ta = 0; ta < NT; ++ta)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int b = 16 * tb + 4 * lg + i, a = 16 * ta + lr;
            if (a < U && b < U) ys[(a + U * b) * YLD + rr] = yacc[tb][ta][i];
          }
    }
    __syncthreads();
    // the chunk's output, consecutive threads along the rows (the stride-1 leg)
    const int nrow = min(ABS_RB, gx.n1 - r10);
    float* ybase = p.Y + r10 * gy.s1 + r2 * gy.s2;
    for (int e = threadIdx.x; e < ABS_RB * U * U; e += ABS_THREADS) {
      const int rr = e % ABS_RB, ab = e / ABS_RB;
      if (rr < nrow)
        ybase[rr * gy.s1 + off2<D>(ab % U, gy.sx1, gy.sx2) + off2<D>(ab / U, gy.sy1, gy.sy2)] = ys[ab * YLD + rr];
    }
  }
}

template <int D>
__global__ __launch_bounds__(ABS_THREADS_W) void absorb_wgrad_kernel(AbsorbWP p) {
  constexpr int U = D * D, LD = U + 1, NT = (U + 15) / 16;
  __shared__ float xs[ABS_RBW * U * LD];
  __shared__ float gs[ABS_RBW * U * LD];
  __shared__ float am[U * U * ABS_PMAX], adm[U * U * ABS_PMAX];
  const int P = p.P;
  const Legs32 gx = legs32(p.x), gg = legs32(p.g);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  gather_site<D>(p.Am, p.sa, am, P, ABS_THREADS_W);
  gather_site<D>(p.Adm, p.sd, adm, P, ABS_THREADS_W);
  v4f abar[ABS_PMAX][NT][NT];   // [p][ta][tu]: Abar^T[a][u]
  v4f adbar[ABS_PMAX][NT][NT];  // [p][tb][tv]: Adbar^T[b][v]
#pragma unroll
  for (int q = 0; q < ABS_PMAX; ++q)
#pragma unroll
    for (int i = 0; i < NT; ++i)
#pragma unroll
      for (int j = 0; j < NT; ++j) {
        abar[q][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
        adbar[q][i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
      }
  const int nb1 = (gx.n1 + ABS_RBW - 1) / ABS_RBW;
  const int nblk = nb1 * gx.n2;
  const ChunkSched sch;
  RowChunk<D, ABS_RBW, ABS_THREADS_W> cx, cg;
  int blk = sch.chunk(0);
  if (blk < nblk) {
    cx.fetch(p.X, gx, (blk % nb1) * ABS_RBW, blk / nb1);
    cg.fetch(p.G, gg, (blk % nb1) * ABS_RBW, blk / nb1);
  }
  for (int t = 0; blk < nblk; blk = sch.chunk(++t)) {
    const int r10 = (blk % nb1) * ABS_RBW;
    __syncthreads();  // previous rows fully consumed
    cx.store(xs);
    cg.store(gs);
    __syncthreads();
    const int nx = sch.chunk(t + 1);
    if (nx < nblk) {
      cx.fetch(p.X, gx, (nx % nb1) * ABS_RBW, nx / nb1);
      cg.fetch(p.G, gg, (nx % nb1) * ABS_RBW, nx / nb1);
    }
    for (int rr = wave; rr < ABS_RBW; rr += ABS_NWW) {
      if (r10 + rr >= gx.n1) break;
      const float* xr = xs + rr * U * LD;  // X_r[u][v]
      const float* gr = gs + rr * U * LD;  // G_r[a][b]
#pragma unroll
      for (int pp = 0; pp < ABS_PMAX; ++pp) {
        if (pp >= P) break;
        // Tt[b][u] = sum_v Ad_p[v][b] X_r[u][v]
        v4f tt[NT][NT];
#pragma unroll
        for (int i = 0; i < NT; ++i)
#pragma unroll
          for (int j = 0; j < NT; ++j) tt[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int ks = 0; ks < (U + 3) / 4; ++ks) {
          const int v = 4 * ks + lg;
          float ao[NT], bo[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int b = 16 * t + lr, u = 16 * t + lr;
            ao[t] = (v < U && b < U) ? adm[v + U * (b + U * pp)] : 0.f;
            bo[t] = (v < U && u < U) ? xr[u * LD + v] : 0.f;
          }
#pragma unroll
          for (int tb = 0; tb < NT; ++tb)
#pragma unroll
            for (int tu = 0; tu < NT; ++tu) tt[tb][tu] = mfma4(ao[tb], bo[tu], tt[tb][tu]);
        }
        // Abar^T[a][u] += sum_b G_r[a][b] Tt[b][u]
#pragma unroll
        for (int tb = 0; tb < NT; ++tb)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int b = 16 * tb + 4 * lg + s;
            float ao[NT];
#pragma unroll
            for (int ta = 0; ta < NT; ++ta) {
              const int a = 16 * ta + lr;
              ao[ta] = (a < U && b < U) ? gr[a * LD + b] : 0.f;
            }
#pragma unroll
            for (int ta = 0; ta < NT; ++ta)
#pragma unroll
              for (int tu = 0; tu < NT; ++tu) abar[pp][ta][tu] = mfma4(ao[ta], tt[tb][tu][s], abar[pp][ta][tu]);
          }
        // S[a][v] = sum_u A_p[u][a] X_r[u][v]
        v4f sa[NT][NT];
#pragma unroll
        for (int i = 0; i < NT; ++i)
#pragma unroll
          for (int j = 0; j < NT; ++j) sa[i][j] = (v4f){0.f, 0.f, 0.f, 0.f};
#pragma unroll 1
        for (int ks = 0; ks < (U + 3) / 4; ++ks) {
          const int u = 4 * ks + lg;
          float ao[NT], bo[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int a = 16 * t + lr, v = 16 * t + lr;
            ao[t] = (u < U && a < U) ? am[u + U * (a + U * pp)] : 0.f;
            bo[t] = (u < U && v < U) ? xr[u * LD + v] : 0.f;
          }
#pragma unroll
          for (int ta = 0; ta < NT; ++ta)
#pragma unroll
            for (int tv = 0; tv < NT; ++tv) sa[ta][tv] = mfma4(ao[ta], bo[tv], sa[ta][tv]);
        }
        // Adbar^T[b][v] += sum_a G_r[a][b] S[a][v]
#pragma unroll
        for (int ta = 0; ta < NT; ++ta)
#pragma unroll
          for (int s = 0; s < 4; ++s) {
            const int a = 16 * ta + 4 * lg + s;
            float ao[NT];
#pragma unroll
            for (int tb = 0; tb < NT; ++tb) {
              const int b = 16 * tb + lr;
              ao[tb] = (a < U && b < U) ? gr[a * LD + b] : 0.f;
            }
#pragma unroll
            for (int tb = 0; tb < NT; ++tb)
#pragma unroll
              for (int tv = 0; tv < NT; ++tv) adbar[pp][tb][tv] = mfma4(ao[tb], sa[ta][tv][s], adbar[pp][tb][tv]);
          }
      }
    }
  }
  // sum the waves in LDS, one physical index at a time (xs and gs each hold
  // RBW * U * (U + 1) >= ABS_NWW * U * U floats)
  static_assert(ABS_RBW * LD >= ABS_NWW * U, "wave reduction buffer");
  float* red = xs;
#pragma unroll
  for (int pp = 0; pp < ABS_PMAX; ++pp) {
    if (pp >= P) break;
    __syncthreads();
#pragma unroll
    for (int t1 = 0; t1 < NT; ++t1)
#pragma unroll
      for (int t2 = 0; t2 < NT; ++t2)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int m = 16 * t1 + 4 * lg + i, n = 16 * t2 + lr;  // (a, u) resp. (b, v)
          if (m < U && n < U) {
            red[wave * U * U + n + U * m] = abar[pp][t1][t2][i];    // Abar[u = n][a = m]
            gs[wave * U * U + n + U * m] = adbar[pp][t1][t2][i];    // Adbar[v = n][b = m]
          }
        }
    __syncthreads();
    float* outA = p.partial + (int64_t)blockIdx.x * 2 * U * U * P + (int64_t)pp * U * U;
    float* outD = outA + (int64_t)U * U * P;
    for (int e = threadIdx.x; e < U * U; e += ABS_THREADS_W) {
      float sa = 0.f, sd = 0.f;
#pragma unroll
      for (int w = 0; w < ABS_NWW; ++w) {
        sa += red[w * U * U + e];
        sd += gs[w * U * U + e];
      }
      outA[e] = sa;
      outD[e] = sd;
    }
  }
}

// out[y][e] = sum of in[b][e] over b in [y F, min(nblk, (y+1) F)), e < n;
// run twice (F = 16, then F = all) so that each thread adds at most 16 terms
// in the first pass and the blocks of the first pass cover (n / 256) x (nblk / 16)
__global__ __launch_bounds__(256) void absorb_reduce_kernel(int nblk, int n, int F, const float* __restrict__ in,
                                                            float* __restrict__ out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= n) return;
  const int b0 = blockIdx.y * F, b1 = min(nblk, b0 + F);
  float s = 0.f;
  for (int b = b0; b < b1; ++b) s += in[(int64_t)b * n + e];
  out[(int64_t)blockIdx.y * n + e] = s;
}

}  // namespace

bool absorb_supported(int D, int P) { return D >= 4 && D * D <= ABS_UMAX && P >= 1 && P <= ABS_PMAX; }

void absorb_fwd(Ctx& c, const AbsorbP& p) {
  IPEPS_CHECK(absorb_supported(p.D, p.P) && p.U == p.D * p.D, "absorb: unsupported D / d");
  const int nb1 = (p.x.n1 + ABS_RB - 1) / ABS_RB;
  const int64_t nb = (int64_t)nb1 * p.x.n2;
  IPEPS_CHECK(nb >= 1 && nb < (1ll << 31), "absorb: grid");
  // one block per CU (LDS-bound); the chunk schedule needs a multiple of 32
  const unsigned grid = 256;
  const double rows = (double)p.x.n1 * p.x.n2;
  TimedLaunch tl(c, "absorb_fwd_kernel", rows * 4.0 * p.U * p.U * p.U * p.P, rows * 8.0 * p.U * p.U);
  const dim3 blk(ABS_THREADS);
  if (p.D == 4) hipLaunchKernelGGL((absorb_fwd_kernel<4>), dim3(grid), blk, 0, c.stream, p);
  else if (p.D == 5) hipLaunchKernelGGL((absorb_fwd_kernel<5>), dim3(grid), blk, 0, c.stream, p);
  else hipLaunchKernelGGL((absorb_fwd_kernel<6>), dim3(grid), blk, 0, c.stream, p);
  HIP_CHECK(hipGetLastError());
}

namespace {
constexpr int ABS_WGRID = 256;  // weight-gradient blocks: one persistent block per CU (multiple of 32)
constexpr int ABS_RF = 16;      // partials summed per thread in the first reduction pass
constexpr int ABS_RG = (ABS_WGRID + ABS_RF - 1) / ABS_RF;

// Scratch of the allocation-free entry (absorb_wgrad_scratch): the per-block
// partial sums, the first reduction pass and the result for the largest
// supported site, (ABS_WGRID + ABS_RG + 1) x 2 U^2 P floats at U = 36, P = 2
// (5.6 MiB).  Each GPU
// the library runs on holds its own copy until the process ends; host code
// uses only the per-device address taken by absorb_scratch_init.
constexpr int64_t ABS_SCRATCH_FLOATS = (int64_t)(ABS_WGRID + ABS_RG + 1) * 2 * ABS_UMAX * ABS_UMAX * ABS_PMAX;
static __device__ __align__(256) float scratch[ABS_SCRATCH_FLOATS];

constexpr int ABS_MAX_DEVICES = 64;
std::mutex scratch_mu;
float* scratch_ptr[ABS_MAX_DEVICES] = {};

// the weight gradient with its partial sums in `partial` (ABS_WGRID x n
// floats) and `partial2` (ABS_RG x n floats)
void wgrad_into(Ctx& c, AbsorbWP p, float* partial, float* partial2, float* Abar_Adbar) {
  IPEPS_CHECK(absorb_supported(p.D, p.P) && p.U == p.D * p.D, "absorb: unsupported D / d");
  const int grid = ABS_WGRID;
  const int n = 2 * p.U * p.U * p.P;
  p.partial = partial;
  const double rows = (double)p.x.n1 * p.x.n2;
  {
    TimedLaunch tl(c, "absorb_wgrad_kernel", rows * 8.0 * p.U * p.U * p.U * p.P, rows * 8.0 * p.U * p.U);
    const dim3 blk(ABS_THREADS_W);
    if (p.D == 4) hipLaunchKernelGGL((absorb_wgrad_kernel<4>), dim3(grid), blk, 0, c.stream, p);
    else if (p.D == 5) hipLaunchKernelGGL((absorb_wgrad_kernel<5>), dim3(grid), blk, 0, c.stream, p);
    else hipLaunchKernelGGL((absorb_wgrad_kernel<6>), dim3(grid), blk, 0, c.stream, p);
    HIP_CHECK(hipGetLastError());
  }
  const int F = ABS_RF, G = ABS_RG;
  hipLaunchKernelGGL(absorb_reduce_kernel, dim3((n + 255) / 256, G), dim3(256), 0, c.stream, grid, n, F, p.partial,
                     partial2);
  hipLaunchKernelGGL(absorb_reduce_kernel, dim3((n + 255) / 256, 1), dim3(256), 0, c.stream, G, n, G,
                     (const float*)partial2, Abar_Adbar);
  HIP_CHECK(hipGetLastError());
}
}  // namespace

void absorb_wgrad(Ctx& c, AbsorbWP p, float* Abar_Adbar) {
  const int64_t n = 2 * (int64_t)p.U * p.U * p.P;
  std::shared_ptr<Buf> ws(new Buf{c.pool->alloc((size_t)ABS_WGRID * n * 4), (size_t)ABS_WGRID * n * 4, c.pool.get()});
  std::shared_ptr<Buf> ws2(new Buf{c.pool->alloc((size_t)ABS_RG * n * 4), (size_t)ABS_RG * n * 4, c.pool.get()});
  wgrad_into(c, p, (float*)ws->p, (float*)ws2->p, Abar_Adbar);
}

void absorb_scratch_init(int device) {
  // with `device` current: each GPU holds its own copy of `scratch`
  std::lock_guard<std::mutex> g(scratch_mu);
  IPEPS_CHECK(device >= 0 && device < ABS_MAX_DEVICES, "absorb scratch: device index out of range");
  if (scratch_ptr[device]) return;
  void* ptr = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  IPEPS_CHECK(ptr != nullptr, "absorb scratch: no device address");
  scratch_ptr[device] = (float*)ptr;
}

int64_t absorb_scratch_floats() { return ABS_SCRATCH_FLOATS; }

float* absorb_wgrad_scratch(Ctx& c, AbsorbWP p) {
  const int64_t n = 2 * (int64_t)p.U * p.U * p.P;
  IPEPS_CHECK(n > 0 && (ABS_WGRID + ABS_RG + 1) * n <= ABS_SCRATCH_FLOATS,
              "absorb (no allocation): the site needs " + std::to_string((ABS_WGRID + ABS_RG + 1) * n * 4) +
                  " scratch bytes, more than the " + std::to_string(ABS_SCRATCH_FLOATS * 4) + " available");
  float* base;
  {
    std::lock_guard<std::mutex> g(scratch_mu);
    IPEPS_CHECK(c.device >= 0 && c.device < ABS_MAX_DEVICES && scratch_ptr[c.device],
                "absorb (no allocation): scratch of this device not initialised");
    base = scratch_ptr[c.device];
  }
  float* out = base + (ABS_WGRID + ABS_RG) * n;
  wgrad_into(c, p, base, base + ABS_WGRID * n, out);
  return out;
}

}  // namespace k
}  // namespace ipeps
