// Index arithmetic of the permute kernels, shared by the device kernels and
// host-side checks.  A permutation is described on the output side after
// size-1 axes are dropped and axes that stay adjacent are merged: output axis
// a has extent od[a] and input stride is[a] (column-major, first axis fastest).
//
// Two kernels cover every permutation of fewer than 2^31 elements:
//  * rows   (is[0] == 1): the fastest output axis is also the fastest input
//           axis.  A block owns RB whole output rows; the row offsets in the
//           input are decoded once per row into LDS, then the block streams
//           RB * od[0] elements with contiguous reads and writes.
//  * tiled  (is[0] != 1): output axis j has input stride 1.  A block owns a
//           TY x TX tile of (leading output axes 0..b-1 flattened) x (axis j)
//           for one index of the remaining axes; it reads rows of the tile
//           along j (contiguous in the input) into LDS and writes columns
//           along the leading axes (contiguous in the output).
// Division by the axis extents uses 32-bit multiply-high magic numbers
// (round-up method, valid for dividends below 2^31).
#pragma once

#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif

namespace ipeps {
namespace k {

struct FastDiv {
  uint32_t d = 1, m = 1, s = 0;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while (s < 31 && (1u << s) < d) ++s;
  f.s = s;
  f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << s) - d)) / d + 1);
  return f;
}

IPEPS_HD inline uint32_t umulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

// n / f.d for n < 2^31
IPEPS_HD inline uint32_t fdiv(uint32_t n, const FastDiv& f) { return (umulhi32(n, f.m) + n) >> f.s; }

constexpr int PERM_MAXD = 8;

// A block covers rb rows times `chunk` elements of each row: whole rows
// (chunk = od0) when rows are short, else one row split in nchunk chunks.
struct PermRowsP {
  int nd;                      // merged output axes; is[0] == 1
  uint32_t n, nrows, od0;      // elements, rows (n / od0), row length
  uint32_t rb, chunk, nchunk;  // rows per block, elements per row per block, chunks per row
  FastDiv fchunk;
  FastDiv fod[PERM_MAXD];      // extents of axes 1..nd-1 (index a-1)
  uint32_t is[PERM_MAXD];      // input strides of axes 1..nd-1 (index a-1)
};

// phase 1: thread t < rb decodes the input offset of the block's t-th row
IPEPS_HD inline void perm_rows_phase1(const PermRowsP& p, uint32_t blk, uint32_t t, uint32_t* row_off) {
  if (t >= p.rb) return;
  uint32_t row = (blk / p.nchunk) * p.rb + t;
  if (row >= p.nrows) return;
  uint32_t r = row, off = 0;
  for (int a = 0; a + 1 < p.nd; ++a) {
    uint32_t q = fdiv(r, p.fod[a]);
    off += (r - q * p.fod[a].d) * p.is[a];
    r = q;
  }
  row_off[t] = off;
}

// element e (0 <= e < rb * chunk) of block blk: (input offset, output offset),
// or false when it falls outside the tensor
IPEPS_HD inline bool perm_rows_elem(const PermRowsP& p, uint32_t blk, uint32_t e, const uint32_t* row_off,
                                    uint32_t& in_off, uint32_t& out_off) {
  uint32_t rr = fdiv(e, p.fchunk);
  uint32_t rb_idx = blk / p.nchunk, ch = blk - rb_idx * p.nchunk;
  uint32_t i0 = ch * p.chunk + (e - rr * p.chunk);
  uint32_t row = rb_idx * p.rb + rr;
  if (rr >= p.rb || row >= p.nrows || i0 >= p.od0) return false;
  in_off = row_off[rr] + i0;
  out_off = row * p.od0 + i0;
  return true;
}

struct PermTileP {
  uint32_t wl, wj;     // flattened leading extent, extent of axis j
  uint32_t ostr_j;     // output stride of axis j
  int nlead;           // leading axes 0..nlead-1
  FastDiv flead[PERM_MAXD];
  uint32_t is_lead[PERM_MAXD];
  int nrest;           // remaining axes (output order, j excluded)
  FastDiv frest[PERM_MAXD];
  uint32_t is_rest[PERM_MAXD], os_rest[PERM_MAXD];
  uint32_t ntx, nty, nrest_total;  // tiles along j, tiles along lead, rest indices
  uint32_t TX, TY;
};

// block (bx over ntx * nty tiles, rest index) -> tile origin and rest offsets
IPEPS_HD inline void perm_tile_block(const PermTileP& p, uint32_t bx, uint32_t rest, uint32_t& x0, uint32_t& y0,
                                     uint32_t& in_rest, uint32_t& out_rest) {
  uint32_t ty = bx / p.ntx, tx = bx - ty * p.ntx;
  x0 = tx * p.TX;
  y0 = ty * p.TY;
  uint32_t r = rest;
  in_rest = 0;
  out_rest = 0;
  for (int a = 0; a < p.nrest; ++a) {
    uint32_t q = fdiv(r, p.frest[a]);
    uint32_t dgt = r - q * p.frest[a].d;
    in_rest += dgt * p.is_rest[a];
    out_rest += dgt * p.os_rest[a];
    r = q;
  }
}

// input offset of leading index y (0 <= y < wl)
IPEPS_HD inline uint32_t perm_tile_lead(const PermTileP& p, uint32_t y) {
  uint32_t off = 0;
  for (int a = 0; a < p.nlead; ++a) {
    uint32_t q = fdiv(y, p.flead[a]);
    off += (y - q * p.flead[a].d) * p.is_lead[a];
    y = q;
  }
  return off;
}

// ---- host-side planning ----------------------------------------------------
// od / is: merged output extents and input strides (nd >= 2, n < 2^31).
// Returns 0 for the rows kernel (rp filled), 1 for the tiled kernel (tp), -1
// when no axis has input stride 1 (not produced by the merging step).
inline int plan_permute(int nd, const int64_t* od, const int64_t* is, int esize, PermRowsP& rp, PermTileP& tp) {
  int64_t n = 1;
  for (int a = 0; a < nd; ++a) n *= od[a];
  if (is[0] == 1) {
    rp = PermRowsP();
    rp.nd = nd;
    rp.n = (uint32_t)n;
    rp.od0 = (uint32_t)od[0];
    rp.nrows = (uint32_t)(n / od[0]);
    const uint32_t target = 4096;
    if (rp.od0 >= target / 2) {
      rp.rb = 1;
      rp.chunk = target;
      rp.nchunk = (rp.od0 + target - 1) / target;
    } else {
      uint32_t rb = target / rp.od0;
      if (rb > 256) rb = 256;
      if (rb > rp.nrows) rb = rp.nrows;
      rp.rb = rb < 1 ? 1 : rb;
      rp.chunk = rp.od0;
      rp.nchunk = 1;
    }
    rp.fchunk = make_fastdiv(rp.chunk);
    for (int a = 1; a < nd; ++a) {
      rp.fod[a - 1] = make_fastdiv((uint32_t)od[a]);
      rp.is[a - 1] = (uint32_t)is[a];
    }
    return 0;
  }
  int j = -1;
  for (int a = 1; a < nd; ++a)
    if (is[a] == 1) j = a;
  tp = PermTileP();
  if (j < 0) return -1;
  int b = 1;
  int64_t wl = od[0];
  while (b < j && wl < 64) wl *= od[b++];
  tp.wl = (uint32_t)wl;
  tp.wj = (uint32_t)od[j];
  int64_t os = 1;
  for (int a = 0; a < j; ++a) os *= od[a];
  tp.ostr_j = (uint32_t)os;
  tp.nlead = b;
  for (int a = 0; a < b; ++a) {
    tp.flead[a] = make_fastdiv((uint32_t)od[a]);
    tp.is_lead[a] = (uint32_t)is[a];
  }
  tp.nrest = 0;
  int64_t ostr = 1, nrest = 1;
  for (int a = 0; a < nd; ++a) {
    if (a >= b && a != j) {
      tp.frest[tp.nrest] = make_fastdiv((uint32_t)od[a]);
      tp.is_rest[tp.nrest] = (uint32_t)is[a];
      tp.os_rest[tp.nrest] = (uint32_t)ostr;
      ++tp.nrest;
      nrest *= od[a];
    }
    ostr *= od[a];
  }
  tp.TX = tp.TY = esize >= 16 ? 32 : 64;
  tp.ntx = (tp.wj + tp.TX - 1) / tp.TX;
  tp.nty = (tp.wl + tp.TY - 1) / tp.TY;
  tp.nrest_total = (uint32_t)nrest;
  return 1;
}

}  // namespace k
}  // namespace ipeps
