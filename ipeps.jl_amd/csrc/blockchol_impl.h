// Blocked Cholesky factor and triangular inverse of a small SPD matrix in one
// workgroup (FP64, n <= 160): the tile layout and index helpers of the device
// kernel small_chol_inv_kernel in elem.hip.
//
// G (n x n, lower triangle read) = L L^T; the kernel returns L^-1 (lower,
// zeros above).  The matrix is padded to NB 16 x 16 tiles with the identity
// (the padded factor and inverse are block-diagonal, so the leading n x n
// block is exact).  The lower tiles live in LDS, tile (I, J), I >= J, at
// index I (I + 1) / 2 + J, each column-major with leading dimension 17.
//
// Cholesky, right-looking over tile columns K:
//  (a) one wave factors the diagonal tile in registers (lane i holds row i,
//      columns broadcast with readlane; 1 / sqrt of each pivot by v_rsq_f64
//      and Newton steps) and replaces it by its inverse Dinv_K;
//  (b) the panel below: L_IK = A_IK Dinv_K^T, a wave per tile;
//  (c) the trailing tiles: A_IJ -= L_IK L_JK^T for K < J <= I, a wave per tile.
// Inverse X = L^-1, over tile columns J = NB-2 .. 0, a wave per row I > J:
//  X_IJ = -(sum_{M=J+1..I} X_IM L_MJ) X_JJ with X_II = Dinv_I; the column's
//  tiles are held in registers until every wave has read column J of L.
// Each tile product is four v_mfma_f64_16x16x4f64 (16 x 16 x 16).  The sum
// is formed transposed, W^T = sum L_MJ^T X_IM^T, whose accumulator layout
// (lane l, element e: W^T[l/16 + 4e][l%16]) is the A operand layout of the
// product with X_JJ at k step e, so W never goes through LDS.
#pragma once

#include <stdint.h>

#ifndef IPEPS_HD
#define IPEPS_HD __host__ __device__
#endif

namespace ipeps {
namespace k {

constexpr int BCH_T = 16;                // tile size (one MFMA output block)
constexpr int BCH_LD = 17;               // tile leading dimension (LDS banks)
constexpr int BCH_TE = BCH_T * BCH_LD;   // doubles per tile
constexpr int BCH_NBMAX = 10;            // tiles per dimension
constexpr int BCH_MAXN = BCH_T * BCH_NBMAX;
constexpr int BCH_NTILES = BCH_NBMAX * (BCH_NBMAX + 1) / 2;
constexpr int BCH_THREADS = 512;

IPEPS_HD inline int bch_tix(int I, int J) { return I * (I + 1) / 2 + J; }
IPEPS_HD inline int bch_at(int I, int J, int r, int c) { return bch_tix(I, J) * BCH_TE + c * BCH_LD + r; }

}  // namespace k
}  // namespace ipeps
