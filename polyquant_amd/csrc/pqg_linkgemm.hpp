// Shared device pieces of the link-GEMM sigma kernels (pqg_linkgemm.hip: FP64/FP32 MFMA
// kernel and host; pqg_linkgemm_b3.hip: BF16x3 kernel).
#pragma once

#include <algorithm>
#include <cstdint>

#include "pqg_ci.hpp"

namespace pqg {

constexpr int LG_NT = 32;  // columns per tile (FP64; FP32 takes LG_NT32 when its LDS allows)
constexpr int LG_NT32 = 64;
constexpr int LG_MAXL = 64;     // links per row tuple (four 16-row M blocks)
constexpr int LG_MAXIN = 160;   // in-row rows
constexpr int LG_MAXBO = 128;   // off-diagonal pairs of the column set (those go through the map)
constexpr int LG_MAXSETS = 3;

// Row tuples range over up to three string sets (electron alpha or beta, positron alpha,
// positron beta): row index = sum_u i_u * rstride[u].  The "active" sets (act[] = their
// position in the row tuple) contribute links, numbered from loff[s] in set order.
template <typename T>
struct LgArgs {
  const T* C;          // this vector, n_r x n_c row-major
  T* U;                // chunk scratch [n_r][L][ldu]
  int nrow;            // sets in the row tuple
  int64_t rstride[LG_MAXSETS];
  int rns[LG_MAXSETS];
  int nsets;           // active sets
  int act[LG_MAXSETS];
  const int2* links[LG_MAXSETS];  // per active set [n_s][L_s]: {J | odd << 31, Q << 16 | rev}
  int Ls[LG_MAXSETS], nvs[LG_MAXSETS], loff[LG_MAXSETS], np[LG_MAXSETS];
  const T* Vrc[LG_MAXSETS];              // (P|Q): P a pair of active set s, Q a pair of the column set
  const T* Vrr[LG_MAXSETS][LG_MAXSETS];  // (P|Q): P of active set s, Q of active set t (charges folded in)
  const T* Aone[LG_MAXSETS];             // one-string part of singles of active set s, [L_s][n_s]
  const T* Xrow[LG_MAXSETS][LG_MAXSETS]; // spectator sums of row set u for singles of active set s, [n_u][np_s]
  const int32_t* map;  // column-set map [nin][n_c]: (column index << 1 | odd) or -1
  const uint64_t* cbits;  // column strings (bit j = orbital j occupied)
  int nbo, ncd, nin;   // in-row rows: off-diagonal pairs, diagonal pairs, + the own row
  int np_c;
  int n_r, n_c, L, K;
  int c0, cw, gcols;
  int ldu;             // U row stride: cw rounded up to 4 elements, so rows start 16-byte aligned
  int accumulate;      // gather: add to S (1) or overwrite it (0)
  int lab_off, kbf_off, crow_off, tile_off;  // LDS byte offsets (links at 0)
  // BF16x3 kernel: this vector as bf16 pairs; the column-set map transposed to [m][kb], each
  // word the byte offset of the target pair in the staged row | odd << 31
  const uint32_t* Cp;
  const int32_t* mapT;
  int part_off;        // LDS byte offset of the partial sums of the K split
  int lq_off;          // LDS byte offset of the links (the staged row is at offset 0)
  uint32_t cp_bytes, mapt_bytes;  // sizes of Cp and mapT (buffer loads past them return 0)
  int bstride;         // LDS tile row stride in 32-bit words (one row per column)
};

// kernel-argument arrays indexed by a run-time set number, without a copy to scratch
template <class X>
__device__ __forceinline__ X pick3(const X (&v)[LG_MAXSETS], int s) {
  return s == 0 ? v[0] : (s == 1 ? v[1] : v[2]);
}
template <class X>
__device__ __forceinline__ X pick33(const X (&v)[LG_MAXSETS][LG_MAXSETS], int s, int t) {
  return s == 0 ? pick3(v[0], t) : (s == 1 ? pick3(v[1], t) : pick3(v[2], t));
}

template <typename T>
__device__ __forceinline__ int row_index(const LgArgs<T>& a, int64_t J, int u) {
  return (int)((J / pick3(a.rstride, u)) % pick3(a.rns, u));
}

// Links of row J into LDS: lq[e] = {target row | odd << 31, Q << 16 | rev}, lab[e] =
// set << 16 | hole number << 8 | particle number.
template <typename T>
__device__ void lg_row_links(const LgArgs<T>& a, int64_t J, int2* lq, int* lab, int tid, int nth) {
  for (int e = tid; e < a.L; e += nth) {
    int s = 0;
    if (a.nsets > 1 && e >= a.loff[1]) s = 1;
    if (a.nsets > 2 && e >= a.loff[2]) s = 2;
    const int el = e - pick3(a.loff, s);
    const int u = pick3(a.act, s);
    const int64_t stride = pick3(a.rstride, u);
    const int64_t i = row_index(a, J, u);
    const int2 l = pick3(a.links, s)[(size_t)i * pick3(a.Ls, s) + el];
    const int64_t tgt = J + ((int64_t)(l.x & 0x7fffffff) - i) * stride;
    lq[e] = make_int2((int)tgt | (l.x & (int)0x80000000), (l.y & ~0xffff) | (pick3(a.loff, s) + (l.y & 0xffff)));
    const int nv = pick3(a.nvs, s);
    const int ae = el / nv;
    lab[e] = (s << 16) | (ae << 8) | (el - ae * nv);
  }
}

// p of the off-diagonal pair number k = p (p - 1) / 2 + q, q < p
__device__ __forceinline__ int offpair_hi(int k) {
  int p = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)k)) * 0.5f);
  while (p * (p - 1) / 2 > k) --p;
  while (p * (p + 1) / 2 <= k) ++p;
  return p;
}

// One-string part of the single excitation of link ep of row J (the own-row K entry): the
// one-body and same-string spectator part plus the other row strings' spectator sums.
template <typename T>
__device__ __forceinline__ T lg_own_entry(const LgArgs<T>& a, int64_t J, int ep, int Qp, int sp) {
  const int el = ep - pick3(a.loff, sp);
  const int us = pick3(a.act, sp);
  T own = pick3(a.Aone, sp)[(size_t)el * pick3(a.rns, us) + row_index(a, J, us)];
  for (int u = 0; u < a.nrow; u++) {
    const T* X = pick33(a.Xrow, sp, u);
    if (X) own += X[(size_t)row_index(a, J, u) * pick3(a.np, sp) + Qp];
  }
  return own;
}

// A_J[ep][k] for a valid link ep (pair Qp, hole/particle numbers abp, set sp) of row J:
// in-row rows k < nin - 1 couple with the column set's pairs, k = nin - 1 is the own row, and
// the row rows k >= nin (the links e = k - nin) carry the same-set doubles and the couplings of
// two row sets; k >= K is padding.
template <typename T>
__device__ __forceinline__ T lg_a_entry(const LgArgs<T>& a, const int2* lq, const int* lab, const int* kbf, int Qp,
                                        int abp, int sp, T own, int k) {
  const int nin = a.nin;
  if (k >= a.K) return T(0);
  if (k < nin - 1) return pick3(a.Vrc, sp)[(size_t)Qp * a.np_c + kbf[k]];
  if (k == nin - 1) return own;
  const int e = k - nin;
  const int abe = lab[e];
  const int se = abe >> 16;
  const int2 l = lq[e];
  if (se == sp) {
    // same-set doubles: the two links share neither hole nor particle
    if (((abe ^ abp) & 0xff00) && ((abe ^ abp) & 0xff))
      return (l.x < 0 ? T(-0.5) : T(0.5)) * pick33(a.Vrr, sp, sp)[(size_t)Qp * pick3(a.np, sp) + (l.y >> 16)];
  } else if (sp < se) {
    // one excitation in each of two row sets: only the path whose first link lies in the
    // lower set carries the coupling
    return (l.x < 0 ? T(-1) : T(1)) * pick33(a.Vrr, sp, se)[(size_t)Qp * pick3(a.np, se) + (l.y >> 16)];
  }
  return T(0);
}

// BF16x3 kernel shape for a pass: MB 32-link M blocks, KS K parts of KQ steps of 16, the LDS
// layout (byte offsets) and the tile column stride in bf16 elements
struct LgB3Shape {
  int mb = 0, ks = 0, kq = 0, stride = 0;
  size_t lds = 0;
  int lq_off = 0, lab_off = 0, kbf_off = 0, tile_off = 0, part_off = 0;
};

constexpr int LG_B3_NT = 32;  // columns per BF16x3 tile
// launches k_lg_b3 for the shape (pqg_linkgemm_b3.hip); k_lg_to_pairs converts a vector
void lg_b3_launch(const LgArgs<float>& a, const LgB3Shape& sh, dim3 grid, hipStream_t st);
void lg_to_pairs(const float* x, uint32_t* y, int64_t n, hipStream_t st);

}  // namespace pqg
