// CI determinant set and the matrix-free sigma build on gfx950.
//
// Layout in HBM
//   * strings: one uint64 per string (orbital o = bit o), per (species, spin) set, sorted by
//     integer value.  A string's index is its colex rank (combinatorial number system),
//     computed from 8-bit chunk tables: rank(S) = sum_c T_c[byte_c][electrons at or below c].
//   * determinant (internal) address: ((i_ea * n_eb + i_eb) * n_pa + i_pa) * n_pb + i_pb.
//     The reference's addresses (level-major string order, determinant_creation.cpp:139-207)
//     are mapped at the boundary; for truncated spaces the internal product space carries a
//     validity mask.
//   * CI vectors: N x k column-major (k vectors of one Davidson block).
//   * integrals: coupling blocks V[t][u] = (P|Q) row-major [npair_t][npair_u] per ordered
//     pair of string sets, charge factor q_t*q_u folded in for mixed species; per-string tables
//     A (one-string part of single-excitation elements), D (double-excitation elements with
//     phase), X (spectator sums sum_{i in occ(I_u)} (P|ii)), E1/DW (diagonal parts).
//
// Sigma (pull form, no atomics, deterministic): sigma = H C is split by which strings change.
//   pass A, one workgroup per "row" string of the first set: stages C(row, :) in LDS and
//     accumulates, for every column determinant, the diagonal, all excitations of the
//     column-side strings (gathered from the staged row) and, for every single excitation
//     row -> J of the row string, the row x column doubles from the staged row C(J, :);
//   pass B, the same kernel on the transposed vector: excitations of the first string only.
// Excitations are enumerated on the fly from the bitstrings (ctz loops), phases from
// popcounts, excited-string indices from the rank tables.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <unordered_map>

#include <hipcub/hipcub.hpp>

#include "pqg_ci.hpp"
#include "pqg_strings.hpp"

namespace pqg {

// ---------------------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ int pairidx(int i, int j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// Visit all single excitations I -> J of one string: (e, h, p, Q = idx2(h, p), odd phase, J).
// e enumerates occupied h (ascending) outer, virtual p (ascending) inner.
template <class F>
__device__ __forceinline__ void for_singles(uint64_t I, int n_orb, F&& f) {
  const uint64_t vm = ~I & low_mask(n_orb);
  int e = 0;
  for (uint64_t occ = I; occ; occ &= occ - 1) {
    const int h = ctz64(occ);
    for (uint64_t vir = vm; vir; vir &= vir - 1) {
      const int p = ctz64(vir);
      const int lo = min(h, p), hi = max(h, p);
      const int odd = __popcll(I & between_mask(lo, hi)) & 1;
      f(e, h, p, hi * (hi + 1) / 2 + lo, odd, I ^ ((1ull << h) | (1ull << p)));
      ++e;
    }
  }
}

// Visit all same-string double excitations: holes h0 < h1, particles p0 < p1.
template <class F>
__device__ __forceinline__ void for_doubles(uint64_t I, int n_orb, F&& f) {
  const uint64_t vm = ~I & low_mask(n_orb);
  int e = 0;
  for (uint64_t o0 = I; o0; o0 &= o0 - 1) {
    const int h0 = ctz64(o0);
    for (uint64_t o1 = o0 & (o0 - 1); o1; o1 &= o1 - 1) {
      const int h1 = ctz64(o1);
      for (uint64_t v0 = vm; v0; v0 &= v0 - 1) {
        const int p0 = ctz64(v0);
        for (uint64_t v1 = v0 & (v0 - 1); v1; v1 &= v1 - 1) {
          const int p1 = ctz64(v1);
          f(e, h0, h1, p0, p1, I ^ ((1ull << h0) | (1ull << h1) | (1ull << p0) | (1ull << p1)));
          ++e;
        }
      }
    }
  }
}

// get_phase for a double excitation (determinant_operations.cpp:47-116), one word.
__device__ __forceinline__ int double_phase_odd(uint64_t I, int h0, int h1, int p0, int p1) {
  int a = min(h0, p0), b = max(h0, p0), c = min(h1, p1), d = max(h1, p1);
  int n = __popcll(I & between_mask(a, b)) + __popcll(I & between_mask(c, d));
  if (a < c && c < b && b < d) n++;
  return n & 1;
}

// ---------------------------------------------------------------------------------------
// table builders (once per CI)
// ---------------------------------------------------------------------------------------
// A[e][I] = h_hp + sum_{i in occ(I)} ((hp|ii) - (hi|ip))
__global__ void k_tab_singles(SlotDev s, const double* __restrict__ h, const double* __restrict__ V, int npair,
                              double* __restrict__ A) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= s.nstr) return;
  const uint64_t b = s.bits[I];
  const int n = s.n_orb;
  for_singles(b, n, [&](int e, int hh, int pp, int Q, int, uint64_t) {
    double v = h[hh * n + pp];
    for (uint64_t o = b; o; o &= o - 1) {
      const int i = ctz64(o);
      v += V[(size_t)Q * npair + pairidx(i, i)] - V[(size_t)pairidx(hh, i) * npair + pairidx(i, pp)];
    }
    A[(size_t)e * s.nstr + I] = v;
  });
}

// D[e][I] = phase * ((h0 p0|h1 p1) - (h0 p1|h1 p0))   (slater_condon_singleparticle.cpp:107-175)
__global__ void k_tab_doubles(SlotDev s, const double* __restrict__ V, int npair, double* __restrict__ D) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= s.nstr) return;
  const uint64_t b = s.bits[I];
  for_doubles(b, s.n_orb, [&](int e, int h0, int h1, int p0, int p1, uint64_t) {
    double v = V[(size_t)pairidx(h0, p0) * npair + pairidx(h1, p1)] - V[(size_t)pairidx(h0, p1) * npair + pairidx(h1, p0)];
    D[(size_t)e * s.nstr + I] = double_phase_odd(b, h0, h1, p0, p1) ? -v : v;
  });
}

// X[I_u][P] = sum_{i in occ(I_u)} V_tu[P][ii]   (V_tu includes the charge factor)
__global__ void k_tab_spectator(SlotDev su, const double* __restrict__ Vtu, int npair_t, int npair_u,
                                double* __restrict__ X) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)su.nstr * npair_t) return;
  const int I = (int)(idx / npair_t), P = (int)(idx % npair_t);
  const uint64_t b = su.bits[I];
  double v = 0.0;
  for (uint64_t o = b; o; o &= o - 1) {
    const int i = ctz64(o);
    v += Vtu[(size_t)P * npair_u + pairidx(i, i)];
  }
  X[idx] = v;
}

// E1[I] = sum_{i in occ} h_ii + 1/2 sum_{i,j in occ} ((ii|jj) - (ij|ij))
__global__ void k_tab_e1(SlotDev s, const double* __restrict__ h, const double* __restrict__ V, int npair,
                         double* __restrict__ E1) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= s.nstr) return;
  const uint64_t b = s.bits[I];
  const int n = s.n_orb;
  double e1 = 0.0, e2 = 0.0;
  for (uint64_t o = b; o; o &= o - 1) {
    const int i = ctz64(o);
    e1 += h[i * n + i];
    for (uint64_t q = b; q; q &= q - 1) {
      const int j = ctz64(q);
      e2 += V[(size_t)pairidx(i, i) * npair + pairidx(j, j)] - V[(size_t)pairidx(i, j) * npair + pairidx(i, j)];
    }
  }
  E1[I] = e1 + 0.5 * e2;
}

// DW[I_u][i] = sum_{j in occ(I_u)} w(i, j):  w = 1/2 (V_tu[ii][jj] + V_ut[jj][ii]) (same species)
// or V_tu[ii][jj] (mixed species, charge factor already in V_tu).
__global__ void k_tab_diagpair(SlotDev su, int n_t, const double* __restrict__ Vtu, const double* __restrict__ Vut,
                               int npair_t, int npair_u, double* __restrict__ DW) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)su.nstr * n_t) return;
  const int I = (int)(idx / n_t), i = (int)(idx % n_t);
  const uint64_t b = su.bits[I];
  double v = 0.0;
  for (uint64_t o = b; o; o &= o - 1) {
    const int j = ctz64(o);
    if (Vut)
      v += 0.5 * (Vtu[(size_t)pairidx(i, i) * npair_u + pairidx(j, j)] + Vut[(size_t)pairidx(j, j) * npair_t + pairidx(i, i)]);
    else
      v += Vtu[(size_t)pairidx(i, i) * npair_u + pairidx(j, j)];
  }
  DW[idx] = v;
}

struct DiagArgs {
  SlotDev sl[kSlots];
  const double* E1[kSlots];
  const double* DW[kSlots][kSlots];
  int64_t stride[kSlots];
  int nslot;
  int64_t n;
  const uint8_t* valid;
};

// diagonal_Hii (slater_condon.cpp:102-153), internal order
__global__ void k_hdiag(DiagArgs a, double* __restrict__ Hd) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= a.n) return;
  if (a.valid && !a.valid[I]) {
    Hd[I] = 1e30;
    return;
  }
  int idx[kSlots];
  uint64_t bits[kSlots];
  int64_t r = I;
  for (int t = a.nslot - 1; t >= 0; --t) {
    idx[t] = (int)(r % a.sl[t].nstr);
    r /= a.sl[t].nstr;
    bits[t] = a.sl[t].bits[idx[t]];
  }
  double e = 0.0;
  for (int t = 0; t < a.nslot; t++) e += a.E1[t][idx[t]];
  for (int t = 0; t < a.nslot; t++)
    for (int u = t + 1; u < a.nslot; u++) {
      const double* dw = a.DW[t][u];
      if (!dw) continue;
      const int nt = a.sl[t].n_orb;
      for (uint64_t o = bits[t]; o; o &= o - 1) e += dw[(size_t)idx[u] * nt + ctz64(o)];
    }
  Hd[I] = e;
}

// ---------------------------------------------------------------------------------------
// sigma pass kernel
// ---------------------------------------------------------------------------------------
constexpr int kCPT = 16;   // max columns per thread
constexpr int kGMax = 8;   // excited rows of the row string staged per block (row x column doubles)

enum PassFlags { F_DIAG = 1, F_COLS = 2, F_MIXED = 4, F_ACCUM = 8, F_NODBL = 16, F_NOPAIR = 32 };

template <typename T>
struct PassArgs {
  SlotDev sl[kSlots];
  Tables<T> tb;
  int nr, rs[3], rn[3];  // row part: slots and string counts
  int nc, cs[3], cn[3];  // column part
  int64_t cstride[3];    // column-index stride of each column slot
  int64_t M;             // columns per row
  int tiles, tile_cols, cpt;
  int G;                 // excited rows staged together (1..kGMax)
  const T* C;
  int64_t ldc;
  T* S;
  int64_t lds;
  const double* Hd;
  int flags;
  int off_x, off_v, off_rt, off_rl;  // LDS byte offsets
  int x_off[3], v_off[3], rt_off[3];
  int vstride;                       // elements per staged coupling row (sum of npair of the column slots)
};

template <int NC>
constexpr int pass_threads() { return NC == 1 ? 1024 : 256; }

// String bit helpers on a W-bit word (W = uint32_t when every orbital count is <= 32).
template <typename W>
__device__ __forceinline__ int ctzw(W x) {
  if constexpr (sizeof(W) == 4)
    return __builtin_ctz(x);
  else
    return __builtin_ctzll(x);
}
template <typename W>
__device__ __forceinline__ int popw(W x) {
  if constexpr (sizeof(W) == 4)
    return __popc(x);
  else
    return __popcll(x);
}
template <typename W>
__device__ __forceinline__ W lowmaskw(int n) {
  return n >= (int)(8 * sizeof(W)) ? ~W(0) : ((W(1) << n) - W(1));
}

// Visit all single excitations I -> J: (e, Q = idx2(h, p), odd phase, J), e enumerating
// occupied h (ascending) outer and virtual p (ascending) inner.
template <typename W, class F>
__device__ __forceinline__ void for_singles_w(W I, int n_orb, F&& f) {
  const W vm = ~I & lowmaskw<W>(n_orb);
  int e = 0;
  for (W occ = I; occ; occ &= occ - 1) {
    const int h = ctzw(occ);
    for (W vir = vm; vir; vir &= vir - 1) {
      const int p = ctzw(vir);
      const int lo = min(h, p), hi = max(h, p);
      // occupied orbitals strictly between h and p
      const W btw = lowmaskw<W>(hi) & ~lowmaskw<W>(lo + 1);
      f(e, hi * (hi + 1) / 2 + lo, popw(I & btw) & 1, I ^ ((W(1) << h) | (W(1) << p)));
      ++e;
    }
  }
}

template <typename W, class F>
__device__ __forceinline__ void for_doubles_w(W I, int n_orb, F&& f) {
  const W vm = ~I & lowmaskw<W>(n_orb);
  int e = 0;
  for (W o0 = I; o0; o0 &= o0 - 1) {
    const W b0 = o0 & (~o0 + 1);
    for (W o1 = o0 & (o0 - 1); o1; o1 &= o1 - 1) {
      const W hh = b0 | (o1 & (~o1 + 1));
      for (W v0 = vm; v0; v0 &= v0 - 1) {
        const W c0 = v0 & (~v0 + 1);
        for (W v1 = v0 & (v0 - 1); v1; v1 &= v1 - 1) {
          f(e, I ^ (hh | c0 | (v1 & (~v1 + 1))));
          ++e;
        }
      }
    }
  }
}

// colex rank of J (chunk tables in LDS) -> index in the set, or -1
template <typename W>
__device__ __forceinline__ int str_index_w(const SlotDev& s, const uint32_t* tab, W J) {
  if (s.nocc == 1 && !s.r2i) return ctzw(J);
  uint32_t r = 0;
  int e = s.nocc;
  const int w = s.nocc + 1;
  for (int c = s.nchunk - 1; c >= 0; --c) {
    const uint32_t x = (uint32_t)(J >> (8 * c)) & 255u;
    r += tab[((c << 8) + x) * w + e];
    e -= __popc(x);
  }
  return s.r2i ? s.r2i[r] : (int)r;
}

// Column strings of column index m.
template <typename T, typename W, int NC>
__device__ __forceinline__ void decode_cols(const PassArgs<T>& a, int64_t m, int (&ci)[NC], W (&cb)[NC]) {
  int64_t r = m;
#pragma unroll
  for (int c = NC - 1; c >= 0; --c) {
    ci[c] = (int)(r % a.cn[c]);
    r /= a.cn[c];
  }
#pragma unroll
  for (int c = 0; c < NC; c++) cb[c] = (W)a.sl[a.cs[c]].bits[ci[c]];
}

// Diagonal and every excitation that changes only column strings, gathered from the row
// C(row, :) (crow: staged in LDS or read in place).
template <typename T, typename W, int NC>
__device__ __forceinline__ T column_terms(const PassArgs<T>& a, int64_t row, int64_t m, const T* crow, const T* xs,
                                          const uint32_t* rt) {
  T sum = 0;
  if (a.flags & F_DIAG) sum = (T)a.Hd[row * a.M + m] * crow[m];
  if (!(a.flags & F_COLS)) return sum;
  int ci[NC];
  W cb[NC];
  decode_cols<T, W, NC>(a, m, ci, cb);
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int t = a.cs[c];
    const SlotDev& s = a.sl[t];
    if (s.nocc == 0 || s.nvirt == 0) continue;
    const uint32_t* tab = rt + a.rt_off[c];
    const T* __restrict__ At = a.tb.A[t];
    const T* xrow = xs + a.x_off[c];
    const int it = ci[c];
    const int64_t cst = a.cstride[c];
    const T* cbase = crow + m - (int64_t)it * cst;  // crow[m + (jt - it) * cst] = cbase[jt * cst]
    // singles of column string c: s * (A + spectator sums) * C(row, m')
    for_singles_w<W>(cb[c], s.n_orb, [&](int e, int Q, int odd, W J) {
      const int jt = str_index_w<W>(s, tab, J);
      if (jt < 0) return;
      T v = At[(size_t)e * s.nstr + it] + xrow[Q];
#pragma unroll
      for (int c2 = 0; c2 < NC; c2++) {
        if (c2 == c) continue;
        const T* X = a.tb.X[t][a.cs[c2]];
        if (X) v += X[(size_t)ci[c2] * s.npair + Q];
      }
      const T cv = cbase[(int64_t)jt * cst];
      sum += odd ? -v * cv : v * cv;
    });
    // same-string doubles (unless the link GEMM forms them)
    if (s.ndouble > 0 && !(a.flags & F_NODBL)) {
      const T* __restrict__ Dt = a.tb.D[t] + it;
      for_doubles_w<W>(cb[c], s.n_orb, [&](int e, W J) {
        const int jt = str_index_w<W>(s, tab, J);
        if (jt < 0) return;
        sum += Dt[(size_t)e * s.nstr] * cbase[(int64_t)jt * cst];
      });
    }
    // one single in each of two column strings (unless the link GEMM forms them)
    if (a.flags & F_NOPAIR) continue;
#pragma unroll
    for (int c2 = c + 1; c2 < NC; c2++) {
      const int u = a.cs[c2];
      const SlotDev& s2 = a.sl[u];
      const T* Vtu = a.tb.V[t][u];
      if (!Vtu || s2.nocc == 0 || s2.nvirt == 0) continue;
      const uint32_t* tab2 = rt + a.rt_off[c2];
      const int iu = ci[c2];
      const int64_t cst2 = a.cstride[c2];
      const W cb2 = cb[c2];
      for_singles_w<W>(cb[c], s.n_orb, [&](int, int Q, int odd, W J) {
        const int jt = str_index_w<W>(s, tab, J);
        if (jt < 0) return;
        const T* cb2base = cbase + (int64_t)jt * cst - (int64_t)iu * cst2;
        const T* vrowQ = Vtu + (size_t)Q * s2.npair;
        T part = 0;
        for_singles_w<W>(cb2, s2.n_orb, [&](int, int Q2, int odd2, W J2) {
          const int ju = str_index_w<W>(s2, tab2, J2);
          if (ju < 0) return;
          const T cv = cb2base[(int64_t)ju * cst2];
          const T v = vrowQ[Q2];
          part += odd2 ? -v * cv : v * cv;
        });
        sum += odd ? -part : part;
      });
    }
  }
  return sum;
}

// For a block of g_n single excitations of the row string (their sign-scaled coupling rows
// staged in vr, the excited rows C(J_g, :) at crows[g]): sum over single excitations of the
// column strings.  Each column excitation is enumerated once for all g_n row excitations.
template <typename T, typename W, int NC>
__device__ __forceinline__ T mixed_terms(const PassArgs<T>& a, int t0, int64_t m, const T* const (&crows)[kGMax],
                                         const T* vr, int g_n, const uint32_t* rt) {
  int ci[NC];
  W cb[NC];
  decode_cols<T, W, NC>(a, m, ci, cb);
  T part = 0;
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const int t = a.cs[c];
    const SlotDev& s = a.sl[t];
    if (s.nocc == 0 || s.nvirt == 0 || !a.tb.V[t0][t]) continue;
    const uint32_t* tab = rt + a.rt_off[c];
    const T* vrow = vr + a.v_off[c];
    const int it = ci[c];
    const int64_t cst = a.cstride[c];
    const int64_t mb = m - (int64_t)it * cst;
    const int vs = a.vstride;
    for_singles_w<W>(cb[c], s.n_orb, [&](int, int Q, int odd, W J) {
      const int jt = str_index_w<W>(s, tab, J);
      if (jt < 0) return;
      const int64_t mj = mb + (int64_t)jt * cst;
      T v = 0;
#pragma unroll
      for (int g = 0; g < kGMax; g++)
        if (g < g_n) v += vrow[g * vs + Q] * crows[g][mj];
      part += odd ? -v : v;
    });
  }
  return part;
}

template <typename T, typename W, int NC, int CPT, bool STAGE>
__global__ __launch_bounds__(pass_threads<NC>()) void k_sigma_pass(const PassArgs<T> a) {
  extern __shared__ __align__(16) char smem[];
  T* rows_lds = reinterpret_cast<T*>(smem);
  T* xs = reinterpret_cast<T*>(smem + a.off_x);
  T* vr = reinterpret_cast<T*>(smem + a.off_v);
  uint32_t* rt = reinterpret_cast<uint32_t*>(smem + a.off_rt);
  int2* rowlist = reinterpret_cast<int2*>(smem + a.off_rl);
  __shared__ int n_rowlist;

  const int root = blockIdx.y;
  const int64_t row = blockIdx.x / a.tiles;
  const int tile = blockIdx.x % a.tiles;
  const int tid = threadIdx.x, nth = blockDim.x;
  const T* __restrict__ Cr = a.C + (int64_t)root * a.ldc;
  const int64_t M = a.M;

  int ri[3];
  {
    int64_t r = row;
    for (int q = a.nr - 1; q >= 0; --q) {
      ri[q] = (int)(r % a.rn[q]);
      r /= a.rn[q];
    }
  }
  // rank tables and spectator sums for the column strings
#pragma unroll
  for (int c = 0; c < NC; c++) {
    const SlotDev& s = a.sl[a.cs[c]];
    const int tsz = s.nchunk * 256 * (s.nocc + 1);
    for (int i = tid; i < tsz; i += nth) rt[a.rt_off[c] + i] = s.rtab[i];
    const int t = a.cs[c];
    for (int P = tid; P < s.npair; P += nth) {
      T v = 0;
      for (int q = 0; q < a.nr; q++) {
        const T* X = a.tb.X[t][a.rs[q]];
        if (X) v += X[(size_t)ri[q] * s.npair + P];
      }
      xs[a.x_off[c] + P] = v;
    }
  }
  const T* crow;
  if constexpr (STAGE) {
    for (int64_t m = tid; m < M; m += nth) rows_lds[m] = Cr[row * M + m];
    crow = rows_lds;
  } else {
    crow = Cr + row * M;
  }
  const bool mixed = (a.flags & F_MIXED) && a.nr == 1;
  if (mixed && tid == 0) {
    // single excitations of the row string: (index offset of the excited row, Q | odd << 30)
    const int t0 = a.rs[0];
    const SlotDev& s0 = a.sl[t0];
    const int i0 = ri[0];
    int n = 0;
    for_singles_w<W>((W)s0.bits[i0], s0.n_orb, [&](int, int Q0, int odd0, W J0) {
      const int j0 = str_index_w<W>(s0, s0.rtab, J0);
      if (j0 < 0) return;
      rowlist[n++] = make_int2(j0 - i0, Q0 | (odd0 << 30));
    });
    n_rowlist = n;
  }
  __syncthreads();

  const int64_t mbase = (int64_t)tile * a.tile_cols + tid;
  const int64_t mend = min(M, (int64_t)(tile + 1) * a.tile_cols);
  T acc[CPT];
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int64_t m = mbase + (int64_t)j * nth;
    acc[j] = (m < mend) ? column_terms<T, W, NC>(a, row, m, crow, xs, rt) : (T)0;
  }

  if (mixed) {
    const int t0 = a.rs[0];
    const int nrl = n_rowlist;
    const int G = a.G;
    for (int a0 = 0; a0 < nrl; a0 += G) {
      const int g_n = min(G, nrl - a0);
      __syncthreads();  // previous block's rows and couplings are no longer read
      const T* crows[kGMax];
#pragma unroll
      for (int g = 0; g < kGMax; g++) {
        const int2 rl = rowlist[a0 + (g < g_n ? g : 0)];
        const T* src = Cr + (row + rl.x) * M;
        if constexpr (STAGE) {
          T* dst = rows_lds + (int64_t)g * M;
          if (g < g_n)
            for (int64_t m = tid; m < M; m += nth) dst[m] = src[m];
          crows[g] = dst;
        } else {
          crows[g] = src;
        }
        if (g < g_n) {
          const int Q0 = rl.y & ((1 << 30) - 1);
          const T sg = (rl.y >> 30) ? (T)-1 : (T)1;
#pragma unroll
          for (int c = 0; c < NC; c++) {
            const T* V0t = a.tb.V[t0][a.cs[c]];
            const int np = a.sl[a.cs[c]].npair;
            for (int P = tid; P < np; P += nth)
              vr[g * a.vstride + a.v_off[c] + P] = V0t ? sg * V0t[(size_t)Q0 * np + P] : (T)0;
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < CPT; j++) {
        const int64_t m = mbase + (int64_t)j * nth;
        if (m < mend) acc[j] += mixed_terms<T, W, NC>(a, t0, m, crows, vr, g_n, rt);
      }
    }
  }

  T* __restrict__ Sr = a.S + (int64_t)root * a.lds;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int64_t m = mbase + (int64_t)j * nth;
    if (m < mend) {
      if (a.flags & F_ACCUM)
        Sr[row * M + m] += acc[j];
      else
        Sr[row * M + m] = acc[j];
    }
  }
}

// out[c * rows + r] (+)= in[r * cols + c] for each of k vectors; 64 x 64 tiles through LDS
template <typename T>
__global__ __launch_bounds__(256) void k_transpose(const T* __restrict__ in, int64_t ldi, T* __restrict__ out,
                                                   int64_t ldo, int64_t rows, int64_t cols, int accumulate) {
  __shared__ T tile[64][65];
  const int64_t r0 = (int64_t)blockIdx.y * 64, c0 = (int64_t)blockIdx.x * 64;
  const T* src = in + (int64_t)blockIdx.z * ldi;
  T* dst = out + (int64_t)blockIdx.z * ldo;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 64; i += 4) {
    const int64_t r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) tile[i][tx] = src[r * cols + c];
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {
    const int64_t c = c0 + i, r = r0 + tx;
    if (r < rows && c < cols) {
      if (accumulate)
        dst[c * rows + r] += tile[tx][i];
      else
        dst[c * rows + r] = tile[tx][i];
    }
  }
}

// ---------------------------------------------------------------------------------------
// boundary permutations, masks, conversions
// ---------------------------------------------------------------------------------------
struct PermArgs {
  const int32_t* r2i[kSlots];
  int64_t nref[kSlots];
  int64_t stride[kSlots];
  int nslot;
  const int64_t* table;  // truncated spaces: explicit map
  int64_t n;
};

__device__ __forceinline__ int64_t ref_to_int(const PermArgs& p, int64_t a) {
  if (p.table) return p.table[a];
  int64_t r = a, addr = 0;
  for (int t = p.nslot - 1; t >= 0; --t) {
    const int64_t d = r % p.nref[t];
    r /= p.nref[t];
    addr += (int64_t)p.r2i[t][d] * p.stride[t];
  }
  return addr;
}

__global__ void k_ref2int(PermArgs p, const double* __restrict__ ref, int64_t ldr, double* __restrict__ in, int64_t ldi) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  const int k = blockIdx.y;
  in[k * ldi + ref_to_int(p, a)] = ref[k * ldr + a];
}

__global__ void k_int2ref(PermArgs p, const double* __restrict__ in, int64_t ldi, double* __restrict__ ref, int64_t ldr) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  const int k = blockIdx.y;
  ref[k * ldr + a] = in[k * ldi + ref_to_int(p, a)];
}

// Full CI with two string sets: the reference address is i_a * n_b + i_b and the internal one
// r2i_a[i_a] * n_b + r2i_b[i_b], so a row moves as a whole and is permuted within itself.  One
// workgroup per (row, vector) stages the row in LDS, so both the read and the write are
// contiguous (the element-wise kernels above scatter one side).  A row of up to 20480 strings
// takes up to 160 KB of LDS, one workgroup per CU, so the workgroup has 1024 threads to keep
// enough loads in flight.
__global__ __launch_bounds__(1024) void k_ref2int_rows(const int32_t* __restrict__ r2i_a, const int32_t* __restrict__ r2i_b,
                                                      int nb, const double* __restrict__ ref, int64_t ldr,
                                                      double* __restrict__ in, int64_t ldi) {
  extern __shared__ double row[];
  const int64_t i = blockIdx.x;
  const double* src = ref + (int64_t)blockIdx.y * ldr + i * nb;
  for (int j = threadIdx.x; j < nb; j += 1024) row[r2i_b[j]] = src[j];
  __syncthreads();
  double* dst = in + (int64_t)blockIdx.y * ldi + (int64_t)r2i_a[i] * nb;
  for (int j = threadIdx.x; j < nb; j += 1024) dst[j] = row[j];
}

__global__ __launch_bounds__(1024) void k_int2ref_rows(const int32_t* __restrict__ r2i_a, const int32_t* __restrict__ r2i_b,
                                                      int nb, const double* __restrict__ in, int64_t ldi,
                                                      double* __restrict__ ref, int64_t ldr) {
  extern __shared__ double row[];
  const int64_t i = blockIdx.x;
  const double* src = in + (int64_t)blockIdx.y * ldi + (int64_t)r2i_a[i] * nb;
  for (int j = threadIdx.x; j < nb; j += 1024) row[j] = src[j];
  __syncthreads();
  double* dst = ref + (int64_t)blockIdx.y * ldr + i * nb;
  for (int j = threadIdx.x; j < nb; j += 1024) dst[j] = row[r2i_b[j]];
}

template <typename T>
__global__ void k_mask(const uint8_t* __restrict__ valid, T* __restrict__ v, int64_t ld, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && !valid[i]) v[blockIdx.y * ld + i] = 0;
}

__global__ void k_d2f(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

// ---------------------------------------------------------------------------------------
// element evaluation (Slater_Condon, slater_condon.cpp:4-100) and parity dumps
// ---------------------------------------------------------------------------------------
struct ElemArgs {
  SlotDev sl[kSlots];
  Tables<double> tb;
  int64_t stride[kSlots];
  int nslot;
  const double* Hd;
};

__device__ __forceinline__ int rank_among(uint64_t set, int bit) { return __popcll(set & low_mask(bit)); }

// Slater_Condon element between the determinants at internal addresses ia and ja
__device__ double element_at(const ElemArgs& a, int64_t ia, int64_t ja) {
  if (ia == ja) return a.Hd[ia];
  int ii[kSlots], jj[kSlots], deg[kSlots];
  uint64_t bi[kSlots], bj[kSlots];
  int64_t ri = ia, rj = ja;
  int total = 0;
  for (int t = a.nslot - 1; t >= 0; --t) {
    ii[t] = (int)(ri % a.sl[t].nstr);
    ri /= a.sl[t].nstr;
    jj[t] = (int)(rj % a.sl[t].nstr);
    rj /= a.sl[t].nstr;
    bi[t] = a.sl[t].bits[ii[t]];
    bj[t] = a.sl[t].bits[jj[t]];
    deg[t] = __popcll(bi[t] ^ bj[t]) / 2;
    total += deg[t];
  }
  double v = 0.0;
  if (total == 1) {
    int t = 0;
    while (deg[t] == 0) t++;
    const SlotDev& s = a.sl[t];
    const int h = ctz64(bi[t] & ~bj[t]), p = ctz64(bj[t] & ~bi[t]);
    const uint64_t vm = ~bi[t] & low_mask(s.n_orb);
    const int e = rank_among(bi[t], h) * s.nvirt + rank_among(vm, p);
    const int Q = pairidx(h, p);
    v = a.tb.A[t][(size_t)e * s.nstr + ii[t]];
    for (int u = 0; u < a.nslot; u++)
      if (u != t && a.tb.X[t][u]) v += a.tb.X[t][u][(size_t)ii[u] * s.npair + Q];
    if (__popcll(bi[t] & between_mask(min(h, p), max(h, p))) & 1) v = -v;
  } else if (total == 2) {
    int t = -1, u = -1;
    for (int k = 0; k < a.nslot; k++)
      if (deg[k]) {
        if (t < 0)
          t = k;
        else
          u = k;
      }
    if (u < 0) {  // double within one string
      const SlotDev& s = a.sl[t];
      const uint64_t hs = bi[t] & ~bj[t], ps = bj[t] & ~bi[t];
      const int h0 = ctz64(hs), h1 = 63 - __clzll(hs), p0 = ctz64(ps), p1 = 63 - __clzll(ps);
      const uint64_t vm = ~bi[t] & low_mask(s.n_orb);
      const int no = s.nocc, nv = s.nvirt;
      const int a0 = rank_among(bi[t], h0), a1 = rank_among(bi[t], h1);
      const int b0 = rank_among(vm, p0), b1 = rank_among(vm, p1);
      const int eh = a0 * (2 * no - a0 - 1) / 2 + (a1 - a0 - 1);
      const int ep = b0 * (2 * nv - b0 - 1) / 2 + (b1 - b0 - 1);
      const int e = eh * (nv * (nv - 1) / 2) + ep;
      v = a.tb.D[t][(size_t)e * s.nstr + ii[t]];
    } else if (a.tb.V[t][u]) {  // one single in each of two strings
      const int ht = ctz64(bi[t] & ~bj[t]), pt = ctz64(bj[t] & ~bi[t]);
      const int hu = ctz64(bi[u] & ~bj[u]), pu = ctz64(bj[u] & ~bi[u]);
      v = a.tb.V[t][u][(size_t)pairidx(ht, pt) * a.sl[u].npair + pairidx(hu, pu)];
      int odd = __popcll(bi[t] & between_mask(min(ht, pt), max(ht, pt))) +
                __popcll(bi[u] & between_mask(min(hu, pu), max(hu, pu)));
      if (odd & 1) v = -v;
    }
  }
  return v;
}

__global__ void k_elements(ElemArgs a, const int64_t* __restrict__ I, const int64_t* __restrict__ J, int64_t npairs,
                           double* __restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  out[q] = element_at(a, I[q], J[q]);
}

// Dense H in reference address order (column j = blockIdx.y), from the element of the lower
// reference address pair (min, max) for both positions, so H is exactly symmetric as the
// reference's mirrored upper triangle is (epci.cpp:634-643).
__global__ void k_dense_h(ElemArgs a, PermArgs p, int64_t n, int64_t j0, double* __restrict__ H, int64_t ldh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = j0 + blockIdx.y;
  if (i >= n) return;
  const int64_t lo = min(i, j), hi = max(i, j);
  H[i + (int64_t)blockIdx.y * ldh] = element_at(a, ref_to_int(p, lo), ref_to_int(p, hi));
}

// Upper-triangle CSR of H in reference order (create_ham, explicit_ham.cpp:51-84 with
// single_species_create_ham_singleshot / two_species_create_ham_singleshot): row r holds the
// diagonal (always) and every connected determinant J with reference address above r whose
// Slater_Condon element is nonzero; with a screening threshold s > 0 every entry with |v| <= s is
// dropped (ham.prune(s, 1.0)).  One thread per row enumerates the singles and doubles of each
// string and the single x single pairs of two strings.
struct CsrArgs {
  ElemArgs e;
  PermArgs p;
  const int64_t* inv;             // truncated spaces: internal address -> reference address (-1: not in space)
  const int32_t* i2r[kSlots];     // full CI: internal -> reference string index per set
  double screen;
};

__device__ __forceinline__ int64_t csr_ref_of(const CsrArgs& c, int64_t addr) {
  if (c.inv) return c.inv[addr];
  int64_t r = 0;
  for (int t = c.e.nslot - 1; t >= 0; --t) {
    const int64_t d = addr % c.e.sl[t].nstr;
    addr /= c.e.sl[t].nstr;
    r += (int64_t)c.i2r[t][d] * c.e.stride[t];
  }
  return r;
}

// Every determinant address connected to the one at internal address ia by a single or double
// excitation within one string, or by singles in two strings (targets outside a truncated set
// skipped; targets outside a truncated product space are the caller's to drop).
template <class F>
__device__ void for_connected(const ElemArgs& a, int64_t ia, F&& visit) {
  int idx[kSlots];
  uint64_t bits[kSlots];
  int64_t rr = ia;
  for (int t = a.nslot - 1; t >= 0; --t) {
    idx[t] = (int)(rr % a.sl[t].nstr);
    rr /= a.sl[t].nstr;
    bits[t] = a.sl[t].bits[idx[t]];
  }
  for (int t = 0; t < a.nslot; t++) {
    const SlotDev& s = a.sl[t];
    for_singles(bits[t], s.n_orb, [&](int, int, int, int, int, uint64_t J) {
      const int j = str_index(s, s.rtab, J);
      if (j >= 0) visit(ia + (int64_t)(j - idx[t]) * a.stride[t]);
    });
    for_doubles(bits[t], s.n_orb, [&](int, int, int, int, int, uint64_t J) {
      const int j = str_index(s, s.rtab, J);
      if (j >= 0) visit(ia + (int64_t)(j - idx[t]) * a.stride[t]);
    });
    for (int u = t + 1; u < a.nslot; u++) {
      const SlotDev& su = a.sl[u];
      for_singles(bits[t], s.n_orb, [&](int, int, int, int, int, uint64_t J) {
        const int j = str_index(s, s.rtab, J);
        if (j < 0) return;
        const int64_t base = ia + (int64_t)(j - idx[t]) * a.stride[t];
        for_singles(bits[u], su.n_orb, [&](int, int, int, int, int, uint64_t K) {
          const int k = str_index(su, su.rtab, K);
          if (k >= 0) visit(base + (int64_t)(k - idx[u]) * a.stride[u]);
        });
      });
    }
  }
}

__device__ __forceinline__ bool csr_keep(const CsrArgs& c, double v, bool diag) {
  return (diag || v != 0.0) && (c.screen <= 0.0 || fabs(v) > c.screen);
}

// Row r (reference order) of the upper triangle: (reference column, value).
template <class F>
__device__ void csr_row(const CsrArgs& c, int64_t r, F&& emit) {
  const int64_t ia = ref_to_int(c.p, r);
  if (csr_keep(c, c.e.Hd[ia], true)) emit(r, c.e.Hd[ia]);
  for_connected(c.e, ia, [&](int64_t addr) {
    const int64_t rj = csr_ref_of(c, addr);
    if (rj <= r) return;  // lower triangle, or not in the space (-1)
    const double v = element_at(c.e, ia, addr);
    if (csr_keep(c, v, false)) emit(rj, v);
  });
}

// Row ia (internal order) of the symmetric matrix with the same entries: (internal column, value),
// each value from the lower reference address of the pair as in csr_row.
template <class F>
__device__ void xh_row(const CsrArgs& c, int64_t ia, F&& emit) {
  const int64_t r = csr_ref_of(c, ia);
  if (r < 0) return;
  if (csr_keep(c, c.e.Hd[ia], true)) emit(ia, c.e.Hd[ia]);
  for_connected(c.e, ia, [&](int64_t addr) {
    const int64_t rj = csr_ref_of(c, addr);
    if (rj < 0) return;
    const double v = rj > r ? element_at(c.e, ia, addr) : element_at(c.e, addr, ia);
    if (csr_keep(c, v, false)) emit(addr, v);
  });
}

__global__ void k_xh_count(CsrArgs c, int64_t n, int64_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t m = 0;
  xh_row(c, r, [&](int64_t, double) { ++m; });
  cnt[r] = m;
}

__global__ void k_xh_fill(CsrArgs c, int64_t n, const int64_t* __restrict__ rowptr, int32_t* __restrict__ cols,
                          double* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t q = rowptr[r];
  xh_row(c, r, [&](int64_t j, double v) {
    cols[q] = (int32_t)j;
    vals[q] = v;
    ++q;
  });
}

// S[:, v] = H C[:, v] over the stored rows, one thread per row, every vector (fixed order).
__global__ void k_xh_spmv(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                          const double* __restrict__ val, int64_t n, const double* __restrict__ C, double* __restrict__ S,
                          int k, int64_t ld) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t q0 = ptr[r], q1 = ptr[r + 1];
  for (int v = 0; v < k; v++) {
    const double* x = C + (int64_t)v * ld;
    double s = 0.0;
    for (int64_t q = q0; q < q1; q++) s += val[q] * x[col[q]];
    S[(int64_t)v * ld + r] = s;
  }
}

__global__ void k_csr_count(CsrArgs c, int64_t n, int64_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t m = 0;
  csr_row(c, r, [&](int64_t, double) { ++m; });
  cnt[r] = m;
}

__global__ void k_csr_fill(CsrArgs c, int64_t n, const int64_t* __restrict__ rowptr, int32_t* __restrict__ cols,
                           double* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t q = rowptr[r];
  csr_row(c, r, [&](int64_t j, double v) {
    cols[q] = (int32_t)j;
    vals[q] = v;
    ++q;
  });
}

__global__ void k_inv_addr(const int64_t* __restrict__ ref_addr, int64_t nref, int64_t* __restrict__ inv) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nref) inv[ref_addr[r]] = r;
}

__global__ void k_fill_i64(int64_t* __restrict__ x, int64_t n, int64_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

// (S^2 - e) between two determinants given by the strings of every slot, for the species whose
// alpha and beta sets are slots ta and tb (create_S_sq_minus_expected_S_sq_matrix_singleshot,
// s2.cpp:156-255).  Diagonal: n_a + sz (sz - 1) - e, with x = alpha ^ beta the singly occupied
// orbitals, n_a = |x & alpha| and sz = n_a - |x| / 2.  Off the diagonal: -phase_a phase_b when
// alpha moves i -> j and beta j -> i with every other string equal, else 0.
__device__ double s2_pen_elem(const uint64_t* bi, const uint64_t* bj, int nslot, int ta, int tb, double e) {
  for (int u = 0; u < nslot; u++)
    if (u != ta && u != tb && bi[u] != bj[u]) return 0.0;
  const uint64_t ai = bi[ta], aj = bj[ta], Bi = bi[tb], Bj = bj[tb];
  if (ai == aj && Bi == Bj) {
    const uint64_t x = ai ^ Bi;
    const int na = __popcll(x & ai);
    const double sz = na - 0.5 * __popcll(x);
    return na + sz * (sz - 1.0) - e;
  }
  if (__popcll(ai ^ aj) != 2 || __popcll(Bi ^ Bj) != 2) return 0.0;
  const int ha = ctz64(ai & ~aj), pa = ctz64(aj & ~ai), hb = ctz64(Bi & ~Bj), pb = ctz64(Bj & ~Bi);
  if (ha != pb || hb != pa) return 0.0;
  const int odd = __popcll(ai & between_mask(min(ha, pa), max(ha, pa))) +
                  __popcll(Bi & between_mask(min(hb, pb), max(hb, pb)));
  return (odd & 1) ? 1.0 : -1.0;
}

struct S2PenArgs {
  int ta, tb, power;
  double e;
  const uint8_t* valid;  // truncated spaces: internal address -> in space; nullptr for full CI
};

// Dense P = S^2 - e (power 1) or P^2 (power 2) of one species in reference address order, from
// the (min, max) reference pair for both positions (exactly symmetric).  P^2 entry (I, J) is
// P_II P_IJ + sum over the spin flips K of I that lie in the space of P_IK P_KJ: those are the
// only nonzero P_IK (the sparse product of s2.cpp:147).
__global__ void k_dense_s2(ElemArgs a, PermArgs p, S2PenArgs q, int64_t n, int64_t j0, double* __restrict__ S,
                           int64_t lds) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t j = j0 + blockIdx.y;
  if (i >= n) return;
  const int64_t lo = min(i, j), hi = max(i, j);
  int64_t ri = ref_to_int(p, lo), rj = ref_to_int(p, hi);
  const int64_t ia = ri;
  int ii[kSlots];
  uint64_t bi[kSlots], bj[kSlots];
  for (int t = a.nslot - 1; t >= 0; --t) {
    ii[t] = (int)(ri % a.sl[t].nstr);
    ri /= a.sl[t].nstr;
    bi[t] = a.sl[t].bits[ii[t]];
    bj[t] = a.sl[t].bits[rj % a.sl[t].nstr];
    rj /= a.sl[t].nstr;
  }
  double v = s2_pen_elem(bi, bj, a.nslot, q.ta, q.tb, q.e);
  if (q.power == 2) {
    double w = s2_pen_elem(bi, bi, a.nslot, q.ta, q.tb, q.e) * v;
    const uint64_t A = bi[q.ta], B = bi[q.tb];
    const SlotDev &sa = a.sl[q.ta], &sb = a.sl[q.tb];
    for (uint64_t xa = A & ~B; xa; xa &= xa - 1) {
      const int oi = ctz64(xa);
      for (uint64_t xb = B & ~A; xb; xb &= xb - 1) {
        const int oj = ctz64(xb);
        uint64_t bk[kSlots];
        for (int t = 0; t < a.nslot; t++) bk[t] = bi[t];
        bk[q.ta] = A ^ (1ull << oi) ^ (1ull << oj);
        bk[q.tb] = B ^ (1ull << oj) ^ (1ull << oi);
        if (q.valid) {
          const int ka = str_index(sa, sa.rtab, bk[q.ta]), kb = str_index(sb, sb.rtab, bk[q.tb]);
          if (ka < 0 || kb < 0) continue;
          const int64_t addr = ia + (int64_t)(ka - ii[q.ta]) * a.stride[q.ta] + (int64_t)(kb - ii[q.tb]) * a.stride[q.tb];
          if (!q.valid[addr]) continue;
        }
        w += s2_pen_elem(bi, bk, a.nslot, q.ta, q.tb, q.e) * s2_pen_elem(bk, bj, a.nslot, q.ta, q.tb, q.e);
      }
    }
    v = w;
  }
  S[i + (int64_t)blockIdx.y * lds] = v;
}

// Row ia (internal order) of P = S^2 - e of one species: the diagonal and the spin flips of I
// that lie in the space, the only nonzero entries (s2_pen_elem).
template <class F>
__device__ void xp_row(const ElemArgs& a, const S2PenArgs& q, int64_t ia, F&& emit) {
  if (q.valid && !q.valid[ia]) return;
  int ii[kSlots];
  uint64_t bi[kSlots];
  int64_t rr = ia;
  for (int t = a.nslot - 1; t >= 0; --t) {
    ii[t] = (int)(rr % a.sl[t].nstr);
    rr /= a.sl[t].nstr;
    bi[t] = a.sl[t].bits[ii[t]];
  }
  emit(ia, s2_pen_elem(bi, bi, a.nslot, q.ta, q.tb, q.e));
  const uint64_t A = bi[q.ta], B = bi[q.tb];
  const SlotDev &sa = a.sl[q.ta], &sb = a.sl[q.tb];
  for (uint64_t xa = A & ~B; xa; xa &= xa - 1) {
    const int oi = ctz64(xa);
    for (uint64_t xb = B & ~A; xb; xb &= xb - 1) {
      const int oj = ctz64(xb);
      uint64_t bk[kSlots];
      for (int t = 0; t < a.nslot; t++) bk[t] = bi[t];
      bk[q.ta] = A ^ (1ull << oi) ^ (1ull << oj);
      bk[q.tb] = B ^ (1ull << oj) ^ (1ull << oi);
      const int ka = str_index(sa, sa.rtab, bk[q.ta]), kb = str_index(sb, sb.rtab, bk[q.tb]);
      if (ka < 0 || kb < 0) continue;
      const int64_t addr = ia + (int64_t)(ka - ii[q.ta]) * a.stride[q.ta] + (int64_t)(kb - ii[q.tb]) * a.stride[q.tb];
      if (q.valid && !q.valid[addr]) continue;
      emit(addr, s2_pen_elem(bi, bk, a.nslot, q.ta, q.tb, q.e));
    }
  }
}

__global__ void k_xp_count(ElemArgs a, S2PenArgs q, int64_t n, int64_t* __restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t m = 0;
  xp_row(a, q, r, [&](int64_t, double) { ++m; });
  cnt[r] = m;
}

__global__ void k_xp_fill(ElemArgs a, S2PenArgs q, int64_t n, const int64_t* __restrict__ rowptr,
                          int32_t* __restrict__ cols, double* __restrict__ vals) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int64_t p = rowptr[r];
  xp_row(a, q, r, [&](int64_t j, double v) {
    cols[p] = (int32_t)j;
    vals[p] = v;
    ++p;
  });
}

// Y[:, v] = s * M X[:, v] (+ Y[:, v] when acc) over stored rows, one thread per row.
__global__ void k_csr_apply(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const double* __restrict__ val, int64_t n, const double* __restrict__ X, double* __restrict__ Y,
                            int k, int64_t ld, double s, int acc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int64_t q0 = ptr[r], q1 = ptr[r + 1];
  for (int v = 0; v < k; v++) {
    const double* x = X + (int64_t)v * ld;
    double t = 0.0;
    for (int64_t q = q0; q < q1; q++) t += val[q] * x[col[q]];
    double* y = Y + (int64_t)v * ld + r;
    *y = acc ? *y + s * t : s * t;
  }
}

// Diagonal of the stored explicit H (entries outside the space keep diagonal_Hii's 1e30).
__global__ void k_xdiag_init(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                             const double* __restrict__ val, const double* __restrict__ hd, int64_t n,
                             double* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int64_t q = ptr[r]; q < ptr[r + 1]; q++)
    if (col[q] == r) d += val[q];
  out[r] = hd[r] >= 1e29 ? hd[r] : d;
}
// out += alpha diag(P) (power 1) or alpha diag(P^2) = alpha sum_k P_rk^2 (power 2, P symmetric)
__global__ void k_xdiag_pen(const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const double* __restrict__ val, int64_t n, double alpha, int power,
                            double* __restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n || out[r] >= 1e29) return;
  double d = 0.0;
  for (int64_t q = ptr[r]; q < ptr[r + 1]; q++) {
    if (power == 1) {
      if (col[q] == r) d += val[q];
    } else {
      d += val[q] * val[q];
    }
  }
  out[r] += alpha * d;
}

// For every string I of a set: its singles (which = 1) or doubles (which = 2) as
// (excited string internal index, phase bit), in enumeration order.
__global__ void k_dump_excitations(SlotDev s, int which, int per, int32_t* __restrict__ J, int8_t* __restrict__ odd) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= s.nstr) return;
  const uint64_t b = s.bits[I];
  if (which == 1) {
    for_singles(b, s.n_orb, [&](int e, int, int, int, int o, uint64_t Jb) {
      J[(size_t)I * per + e] = str_index(s, s.rtab, Jb);
      odd[(size_t)I * per + e] = (int8_t)o;
    });
  } else {
    for_doubles(b, s.n_orb, [&](int e, int h0, int h1, int p0, int p1, uint64_t Jb) {
      J[(size_t)I * per + e] = str_index(s, s.rtab, Jb);
      odd[(size_t)I * per + e] = (int8_t)double_phase_odd(b, h0, h1, p0, p1);
    });
  }
}

// ---------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------
static uint64_t binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  k = std::min(k, n - k);
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return (uint64_t)r;
}

static int grid1(int64_t n, int b) { return (int)((n + b - 1) / b); }

// Level-major strings (create_det + create_unique_excitation, determinant_creation.cpp:6-146):
// the HF string, then for k = 1..max_level every choice of k occupied orbitals (outer, in
// lexicographic order) and k virtual orbitals (inner, lexicographic).
static std::vector<uint64_t> reference_strings(int n, int nocc, int max_level) {
  std::vector<uint64_t> out;
  const uint64_t hf = nocc ? ((nocc >= 64) ? ~0ull : ((1ull << nocc) - 1)) : 0ull;
  out.push_back(hf);
  const int nv = n - nocc;
  for (int k = 1; k <= max_level; k++) {
    PQG_REQUIRE(k <= nv, PQG_ERR_ARG, "Excitation level exceeds virtual size!");
    if (k > nocc) continue;  // no combination of k occupied orbitals
    // lexicographic k-subsets of {0..m-1} as index vectors
    auto subsets = [](int m, int kk) {
      std::vector<std::vector<int>> all;
      std::vector<int> c(kk);
      std::iota(c.begin(), c.end(), 0);
      while (true) {
        all.push_back(c);
        int i = kk - 1;
        while (i >= 0 && c[i] == m - kk + i) i--;
        if (i < 0) break;
        c[i]++;
        for (int j = i + 1; j < kk; j++) c[j] = c[j - 1] + 1;
      }
      return all;
    };
    const auto occs = subsets(nocc, k);
    const auto virts = subsets(nv, k);
    for (const auto& oc : occs)
      for (const auto& vc : virts) {
        uint64_t s = hf;
        for (int o : oc) s &= ~(1ull << o);
        for (int v : vc) s |= 1ull << (nocc + v);
        out.push_back(s);
      }
  }
  return out;
}

// colex rank of a string with k set bits
static uint64_t colex_rank(uint64_t s) {
  uint64_t r = 0;
  int j = 1;
  for (uint64_t b = s; b; b &= b - 1, j++) r += binom(__builtin_ctzll(b), j);
  return r;
}

static void build_slot(pqg_ci* ci, Slot& sl, int n, int nocc, int max_level) {
  pqg_ctx* ctx = ci->ctx;
  PQG_REQUIRE(n >= 1 && n <= 63, PQG_ERR_UNSUPPORTED, "n_orb must be in 1..63 (one 64-bit word per string)");
  PQG_REQUIRE(nocc >= 0 && nocc <= n, PQG_ERR_ARG, "particle count exceeds the orbital count");
  sl.n_orb = n;
  sl.nocc = nocc;
  sl.nvirt = n - nocc;
  sl.npair = n * (n + 1) / 2;
  sl.nchunk = (n + 7) / 8;
  sl.ref = reference_strings(n, nocc, max_level);
  const uint64_t total = binom(n, nocc);
  PQG_REQUIRE(total < (1ull << 31), PQG_ERR_UNSUPPORTED, "string set too large for 32-bit string indices");
  sl.nstr = (int)sl.ref.size();
  sl.complete = ((uint64_t)sl.nstr == total);
  sl.sorted = sl.ref;
  std::sort(sl.sorted.begin(), sl.sorted.end());
  std::unordered_map<uint64_t, int> pos;
  pos.reserve(sl.nstr * 2);
  for (int i = 0; i < sl.nstr; i++) pos[sl.sorted[i]] = i;
  sl.ref2int.resize(sl.nstr);
  sl.int2ref.resize(sl.nstr);
  sl.deg.resize(sl.nstr);
  for (int r = 0; r < sl.nstr; r++) {
    const int i = pos[sl.ref[r]];
    sl.ref2int[r] = i;
    sl.int2ref[i] = r;
    sl.deg[r] = __builtin_popcountll(sl.ref[r] ^ sl.ref[0]) / 2;
  }
  // rank tables: T_c[x][e] = sum_j C(g_j, e - m + j) over the m set bits g_1 < .. < g_m of x
  const int w = nocc + 1;
  std::vector<uint32_t> tab((size_t)sl.nchunk * 256 * w, 0);
  for (int c = 0; c < sl.nchunk; c++)
    for (int x = 0; x < 256; x++) {
      const int m = __builtin_popcount(x);
      for (int e = m; e <= nocc; e++) {
        uint64_t v = 0;
        int j = 1;
        for (int bb = 0; bb < 8; bb++)
          if ((x >> bb) & 1) {
            v += binom(8 * c + bb, e - m + j);
            j++;
          }
        tab[((size_t)c * 256 + x) * w + e] = (uint32_t)v;
      }
    }
  sl.rtab.reset(ctx, tab.size());
  upload(ci->stream, sl.rtab.get(), tab.data(), tab.size() * 4);
  sl.bits.reset(ctx, sl.nstr);
  upload(ci->stream, sl.bits.get(), sl.sorted.data(), sl.nstr * 8);
  if (!sl.complete) {
    std::vector<int32_t> r2i(total, -1);
    for (int i = 0; i < sl.nstr; i++) r2i[colex_rank(sl.sorted[i])] = i;
    sl.r2i.reset(ctx, total);
    upload(ci->stream, sl.r2i.get(), r2i.data(), total * 4);
  }
  sl.d_ref2int.reset(ctx, sl.nstr);
  upload(ci->stream, sl.d_ref2int.get(), sl.ref2int.data(), sl.nstr * 4);
  sl.nsingle = (int64_t)sl.nocc * sl.nvirt;
  sl.ndouble = (int64_t)(sl.nocc * (sl.nocc - 1) / 2) * (sl.nvirt * (sl.nvirt - 1) / 2);
}

static PermArgs perm_args(const pqg_ci* ci) {
  PermArgs p{};
  p.nslot = ci->nslot;
  for (int t = 0; t < ci->nslot; t++) {
    p.r2i[t] = ci->slot[t].d_ref2int.get();
    p.nref[t] = ci->slot[t].nstr;
    p.stride[t] = ci->stride[t];
  }
  p.table = ci->fci ? nullptr : ci->ref_addr.get();
  p.n = ci->n_ref;
  return p;
}

// the row-staged permutation applies to full CI with two string sets whose beta row fits LDS
static bool rows_permutable(const pqg_ci* ci) {
  const int nb = ci->slot[1].nstr;
  return ci->fci && ci->nslot == 2 && nb >= 256 && (size_t)nb * 8 <= 160 * 1024 && ci->slot[0].nstr < (1 << 30);
}

static void set_row_lds(const void* kern, size_t bytes) {
  if (bytes > 64 * 1024)
    PQG_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

void to_internal(pqg_ci* ci, const double* ref, int64_t ldr, double* in, int64_t ldi, int k, hipStream_t st) {
  if (rows_permutable(ci)) {
    const int nb = ci->slot[1].nstr;
    const size_t lds = (size_t)nb * 8;
    set_row_lds(reinterpret_cast<const void*>(k_ref2int_rows), lds);
    k_ref2int_rows<<<dim3((unsigned)ci->slot[0].nstr, (unsigned)k), 1024, lds, st>>>(
        ci->slot[0].d_ref2int.get(), ci->slot[1].d_ref2int.get(), nb, ref, ldr, in, ldi);
    PQG_HIP(hipGetLastError());
    return;
  }
  if (!ci->fci)
    for (int j = 0; j < k; j++) PQG_HIP(hipMemsetAsync(in + j * ldi, 0, ci->n_int * 8, st));
  dim3 g(grid1(ci->n_ref, 256), k);
  k_ref2int<<<g, 256, 0, st>>>(perm_args(ci), ref, ldr, in, ldi);
  PQG_HIP(hipGetLastError());
}

void to_reference(pqg_ci* ci, const double* in, int64_t ldi, double* ref, int64_t ldr, int k, hipStream_t st) {
  if (rows_permutable(ci)) {
    const int nb = ci->slot[1].nstr;
    const size_t lds = (size_t)nb * 8;
    set_row_lds(reinterpret_cast<const void*>(k_int2ref_rows), lds);
    k_int2ref_rows<<<dim3((unsigned)ci->slot[0].nstr, (unsigned)k), 1024, lds, st>>>(
        ci->slot[0].d_ref2int.get(), ci->slot[1].d_ref2int.get(), nb, in, ldi, ref, ldr);
    PQG_HIP(hipGetLastError());
    return;
  }
  dim3 g(grid1(ci->n_ref, 256), k);
  k_int2ref<<<g, 256, 0, st>>>(perm_args(ci), in, ldi, ref, ldr);
  PQG_HIP(hipGetLastError());
}

template <typename T>
void mask_invalid(pqg_ci* ci, T* v, int64_t ld, int k, hipStream_t st) {
  if (ci->fci) return;
  dim3 g(grid1(ci->n_int, 256), k);
  k_mask<T><<<g, 256, 0, st>>>(ci->valid.get(), v, ld, ci->n_int);
  PQG_HIP(hipGetLastError());
}
template void mask_invalid<double>(pqg_ci*, double*, int64_t, int, hipStream_t);
template void mask_invalid<float>(pqg_ci*, float*, int64_t, int, hipStream_t);

// row-major device block V_tu = (P|Q) with P a pair of set t, Q of set u; charge factor f.
static void upload_blocks(pqg_ci* ci, const pqg_ci_desc* d) {
  pqg_ctx* ctx = ci->ctx;
  for (int t = 0; t < ci->nslot; t++)
    for (int u = 0; u < ci->nslot; u++) {
      const Slot &st = ci->slot[t], &su = ci->slot[u];
      int a = st.species, sa = st.spinblk, b = su.species, sb = su.spinblk;
      bool transpose = false;
      if (b < a) {
        std::swap(a, b);
        std::swap(sa, sb);
        transpose = true;
      }
      const double* src = d->eri[a][sa][b][sb];
      PQG_REQUIRE(src != nullptr, PQG_ERR_ARG,
                  "missing two-body block eri[" + std::to_string(a) + "][" + std::to_string(sa) + "][" +
                      std::to_string(b) + "][" + std::to_string(sb) + "]");
      const int np_t = st.npair, np_u = su.npair;
      const int na = transpose ? np_u : np_t;  // rows of the stored (column-major) block
      const double f = (st.species == su.species) ? 1.0 : ci->charge[0] * ci->charge[1];
      std::vector<double> h((size_t)np_t * np_u);
      for (int P = 0; P < np_t; P++)
        for (int Q = 0; Q < np_u; Q++)
          h[(size_t)P * np_u + Q] = f * (transpose ? src[(size_t)Q + (size_t)P * na] : src[(size_t)P + (size_t)Q * na]);
      ci->V64[t][u].reset(ctx, h.size());
      upload(ci->stream, ci->V64[t][u].get(), h.data(), h.size() * 8);
    }
}

Tables<double> tables64(pqg_ci* ci) {
  Tables<double> tb;
  for (int t = 0; t < ci->nslot; t++) {
    tb.A[t] = ci->slot[t].A.get();
    tb.D[t] = ci->slot[t].D.get();
    for (int u = 0; u < ci->nslot; u++) {
      tb.X[t][u] = ci->X64[t][u].get();
      tb.V[t][u] = (t != u) ? ci->V64[t][u].get() : nullptr;
    }
  }
  return tb;
}

Tables<float> tables32(pqg_ci* ci) {
  Tables<float> tb;
  for (int t = 0; t < ci->nslot; t++) {
    tb.A[t] = ci->slot[t].A32.get();
    tb.D[t] = ci->slot[t].D32.get();
    for (int u = 0; u < ci->nslot; u++) {
      tb.X[t][u] = ci->X32[t][u].get();
      tb.V[t][u] = (t != u) ? ci->V32[t][u].get() : nullptr;
    }
  }
  return tb;
}

static void to_f32(pqg_ctx* ctx, hipStream_t st, const DevBuf<double>& src, DevBuf<float>& dst) {
  if (!src.get()) return;
  dst.reset(ctx, src.n);
  k_d2f<<<grid1(src.n, 256), 256, 0, st>>>(src.get(), dst.get(), src.n);
  PQG_HIP(hipGetLastError());
}

void ensure_mixed(pqg_ci* ci) {
  if (ci->mixed_ready) return;
  pqg_ctx* ctx = ci->ctx;
  for (int t = 0; t < ci->nslot; t++) {
    to_f32(ctx, ci->stream, ci->slot[t].A, ci->slot[t].A32);
    to_f32(ctx, ci->stream, ci->slot[t].D, ci->slot[t].D32);
    for (int u = 0; u < ci->nslot; u++) {
      to_f32(ctx, ci->stream, ci->X64[t][u], ci->X32[t][u]);
      if (t != u) to_f32(ctx, ci->stream, ci->V64[t][u], ci->V32[t][u]);
    }
  }
  if (ci->lg.enabled)
    for (int t = 0; t < ci->nslot; t++) to_f32(ctx, ci->stream, ci->V64[t][t], ci->lg.V32s[t]);
  PQG_HIP(hipStreamSynchronize(ci->stream));
  ci->mixed_ready = true;
}

pqg_ci* create_ci(pqg_ctx* ctx, const pqg_ci_desc* d, hipStream_t stream) {
  PQG_REQUIRE(ctx && d, PQG_ERR_ARG, "NULL argument");
  PQG_REQUIRE(d->nspecies == 1 || d->nspecies == 2, PQG_ERR_ARG,
              "The CI code can currently handle up to a maximum of 2 unique quantum particle types.");
  PQG_HIP(hipSetDevice(ctx->device));
  std::unique_ptr<pqg_ci> ci(new pqg_ci());
  ci->ctx = ctx;
  ci->ctx_ref.set(ctx);
  ci->stream = stream ? stream : ctx->stream;
  ci->nsp = d->nspecies;
  ci->nslot = 2 * d->nspecies;
  for (int p = 0; p < ci->nsp; p++) {
    ci->charge[p] = d->charge[p];
    ci->nspin[p] = d->nspin_blocks[p];
    PQG_REQUIRE(ci->nspin[p] == 1 || ci->nspin[p] == 2, PQG_ERR_ARG, "nspin_blocks must be 1 or 2");
    PQG_REQUIRE(d->n_orb[p] >= 1 && d->n_orb[p] <= 63, PQG_ERR_UNSUPPORTED,
                "n_orb must be in 1..63 (one 64-bit word per string)");
    for (int s = 0; s < 2; s++) {
      Slot& sl = ci->slot[2 * p + s];
      sl.species = p;
      sl.spin = s;
      sl.spinblk = s % ci->nspin[p];
      build_slot(ci.get(), sl, d->n_orb[p], s == 0 ? d->n_alpha[p] : d->n_beta[p], d->ex_level[p][s]);
    }
  }
  // ---- determinant space (create_excitation, determinant_creation.cpp:148-207) ----------
  for (int t = 0; t < kSlots; t++) ci->nref_str[t] = (t < ci->nslot) ? ci->slot[t].nstr : 1;
  ci->n_int = 1;
  for (int t = ci->nslot - 1; t >= 0; --t) {
    ci->stride[t] = ci->n_int;
    ci->n_int *= ci->slot[t].nstr;
  }
  auto maxdeg = [&](int t) { return *std::max_element(ci->slot[t].deg.begin(), ci->slot[t].deg.end()); };
  const int tot0 = d->ex_level[0][2];
  const int tot1 = ci->nsp == 2 ? d->ex_level[1][2] : 0;
  const int md0 = maxdeg(0) + maxdeg(1);
  const int md1 = ci->nsp == 2 ? maxdeg(2) + maxdeg(3) : 0;
  ci->fci = md0 <= tot0 && (ci->nsp == 1 || md1 <= tot1) && md0 + md1 <= d->max_collective;
  if (ci->fci) {
    ci->n_ref = ci->n_int;
  } else {
    std::vector<int64_t> addr;
    std::vector<uint8_t> valid(ci->n_int, 0);
    const auto& s0 = ci->slot[0];
    const auto& s1 = ci->slot[1];
    const int n2 = ci->nsp == 2 ? ci->slot[2].nstr : 1, n3 = ci->nsp == 2 ? ci->slot[3].nstr : 1;
    for (int i = 0; i < s0.nstr; i++)
      for (int j = 0; j < s1.nstr; j++) {
        const int g0 = s0.deg[i] + s1.deg[j];
        if (g0 > tot0) continue;
        for (int k = 0; k < n2; k++)
          for (int l = 0; l < n3; l++) {
            int g1 = 0;
            if (ci->nsp == 2) {
              g1 = ci->slot[2].deg[k] + ci->slot[3].deg[l];
              if (g1 > tot1) continue;
            }
            if (g0 + g1 > d->max_collective) continue;
            int64_t ia = (int64_t)s0.ref2int[i] * ci->stride[0] + (int64_t)s1.ref2int[j] * ci->stride[1];
            if (ci->nsp == 2)
              ia += (int64_t)ci->slot[2].ref2int[k] * ci->stride[2] + (int64_t)ci->slot[3].ref2int[l] * ci->stride[3];
            addr.push_back(ia);
            valid[ia] = 1;
            ci->host_unfold.insert(ci->host_unfold.end(), {i, j, k, l});
          }
      }
    ci->n_ref = (int64_t)addr.size();
    ci->ref_addr.reset(ctx, addr.size());
    upload(ci->stream, ci->ref_addr.get(), addr.data(), addr.size() * 8);
    ci->valid.reset(ctx, valid.size());
    upload(ci->stream, ci->valid.get(), valid.data(), valid.size());
  }

  // ---- integrals and per-string tables ---------------------------------------------------
  upload_blocks(ci.get(), d);
  std::vector<DevBuf<double>> hdev(ci->nslot);
  for (int t = 0; t < ci->nslot; t++) {
    const Slot& sl = ci->slot[t];
    const double* hs = d->h[sl.species][sl.spinblk];
    PQG_REQUIRE(hs != nullptr, PQG_ERR_ARG, "missing one-body block");
    const int n = sl.n_orb;
    std::vector<double> hr((size_t)n * n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) hr[(size_t)i * n + j] = hs[i + (size_t)j * n];
    hdev[t].reset(ctx, hr.size());
    upload(ci->stream, hdev[t].get(), hr.data(), hr.size() * 8);
  }
  hipStream_t st = ci->stream;
  for (int t = 0; t < ci->nslot; t++) {
    Slot& sl = ci->slot[t];
    const double* Vtt = ci->V64[t][t].get();
    const SlotDev sd = sl.dev();
    if (sl.nsingle > 0) {
      sl.A.reset(ctx, (size_t)sl.nsingle * sl.nstr);
      k_tab_singles<<<grid1(sl.nstr, 128), 128, 0, st>>>(sd, hdev[t].get(), Vtt, sl.npair, sl.A.get());
      PQG_HIP(hipGetLastError());
    }
    if (sl.ndouble > 0) {
      sl.D.reset(ctx, (size_t)sl.ndouble * sl.nstr);
      k_tab_doubles<<<grid1(sl.nstr, 128), 128, 0, st>>>(sd, Vtt, sl.npair, sl.D.get());
      PQG_HIP(hipGetLastError());
    }
    sl.E1.reset(ctx, sl.nstr);
    k_tab_e1<<<grid1(sl.nstr, 128), 128, 0, st>>>(sd, hdev[t].get(), Vtt, sl.npair, sl.E1.get());
    PQG_HIP(hipGetLastError());
  }
  for (int t = 0; t < ci->nslot; t++)
    for (int u = 0; u < ci->nslot; u++) {
      if (u == t) continue;
      const Slot &stt = ci->slot[t], &su = ci->slot[u];
      if (su.nocc > 0 && stt.nsingle > 0) {  // spectator sums for singles of t
        const int64_t n = (int64_t)su.nstr * stt.npair;
        ci->X64[t][u].reset(ctx, n);
        k_tab_spectator<<<grid1(n, 256), 256, 0, st>>>(su.dev(), ci->V64[t][u].get(), stt.npair, su.npair,
                                                       ci->X64[t][u].get());
        PQG_HIP(hipGetLastError());
      }
      if (u > t && su.nocc > 0 && stt.nocc > 0) {  // diagonal pair terms
        const int64_t n = (int64_t)su.nstr * stt.n_orb;
        ci->DW[t][u].reset(ctx, n);
        const bool same = stt.species == su.species;
        k_tab_diagpair<<<grid1(n, 256), 256, 0, st>>>(su.dev(), stt.n_orb, ci->V64[t][u].get(),
                                                      same ? ci->V64[u][t].get() : nullptr, stt.npair, su.npair,
                                                      ci->DW[t][u].get());
        PQG_HIP(hipGetLastError());
      }
    }
  // ---- single-particle strings: dense pair matrices -----------------------------------------
  if (tensor_eligible(ci.get()) && !std::getenv("PQG_NO_TENSOR_PATH")) {
    std::vector<const double*> h1(kSlots, nullptr);
    for (int t = 0; t < ci->nslot; t++) h1[t] = hdev[t].get();
    tensor_setup(ci.get(), h1);
  }
  if (!ci->tensor.enabled) link_gemm_setup(ci.get());
  // ---- diagonal ----------------------------------------------------------------------------
  DiagArgs da{};
  da.nslot = ci->nslot;
  da.n = ci->n_int;
  da.valid = ci->fci ? nullptr : ci->valid.get();
  for (int t = 0; t < ci->nslot; t++) {
    da.sl[t] = ci->slot[t].dev();
    da.E1[t] = ci->slot[t].E1.get();
    da.stride[t] = ci->stride[t];
    for (int u = 0; u < ci->nslot; u++) da.DW[t][u] = ci->DW[t][u].get();
  }
  ci->hdiag.reset(ctx, ci->n_int);
  k_hdiag<<<grid1(ci->n_int, 256), 256, 0, st>>>(da, ci->hdiag.get());
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipStreamSynchronize(st));
  return ci.release();
}

// ---------------------------------------------------------------------------------------
// sigma driver
// ---------------------------------------------------------------------------------------
template <typename T>
static void launch_pass(pqg_ci* ci, const Tables<T>& tb, const std::vector<int>& rows, const std::vector<int>& cols,
                        const T* C, int64_t ldc, T* S, int64_t lds, int k, int flags, hipStream_t st) {
  PassArgs<T> a{};
  for (int t = 0; t < ci->nslot; t++) a.sl[t] = ci->slot[t].dev();
  a.tb = tb;
  a.nr = (int)rows.size();
  a.nc = (int)cols.size();
  int64_t nrows = 1, M = 1;
  for (int q = 0; q < a.nr; q++) {
    a.rs[q] = rows[q];
    a.rn[q] = ci->slot[rows[q]].nstr;
    nrows *= a.rn[q];
  }
  for (int c = 0; c < a.nc; c++) {
    a.cs[c] = cols[c];
    a.cn[c] = ci->slot[cols[c]].nstr;
    M *= a.cn[c];
  }
  int64_t s = 1;
  for (int c = a.nc - 1; c >= 0; --c) {
    a.cstride[c] = s;
    s *= a.cn[c];
  }
  a.M = M;
  a.C = C;
  a.ldc = ldc;
  a.S = S;
  a.lds = lds;
  a.Hd = ci->hdiag.get();
  a.flags = flags;
  // block: about 4 columns per thread for small rows, up to 1024 threads
  const int max_threads = a.nc == 1 ? pass_threads<1>() : pass_threads<3>();
  int nth = (int)std::min<int64_t>(max_threads, std::max<int64_t>(64, ((M + 3) / 4 + 63) / 64 * 64));
  // columns per thread: up to 16 for one column string; the three-string variant keeps
  // registers (no scratch) only up to 2, so it uses more column tiles instead
  const int cpt_max = a.nc == 1 ? kCPT : 2;
  if (a.nc != 1) nth = max_threads;
  int64_t tile_cols = std::min<int64_t>(M, (int64_t)nth * cpt_max);
  a.tiles = (int)((M + tile_cols - 1) / tile_cols);
  a.tile_cols = (int)tile_cols;
  a.cpt = (int)((tile_cols + nth - 1) / nth);
  // LDS layout: [rows: G x M][spectator sums][G coupling rows][rank tables][row singles]
  size_t xo = 0;
  for (int c = 0; c < a.nc; c++) {
    a.x_off[c] = (int)xo;
    a.v_off[c] = (int)xo;
    xo += ci->slot[cols[c]].npair;
  }
  a.vstride = (int)xo;
  size_t ro = 0;
  for (int c = 0; c < a.nc; c++) {
    a.rt_off[c] = (int)ro;
    const Slot& sl = ci->slot[cols[c]];
    ro += (size_t)sl.nchunk * 256 * (sl.nocc + 1);
  }
  const Slot& s0 = ci->slot[rows[0]];
  const size_t nrowsingles = (flags & F_MIXED) && a.nr == 1 ? (size_t)std::max<int64_t>(1, s0.nsingle) : 1;
  auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
  const size_t kLds = 160 * 1024 - 256;  // leaves room for the kernel's static LDS
  auto layout = [&](int G, bool stage) {
    size_t off = stage ? r16((size_t)G * M * sizeof(T)) : 0;
    a.off_x = (int)off;
    off = r16(off + xo * sizeof(T));
    a.off_v = (int)off;
    off = r16(off + (size_t)G * xo * sizeof(T));
    a.off_rt = (int)off;
    off = r16(off + ro * 4);
    a.off_rl = (int)off;
    off += nrowsingles * sizeof(int2);
    return off;
  };
  // Stage rows in LDS only while the block keeps about 16 waves per CU resident; otherwise read
  // them in place (they stay in L2 while the block works on them).  Without the row x column
  // term only the row itself is staged.
  const int waves = (nth + 63) / 64;
  const int want_blocks = std::max(1, (16 + waves - 1) / waves);
  const size_t budget = std::min(kLds, (size_t)(160 * 1024) / (size_t)want_blocks);
  const bool mixed = (flags & F_MIXED) && a.nr == 1;
  bool stage = true;
  int G = mixed ? kGMax : 1;
  while (G > 1 && layout(G, true) > budget) G--;
  if (layout(G, true) > budget || (mixed && G < 4)) {
    stage = false;
    G = mixed ? kGMax : 1;
  }
  const size_t off = layout(G, stage);
  PQG_REQUIRE(off <= kLds, PQG_ERR_UNSUPPORTED, "sigma pass tables do not fit in LDS");
  a.G = G;
  dim3 grid((unsigned)(nrows * a.tiles), (unsigned)k);
  ProfScope prof(ci->ctx, st, a.nc == 1 ? (sizeof(T) == 8 ? "sigma_pass_nc1_f64" : "sigma_pass_nc1_f32")
                                        : (sizeof(T) == 8 ? "sigma_pass_nc3_f64" : "sigma_pass_nc3_f32"),
                 k);
  int max_orb = 0;
  for (int t = 0; t < ci->nslot; t++) max_orb = std::max(max_orb, ci->slot[t].n_orb);
  const bool narrow = max_orb <= 32;
  auto go = [&](auto kern) {
    PQG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)off));
    kern<<<grid, nth, off, st>>>(a);
  };
  auto pick = [&](auto w_tag, auto stage_tag) {
    using W = decltype(w_tag);
    constexpr bool S = decltype(stage_tag)::value;
    const int cpt = a.cpt;
    if (a.nc == 1) {
      if (cpt <= 1) go(k_sigma_pass<T, W, 1, 1, S>);
      else if (cpt <= 2) go(k_sigma_pass<T, W, 1, 2, S>);
      else if (cpt <= 4) go(k_sigma_pass<T, W, 1, 4, S>);
      else if (cpt <= 8) go(k_sigma_pass<T, W, 1, 8, S>);
      else go(k_sigma_pass<T, W, 1, 16, S>);
    } else {
      if (cpt <= 1) go(k_sigma_pass<T, W, 3, 1, S>);
      else go(k_sigma_pass<T, W, 3, 2, S>);
    }
  };
  if (narrow) {
    if (stage) pick(uint32_t{}, std::true_type{});
    else pick(uint32_t{}, std::false_type{});
  } else {
    if (stage) pick(uint64_t{}, std::true_type{});
    else pick(uint64_t{}, std::false_type{});
  }
  PQG_HIP(hipGetLastError());
}

template <typename T>
static void transpose(const T* in, int64_t ldi, T* out, int64_t ldo, int64_t rows, int64_t cols, int k, bool acc,
                      hipStream_t st) {
  dim3 g((unsigned)((cols + 63) / 64), (unsigned)((rows + 63) / 64), (unsigned)k);
  k_transpose<T><<<g, 256, 0, st>>>(in, ldi, out, ldo, rows, cols, acc ? 1 : 0);
  PQG_HIP(hipGetLastError());
}

template <typename T>
static void sigma_impl(pqg_ci* ci, const Tables<T>& tb, const T* C, T* S, int k, int64_t ld, hipStream_t st) {
  ProfScope prof(ci->ctx, st, sizeof(T) == 8 ? "sigma_f64" : "sigma_f32", k);
  if (ci->xh) {
    if constexpr (std::is_same_v<T, double>) {
      const int64_t n = ci->n_int;
      k_xh_spmv<<<grid1(n, 256), 256, 0, st>>>(ci->xh_ptr.get(), ci->xh_col.get(), ci->xh_val.get(), n, C, S, k, ld);
      PQG_HIP(hipGetLastError());
      for (const auto& pen : ci->xpen) {
        if (pen.power == 1) {
          k_csr_apply<<<grid1(n, 256), 256, 0, st>>>(pen.ptr.get(), pen.col.get(), pen.val.get(), n, C, S, k, ld,
                                                      pen.alpha, 1);
        } else {
          // T1 = P C with the leading dimension of C and S (the kernel takes one for both)
          grow(ci->scratch, ci->ctx, (size_t)ld * k * sizeof(double), st);
          double* T1 = reinterpret_cast<double*>(ci->scratch.get());
          k_csr_apply<<<grid1(n, 256), 256, 0, st>>>(pen.ptr.get(), pen.col.get(), pen.val.get(), n, C, T1, k, ld,
                                                      1.0, 0);
          k_csr_apply<<<grid1(n, 256), 256, 0, st>>>(pen.ptr.get(), pen.col.get(), pen.val.get(), n, T1, S, k, ld,
                                                      pen.alpha, 1);
        }
        PQG_HIP(hipGetLastError());
      }
      return;
    } else {
      PQG_REQUIRE(false, PQG_ERR_UNSUPPORTED, "the stored explicit Hamiltonian is FP64 only (PQG_FP64)");
    }
  }
  if (ci->tensor.enabled) {
    tensor_sigma(ci, C, S, k, ld, st);
    return;
  }
  const int64_t N = ci->n_int;
  const int64_t n0 = ci->slot[0].nstr;
  const int64_t M = N / n0;
  std::vector<int> rowsA = {0}, colsA, rowsB;
  for (int t = 1; t < ci->nslot; t++) {
    colsA.push_back(t);
    rowsB.push_back(t);
  }
  if (ci->lg.enabled) {
    // at most four vectors at a time, so the work copies stay bounded for large N
    constexpr int kGroup = 4;
    const int64_t n1 = ci->slot[1].nstr;
    const int64_t nP = M / n1;  // positron strings per (alpha, beta) pair
    const int kmax = std::min(k, kGroup);
    grow(ci->scratch, ci->ctx, (nP > 1 ? 4 : 2) * (size_t)N * kmax * sizeof(T), st);
    T* CT = reinterpret_cast<T*>(ci->scratch.get());
    T* ST = CT + (size_t)N * kmax;
    T* CP = ST + (size_t)N * kmax;
    T* SP = CP + (size_t)N * kmax;
    for (int v0 = 0; v0 < k; v0 += kGroup) {
      const int kg = std::min(kGroup, k - v0);
      const T* Cg = C + (size_t)v0 * ld;
      T* Sg = S + (size_t)v0 * ld;
      // beta singles and doubles on the transpose
      transpose<T>(Cg, ld, CT, N, n0, M, kg, false, st);
      for (int v = 0; v < kg; v++) link_gemm_pass(ci, 1, CT + (size_t)v * N, ST + (size_t)v * N, false, st);
      transpose<T>(ST, N, Sg, ld, M, n0, kg, false, st);
      // everything else but the diagonal; with positrons on C permuted to (alpha, positron
      // strings, beta)
      for (int v = 0; v < kg; v++) {
        const T* Cv = Cg + (size_t)v * ld;
        T* Sv = Sg + (size_t)v * ld;
        if (nP == 1) {
          link_gemm_pass(ci, 0, Cv, Sv, true, st);
          continue;
        }
        T* Cp = CP + (size_t)v * N;
        T* Sp = SP + (size_t)v * N;
        for (int64_t z0 = 0; z0 < n0; z0 += 65535) {
          const int64_t z = std::min<int64_t>(65535, n0 - z0);
          transpose<T>(Cv + z0 * M, M, Cp + z0 * M, M, n1, nP, (int)z, false, st);
        }
        link_gemm_pass(ci, 0, Cp, Sp, false, st);
        for (int64_t z0 = 0; z0 < n0; z0 += 65535) {
          const int64_t z = std::min<int64_t>(65535, n0 - z0);
          transpose<T>(Sp + z0 * M, M, Sv + z0 * M, M, nP, n1, (int)z, true, st);
        }
      }
      link_gemm_diag(ci, Cg, ld, Sg, ld, kg, st);
    }
    return;
  }
  const bool need_b = ci->slot[0].nsingle > 0;
  // pass A: diagonal, column-side excitations, row x column doubles
  launch_pass<T>(ci, tb, rowsA, colsA, C, ld, S, ld, k, F_DIAG | F_COLS | F_MIXED, st);
  if (need_b) {
    const size_t need = 2 * (size_t)N * k * sizeof(T);
    grow(ci->scratch, ci->ctx, need, st);
    T* CT = reinterpret_cast<T*>(ci->scratch.get());
    T* ST = CT + (size_t)N * k;
    transpose<T>(C, ld, CT, N, n0, M, k, false, st);
    // pass B: excitations of the first string, on the transposed vector
    launch_pass<T>(ci, tb, rowsB, {0}, CT, N, ST, N, k, F_COLS, st);
    transpose<T>(ST, N, S, ld, M, n0, k, true, st);
  }
  mask_invalid<T>(ci, S, ld, k, st);
}

void sigma_internal(pqg_ci* ci, const double* C, double* S, int k, int64_t ld, hipStream_t st) {
  sigma_impl<double>(ci, tables64(ci), C, S, k, ld, st);
}
void sigma_internal(pqg_ci* ci, const float* C, float* S, int k, int64_t ld, hipStream_t st) {
  ensure_mixed(ci);
  sigma_impl<float>(ci, tables32(ci), C, S, k, ld, st);
}

__global__ void k_f2d(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}

// Device bytes one solve allocates (admission control of the serving loop): the string
// tables and integrals, the Davidson subspace and work blocks (as Davidson::Davidson sizes them),
// the sigma work copies of the largest block and the link-GEMM chunk.
uint64_t estimate_ci_bytes(const pqg_ci_desc* d, const pqg_dav_params* p) {
  uint64_t N = 1, tables = 0;
  int one_particle_sets = 0;
  bool positrons = d->nspecies == 2;
  for (int sp = 0; sp < d->nspecies; sp++) {
    const int n = d->n_orb[sp];
    const uint64_t np = (uint64_t)n * (n + 1) / 2;
    for (int s = 0; s < 2; s++) {
      const int k = s == 0 ? d->n_alpha[sp] : d->n_beta[sp];
      const uint64_t ns = binom(n, k);
      N *= ns;
      const uint64_t nv = n - k;
      one_particle_sets += k == 1;
      tables += ns * (8 + 4 + 8) + (uint64_t)((n + 7) / 8) * 256 * (k + 1) * 4;
      tables += ns * (uint64_t)k * nv * (12 + 8);                            // A (+FP32), links
      tables += ns * (uint64_t)(k * (k - 1) / 2) * (nv * (nv - 1) / 2) * 12;  // D (+FP32)
      tables += ns * np * 12 * 2 + ns * (uint64_t)n * (n - 1) / 2 * 4;        // spectator sums, pair map
    }
    tables += np * np * 12 * 8;  // coupling blocks
  }
  const uint64_t nev = (uint64_t)std::max(1, p->nev);
  const uint64_t ninit = std::min<uint64_t>(std::max<uint64_t>(p->nvec_init, nev), N);
  uint64_t nmax = p->nvec_max > 0 ? (uint64_t)p->nvec_max : std::max<uint64_t>(2ull * p->nvec_init, 10 * nev);
  nmax = std::min<uint64_t>(std::max(nmax, ninit + nev), std::min<uint64_t>(N, 64));
  uint64_t dav = N * 8 * (2 * nmax + 4 * nev + ninit);  // V, W, X, AX, [R T], restart block
  if (p->precision == PQG_MIXED) dav += N * 8 * nev + 2 * N * 4 * std::max(nev, ninit);
  const uint64_t kb = std::max(ninit, nev);
  // sigma work copies (FP64 is the larger): transposes (+ positron permutations) of up to four
  // vectors, the enumeration kernel's transposes of the whole block, or the pair-GEMM partials
  uint64_t sig = std::max<uint64_t>((positrons ? 4 : 2) * N * std::min<uint64_t>(kb, 4) * 8, 2 * N * kb * 8);
  if (one_particle_sets >= 2) sig = std::max<uint64_t>(sig, 16ull * 6 * N * kb * 8);
  // link-GEMM chunk: U holds rows x links x chunk columns, at most the budget's worth
  uint64_t lmax = 0;
  for (int sp = 0; sp < d->nspecies; sp++)
    for (int s = 0; s < 2; s++) {
      const int k = s == 0 ? d->n_alpha[sp] : d->n_beta[sp];
      lmax += (uint64_t)k * (d->n_orb[sp] - k);
    }
  const uint64_t other = N * 8 /* diagonal */ + std::min<uint64_t>(lg_budget_bytes((int64_t)N, 8), N * lmax * 8);
  return tables + dav + sig + other + (64ull << 20);
}

}  // namespace pqg

using namespace pqg;

static ElemArgs elem_args(pqg_ci* ci) {
  ElemArgs a{};
  a.nslot = ci->nslot;
  a.tb = tables64(ci);
  a.Hd = ci->hdiag.get();
  for (int t = 0; t < ci->nslot; t++) {
    a.sl[t] = ci->slot[t].dev();
    a.stride[t] = ci->stride[t];
  }
  return a;
}

// Fills an N x N reference-order matrix column block by column block: at most 65535 columns per
// launch (grid.y) and, for a host matrix, a device staging block of at most 512 MiB.
// launch(j0, ncols, dst, ld) fills columns j0 .. j0 + ncols - 1 into dst.
template <typename L>
static void dense_fill(pqg_ci* ci, double* M, int64_t ldm, int on_device, L&& launch) {
  const int64_t n = ci->n_ref;
  PQG_REQUIRE(ldm >= n, PQG_ERR_ARG, "leading dimension smaller than N");
  hipStream_t st = ci->stream;
  int64_t ncol = std::min<int64_t>(n, 65535);
  if (!on_device) ncol = std::max<int64_t>(1, std::min<int64_t>(ncol, (int64_t(512) << 20) / (8 * n)));
  DevBuf<double> stage;
  if (!on_device) stage.reset(ci->ctx, (size_t)n * ncol);
  for (int64_t j0 = 0; j0 < n; j0 += ncol) {
    const int64_t nc = std::min(ncol, n - j0);
    double* dst = on_device ? M + j0 * ldm : stage.get();
    launch(j0, nc, dst, on_device ? ldm : n);
    PQG_HIP(hipGetLastError());
    if (!on_device)
      PQG_HIP(hipMemcpy2DAsync(M + j0 * ldm, ldm * 8, stage.get(), n * 8, n * 8, nc, hipMemcpyDeviceToHost, st));
  }
  PQG_HIP(hipStreamSynchronize(st));
}

// Device lookups of CsrArgs (kept alive by the caller): the per-set internal -> reference maps
// of a full CI, or the inverted address table of a truncated space.
struct CsrLookups {
  DevBuf<int64_t> inv;
  DevBuf<int32_t> i2r[kSlots];
};

static CsrArgs csr_setup(pqg_ci* ci, double screening, CsrLookups& lk) {
  hipStream_t st = ci->stream;
  CsrArgs c{};
  c.e = elem_args(ci);
  c.p = perm_args(ci);
  c.screen = screening;
  if (ci->fci) {
    for (int t = 0; t < ci->nslot; t++) {
      const Slot& sl = ci->slot[t];
      lk.i2r[t].reset(ci->ctx, sl.nstr);
      PQG_HIP(hipMemcpyAsync(lk.i2r[t].get(), sl.int2ref.data(), (size_t)sl.nstr * 4, hipMemcpyHostToDevice, st));
      c.i2r[t] = lk.i2r[t].get();
    }
  } else {
    lk.inv.reset(ci->ctx, ci->n_int);
    k_fill_i64<<<grid1(ci->n_int, 256), 256, 0, st>>>(lk.inv.get(), ci->n_int, -1);
    k_inv_addr<<<grid1(ci->n_ref, 256), 256, 0, st>>>(ci->ref_addr.get(), ci->n_ref, lk.inv.get());
    PQG_HIP(hipGetLastError());
    c.inv = lk.inv.get();
  }
  return c;
}

extern "C" {

pqg_status pqg_ci_create(pqg_ctx* ctx, const pqg_ci_desc* desc, pqg_ci** out) {
  return guard([&] {
    PQG_REQUIRE(out, PQG_ERR_ARG, "out is NULL");
    *out = create_ci(ctx, desc, nullptr);
  });
}

pqg_status pqg_ci_destroy(pqg_ci* ci) {
  return guard([&] {
    if (!ci) return;
    (void)hipStreamSynchronize(ci->stream);
    delete ci;
  });
}

int64_t pqg_ci_ndets(const pqg_ci* ci) { return ci ? ci->n_ref : -1; }

int pqg_ci_sigma_path(const pqg_ci* ci) {
  if (!ci) return -1;
  if (ci->xh) return PQG_SIGMA_STORED_H;  // the order sigma_impl tests them in
  if (ci->tensor.enabled) return PQG_SIGMA_PAIR_GEMM;
  if (ci->lg.enabled) return PQG_SIGMA_LINK_GEMM;
  return PQG_SIGMA_ENUMERATION;
}

int64_t pqg_ci_nstrings(const pqg_ci* ci, int part, int spin) {
  if (!ci || part < 0 || part >= ci->nsp || spin < 0 || spin > 1) return -1;
  return ci->slot[2 * part + spin].nstr;
}

pqg_status pqg_ci_diag(pqg_ci* ci, double* hdiag) {
  return guard([&] {
    PQG_REQUIRE(ci && hdiag, PQG_ERR_ARG, "NULL argument");
    hipStream_t st = ci->stream;
    DevBuf<double> out(ci->ctx, ci->n_ref);
    to_reference(ci, ci->hdiag.get(), ci->n_int, out.get(), ci->n_ref, 1, st);
    PQG_HIP(hipMemcpyAsync(hdiag, out.get(), ci->n_ref * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  });
}

static int64_t ref_to_int_host(const pqg_ci* ci, int64_t a) {
  int t4[4] = {0, 0, 0, 0};
  if (ci->fci) {
    int64_t r = a;
    for (int t = ci->nslot - 1; t >= 0; --t) {
      t4[t] = (int)(r % ci->slot[t].nstr);
      r /= ci->slot[t].nstr;
    }
  } else {
    for (int t = 0; t < 4; t++) t4[t] = ci->host_unfold[(size_t)a * 4 + t];
  }
  int64_t ia = 0;
  for (int t = 0; t < ci->nslot; t++) ia += (int64_t)ci->slot[t].ref2int[t4[t]] * ci->stride[t];
  return ia;
}

pqg_status pqg_ci_dense_h(pqg_ci* ci, double* H, int64_t ldh, int on_device) {
  return guard([&] {
    PQG_REQUIRE(ci && H, PQG_ERR_ARG, "NULL argument");
    const ElemArgs a = elem_args(ci);
    const PermArgs p = perm_args(ci);
    const int64_t n = ci->n_ref;
    dense_fill(ci, H, ldh, on_device, [&](int64_t j0, int64_t nc, double* dst, int64_t ld) {
      k_dense_h<<<dim3(grid1(n, 256), (unsigned)nc), 256, 0, ci->stream>>>(a, p, n, j0, dst, ld);
    });
  });
}

pqg_status pqg_ci_set_explicit(pqg_ci* ci, int enable, double screening, int64_t* nnz) {
  return guard([&] {
    PQG_REQUIRE(ci, PQG_ERR_ARG, "NULL argument");
    ci->xh = false;
    ci->xpen.clear();
    ci->xh_ptr.reset(ci->ctx, 0);
    ci->xh_col.reset(ci->ctx, 0);
    ci->xh_val.reset(ci->ctx, 0);
    ci->xdiag.reset(ci->ctx, 0);
    if (nnz) *nnz = 0;
    if (!enable) return;
    PQG_REQUIRE(ci->n_int < (int64_t(1) << 31) - 1, PQG_ERR_UNSUPPORTED, "N above 2^31 - 2 (int32 indices)");
    hipStream_t st = ci->stream;
    const int64_t n = ci->n_int;
    CsrLookups lk;
    const CsrArgs c = csr_setup(ci, screening, lk);
    DevBuf<int64_t> cnt(ci->ctx, n + 1);
    ci->xh_ptr.reset(ci->ctx, n + 1);
    k_fill_i64<<<grid1(n + 1, 256), 256, 0, st>>>(cnt.get(), n + 1, 0);
    k_xh_count<<<grid1(n, 128), 128, 0, st>>>(c, n, cnt.get());
    PQG_HIP(hipGetLastError());
    size_t tb = 0;
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.get(), ci->xh_ptr.get(), (int)(n + 1), st));
    DevBuf<uint8_t> tmp(ci->ctx, std::max<size_t>(tb, 1));
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, cnt.get(), ci->xh_ptr.get(), (int)(n + 1), st));
    int64_t total = 0;
    PQG_HIP(hipMemcpyAsync(&total, ci->xh_ptr.get() + n, 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    ci->xh_col.reset(ci->ctx, std::max<int64_t>(total, 1));
    ci->xh_val.reset(ci->ctx, std::max<int64_t>(total, 1));
    k_xh_fill<<<grid1(n, 128), 128, 0, st>>>(c, n, ci->xh_ptr.get(), ci->xh_col.get(), ci->xh_val.get());
    PQG_HIP(hipGetLastError());
    ci->xdiag.reset(ci->ctx, n);
    k_xdiag_init<<<grid1(n, 256), 256, 0, st>>>(ci->xh_ptr.get(), ci->xh_col.get(), ci->xh_val.get(), ci->hdiag.get(),
                                                n, ci->xdiag.get());
    PQG_HIP(hipGetLastError());
    PQG_HIP(hipStreamSynchronize(st));
    ci->xh = true;
    if (nnz) *nnz = total;
  });
}

pqg_status pqg_ci_add_spin_penalty(pqg_ci* ci, int part, double expected, int power, double alpha) {
  return guard([&] {
    PQG_REQUIRE(ci, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(ci->xh, PQG_ERR_ARG, "the spin penalty is added to the stored explicit H: pqg_ci_set_explicit first");
    PQG_REQUIRE(part >= 0 && part < ci->nsp, PQG_ERR_ARG, "species index out of range");
    PQG_REQUIRE(power == 1 || power == 2, PQG_ERR_ARG, "power must be 1 or 2");
    hipStream_t st = ci->stream;
    const int64_t n = ci->n_int;
    const ElemArgs a = elem_args(ci);
    S2PenArgs q{};
    q.ta = 2 * part;
    q.tb = 2 * part + 1;
    q.power = power;
    q.e = expected;
    q.valid = ci->fci ? nullptr : ci->valid.get();
    pqg_ci::XPen pen;
    pen.alpha = alpha;
    pen.power = power;
    DevBuf<int64_t> cnt(ci->ctx, n + 1);
    pen.ptr.reset(ci->ctx, n + 1);
    k_fill_i64<<<grid1(n + 1, 256), 256, 0, st>>>(cnt.get(), n + 1, 0);
    k_xp_count<<<grid1(n, 256), 256, 0, st>>>(a, q, n, cnt.get());
    PQG_HIP(hipGetLastError());
    size_t tb = 0;
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.get(), pen.ptr.get(), (int)(n + 1), st));
    DevBuf<uint8_t> tmp(ci->ctx, std::max<size_t>(tb, 1));
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, cnt.get(), pen.ptr.get(), (int)(n + 1), st));
    int64_t total = 0;
    PQG_HIP(hipMemcpyAsync(&total, pen.ptr.get() + n, 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    pen.col.reset(ci->ctx, std::max<int64_t>(total, 1));
    pen.val.reset(ci->ctx, std::max<int64_t>(total, 1));
    k_xp_fill<<<grid1(n, 256), 256, 0, st>>>(a, q, n, pen.ptr.get(), pen.col.get(), pen.val.get());
    PQG_HIP(hipGetLastError());
    k_xdiag_pen<<<grid1(n, 256), 256, 0, st>>>(pen.ptr.get(), pen.col.get(), pen.val.get(), n, alpha, power,
                                               ci->xdiag.get());
    PQG_HIP(hipGetLastError());
    PQG_HIP(hipStreamSynchronize(st));
    ci->xpen.push_back(std::move(pen));
  });
}

pqg_status pqg_ci_csr(pqg_ci* ci, double screening, int32_t* rowptr, int32_t* cols, double* vals, int64_t cap,
                      int64_t* nnz) {
  return guard([&] {
    PQG_REQUIRE(ci && nnz, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(ci->n_ref < (int64_t(1) << 31) - 1, PQG_ERR_UNSUPPORTED, "N above 2^31 - 2 (int32 CSR indices)");
    hipStream_t st = ci->stream;
    const int64_t n = ci->n_ref;
    CsrLookups lk;
    const CsrArgs c = csr_setup(ci, screening, lk);
    // counts -> row pointers (exclusive scan over n + 1 entries, the last count 0)
    DevBuf<int64_t> cnt(ci->ctx, n + 1), ptr(ci->ctx, n + 1);
    k_fill_i64<<<grid1(n + 1, 256), 256, 0, st>>>(cnt.get(), n + 1, 0);
    k_csr_count<<<grid1(n, 128), 128, 0, st>>>(c, n, cnt.get());
    PQG_HIP(hipGetLastError());
    size_t tb = 0;
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.get(), ptr.get(), (int)(n + 1), st));
    DevBuf<uint8_t> tmp(ci->ctx, std::max<size_t>(tb, 1));
    PQG_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.get(), tb, cnt.get(), ptr.get(), (int)(n + 1), st));
    std::vector<int64_t> hptr(n + 1);
    PQG_HIP(hipMemcpyAsync(hptr.data(), ptr.get(), (n + 1) * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    const int64_t total = hptr[n];
    *nnz = total;
    PQG_REQUIRE(total < (int64_t(1) << 31), PQG_ERR_UNSUPPORTED, "more than 2^31 - 1 nonzeros (int32 CSR indices)");
    if (rowptr)
      for (int64_t r = 0; r <= n; r++) rowptr[r] = (int32_t)hptr[r];
    if (!cols && !vals) return;
    PQG_REQUIRE(cols && vals, PQG_ERR_ARG, "cols and vals go together");
    PQG_REQUIRE(cap >= total, PQG_ERR_ARG, "capacity below the number of nonzeros");
    if (total == 0) return;
    DevBuf<int32_t> c0(ci->ctx, total), c1(ci->ctx, total);
    DevBuf<double> v0(ci->ctx, total), v1(ci->ctx, total);
    k_csr_fill<<<grid1(n, 128), 128, 0, st>>>(c, n, ptr.get(), c0.get(), v0.get());
    PQG_HIP(hipGetLastError());
    // columns ascending within each row (the compressed row-major order of the reference)
    size_t sb = 0;
    PQG_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, sb, c0.get(), c1.get(), v0.get(), v1.get(), (int)total,
                                                        (int)n, ptr.get(), ptr.get() + 1, 0, 32, st));
    DevBuf<uint8_t> stmp(ci->ctx, std::max<size_t>(sb, 1));
    PQG_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(stmp.get(), sb, c0.get(), c1.get(), v0.get(), v1.get(),
                                                        (int)total, (int)n, ptr.get(), ptr.get() + 1, 0, 32, st));
    PQG_HIP(hipMemcpyAsync(cols, c1.get(), total * 4, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipMemcpyAsync(vals, v1.get(), total * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  });
}

pqg_status pqg_ci_dense_s2(pqg_ci* ci, int part, double expected, int power, double* S, int64_t lds, int on_device) {
  return guard([&] {
    PQG_REQUIRE(ci && S, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(part >= 0 && part < ci->nsp, PQG_ERR_ARG, "species index out of range");
    PQG_REQUIRE(power == 1 || power == 2, PQG_ERR_ARG, "power must be 1 or 2");
    const ElemArgs a = elem_args(ci);
    const PermArgs p = perm_args(ci);
    S2PenArgs q{};
    q.ta = 2 * part;
    q.tb = 2 * part + 1;
    q.power = power;
    q.e = expected;
    q.valid = ci->fci ? nullptr : ci->valid.get();
    const int64_t n = ci->n_ref;
    dense_fill(ci, S, lds, on_device, [&](int64_t j0, int64_t nc, double* dst, int64_t ld) {
      k_dense_s2<<<dim3(grid1(n, 256), (unsigned)nc), 256, 0, ci->stream>>>(a, p, q, n, j0, dst, ld);
    });
  });
}

pqg_status pqg_ci_elements(pqg_ci* ci, const int64_t* i, const int64_t* j, int64_t npairs, double* out) {
  return guard([&] {
    PQG_REQUIRE(ci && i && j && out, PQG_ERR_ARG, "NULL argument");
    if (npairs <= 0) return;
    std::vector<int64_t> ii(npairs), jj(npairs);
    for (int64_t q = 0; q < npairs; q++) {
      PQG_REQUIRE(i[q] >= 0 && i[q] < ci->n_ref && j[q] >= 0 && j[q] < ci->n_ref, PQG_ERR_ARG,
                  "determinant index out of range");
      ii[q] = ref_to_int_host(ci, i[q]);
      jj[q] = ref_to_int_host(ci, j[q]);
    }
    hipStream_t st = ci->stream;
    DevBuf<int64_t> di(ci->ctx, npairs), dj(ci->ctx, npairs);
    DevBuf<double> dout(ci->ctx, npairs);
    PQG_HIP(hipMemcpyAsync(di.get(), ii.data(), npairs * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(dj.get(), jj.data(), npairs * 8, hipMemcpyHostToDevice, st));
    const ElemArgs a = elem_args(ci);
    k_elements<<<grid1(npairs, 256), 256, 0, st>>>(a, di.get(), dj.get(), npairs, dout.get());
    PQG_HIP(hipGetLastError());
    PQG_HIP(hipMemcpyAsync(out, dout.get(), npairs * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  });
}

pqg_status pqg_ci_sigma(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S, int64_t lds, int on_device,
                        int precision) {
  return guard([&] {
    PQG_REQUIRE(ci && C && S && k >= 1, PQG_ERR_ARG, "invalid argument");
    PQG_REQUIRE(ldc >= ci->n_ref && lds >= ci->n_ref, PQG_ERR_ARG, "leading dimension smaller than N");
    PQG_REQUIRE(precision == PQG_FP64 || precision == PQG_MIXED, PQG_ERR_ARG, "unknown precision");
    hipStream_t st = ci->stream;
    const int64_t N = ci->n_int, Nr = ci->n_ref;
    DevBuf<double> cin(ci->ctx, (size_t)N * k), sout(ci->ctx, (size_t)N * k);
    const double* Cd = C;
    DevBuf<double> cref, sref;
    if (!on_device) {
      cref.reset(ci->ctx, (size_t)Nr * k);
      PQG_HIP(hipMemcpy2DAsync(cref.get(), Nr * 8, C, ldc * 8, Nr * 8, k, hipMemcpyHostToDevice, st));
      Cd = cref.get();
      ldc = Nr;
    }
    to_internal(ci, Cd, ldc, cin.get(), N, k, st);
    if (precision == PQG_FP64) {
      sigma_internal(ci, cin.get(), sout.get(), k, N, st);
    } else {
      DevBuf<float> c32(ci->ctx, (size_t)N * k), s32(ci->ctx, (size_t)N * k);
      k_d2f<<<grid1((int64_t)N * k, 256), 256, 0, st>>>(cin.get(), c32.get(), (int64_t)N * k);
      sigma_internal(ci, c32.get(), s32.get(), k, N, st);
      k_f2d<<<grid1((int64_t)N * k, 256), 256, 0, st>>>(s32.get(), sout.get(), (int64_t)N * k);
      PQG_HIP(hipStreamSynchronize(st));
    }
    if (on_device) {
      to_reference(ci, sout.get(), N, S, lds, k, st);
    } else {
      sref.reset(ci->ctx, (size_t)Nr * k);
      to_reference(ci, sout.get(), N, sref.get(), Nr, k, st);
      PQG_HIP(hipMemcpy2DAsync(S, lds * 8, sref.get(), Nr * 8, Nr * 8, k, hipMemcpyDeviceToHost, st));
    }
    PQG_HIP(hipStreamSynchronize(st));
  });
}

// ---------------------------------------------------------------------------------------
// allocation-free sigma (pqg_ci_sigma_noalloc): every scratch buffer of the call lives in this
// fixed device array.  Each GPU the library runs on holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives for that device.
// ---------------------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[PQG_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// device address of `device`'s copy of scratch (that device must be current)
char* noalloc_scratch(int device) {
  PQG_REQUIRE(device >= 0 && device < kMaxDevices, PQG_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    PQG_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    PQG_REQUIRE(ptr, PQG_ERR_HIP, "hipGetSymbolAddress returned no address for the fixed scratch");
    g_scratch_ptr[device] = static_cast<char*>(ptr);
  }
  return g_scratch_ptr[device];
}

// installs a fixed workspace on this thread and lends the handle's own scratch buffers out of
// the way for the call (they come back, unchanged, when the scope ends)
struct NoAllocScope {
  FixedScratch fs;
  pqg_ci* ci;
  DevBuf<char> own_scratch, own_u;
  DevBuf<uint32_t> own_p;
  NoAllocScope(pqg_ci* c, char* base, size_t size) : ci(c) {
    fs.base = base;
    fs.size = size;
    own_scratch = std::move(ci->scratch);
    own_u = std::move(ci->lg.U);
    own_p = std::move(ci->lg.P);
    fixed_scratch() = &fs;
  }
  ~NoAllocScope() {
    fixed_scratch() = nullptr;
    ci->scratch = std::move(own_scratch);
    ci->lg.U = std::move(own_u);
    ci->lg.P = std::move(own_p);
  }
};
}  // namespace

pqg_status pqg_ci_sigma_noalloc(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S, int64_t lds,
                                int precision) {
  return guard([&] {
    PQG_REQUIRE(ci && C && S && k >= 1, PQG_ERR_ARG, "invalid argument");
    PQG_REQUIRE(ldc >= ci->n_ref && lds >= ci->n_ref, PQG_ERR_ARG, "leading dimension smaller than N");
    PQG_REQUIRE(precision == PQG_FP64 || precision == PQG_MIXED, PQG_ERR_ARG, "unknown precision");
    PQG_REQUIRE(precision == PQG_FP64 || (ci->mixed_ready && (!ci->tensor.enabled || ci->tensor.f32_ready)),
                PQG_ERR_UNSUPPORTED,
                "PQG_MIXED without allocation needs the FP32 tables a first mixed-precision pqg_ci_sigma or "
                "pqg_ci_davidson call builds");
    hipStream_t st = ci->stream;
    PQG_HIP(hipSetDevice(ci->ctx->device));
    const int64_t N = ci->n_int;
    const size_t vec64 = (size_t)N * k * sizeof(double);
    const size_t need_min = 2 * vec64 + (precision == PQG_MIXED ? vec64 : 0);
    PQG_REQUIRE(need_min <= PQG_NOALLOC_SCRATCH_BYTES, PQG_ERR_ARG,
                "N x k too large for the allocation-free sigma: its internal-order vector copies alone exceed the " +
                    std::to_string(PQG_NOALLOC_SCRATCH_BYTES >> 20) + " MiB fixed scratch");
    NoAllocScope scope(ci, noalloc_scratch(ci->ctx->device), PQG_NOALLOC_SCRATCH_BYTES);
    double* cin = reinterpret_cast<double*>(scope.fs.take(vec64));
    double* sout = reinterpret_cast<double*>(scope.fs.take(vec64));
    to_internal(ci, C, ldc, cin, N, k, st);
    if (precision == PQG_FP64) {
      sigma_internal(ci, cin, sout, k, N, st);
    } else {
      float* c32 = reinterpret_cast<float*>(scope.fs.take((size_t)N * k * sizeof(float)));
      float* s32 = reinterpret_cast<float*>(scope.fs.take((size_t)N * k * sizeof(float)));
      k_d2f<<<grid1((int64_t)N * k, 256), 256, 0, st>>>(cin, c32, (int64_t)N * k);
      PQG_HIP(hipGetLastError());
      sigma_internal(ci, c32, s32, k, N, st);
      k_f2d<<<grid1((int64_t)N * k, 256), 256, 0, st>>>(s32, sout, (int64_t)N * k);
      PQG_HIP(hipGetLastError());
    }
    to_reference(ci, sout, N, S, lds, k, st);
    PQG_HIP(hipStreamSynchronize(st));
  });
}

pqg_status pqg_ci_dump_strings(pqg_ci* ci, int part, int spin, uint64_t* out) {
  return guard([&] {
    PQG_REQUIRE(ci && out && part >= 0 && part < ci->nsp && (spin == 0 || spin == 1), PQG_ERR_ARG, "invalid argument");
    const Slot& sl = ci->slot[2 * part + spin];
    std::vector<uint64_t> dev(sl.nstr);
    PQG_HIP(hipMemcpyAsync(dev.data(), sl.bits.get(), sl.nstr * 8, hipMemcpyDeviceToHost, ci->stream));
    PQG_HIP(hipStreamSynchronize(ci->stream));
    for (int r = 0; r < sl.nstr; r++) out[r] = dev[sl.ref2int[r]];
  });
}

pqg_status pqg_ci_dump_unfold(pqg_ci* ci, int32_t* out) {
  return guard([&] {
    PQG_REQUIRE(ci && out, PQG_ERR_ARG, "invalid argument");
    if (!ci->fci) {
      std::memcpy(out, ci->host_unfold.data(), ci->host_unfold.size() * 4);
      return;
    }
    for (int64_t a = 0; a < ci->n_ref; a++) {
      int64_t r = a;
      int t4[4] = {0, 0, 0, 0};
      for (int t = ci->nslot - 1; t >= 0; --t) {
        t4[t] = (int)(r % ci->slot[t].nstr);
        r /= ci->slot[t].nstr;
      }
      for (int t = 0; t < 4; t++) out[a * 4 + t] = t4[t];
    }
  });
}

pqg_status pqg_ci_dump_excitations(pqg_ci* ci, int which, int part, int spin, int64_t* off, int32_t* J, int8_t* phase) {
  return guard([&] {
    PQG_REQUIRE(ci && off && (which == 1 || which == 2) && part >= 0 && part < ci->nsp && (spin == 0 || spin == 1),
                PQG_ERR_ARG, "invalid argument");
    const Slot& sl = ci->slot[2 * part + spin];
    const int per = (int)(which == 1 ? sl.nsingle : sl.ndouble);
    std::vector<int32_t> hj((size_t)sl.nstr * std::max(per, 1));
    std::vector<int8_t> ho((size_t)sl.nstr * std::max(per, 1));
    if (per > 0) {
      DevBuf<int32_t> dj(ci->ctx, (size_t)sl.nstr * per);
      DevBuf<int8_t> dodd(ci->ctx, (size_t)sl.nstr * per);
      hipStream_t st = ci->stream;
      k_dump_excitations<<<grid1(sl.nstr, 128), 128, 0, st>>>(sl.dev(), which, per, dj.get(), dodd.get());
      PQG_HIP(hipGetLastError());
      PQG_HIP(hipMemcpyAsync(hj.data(), dj.get(), hj.size() * 4, hipMemcpyDeviceToHost, st));
      PQG_HIP(hipMemcpyAsync(ho.data(), dodd.get(), ho.size(), hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
    }
    // the reference skips doubles when the HF string's word is < 2 (determinant_creation.cpp:306)
    const bool skip = (which == 2 && sl.ref[0] < 2);
    int64_t pos = 0;
    std::vector<std::pair<int32_t, int8_t>> row;
    for (int r = 0; r < sl.nstr; r++) {
      off[r] = pos;
      row.clear();
      if (!skip) {
        const int I = sl.ref2int[r];
        for (int e = 0; e < per; e++) {
          const int32_t ji = hj[(size_t)I * per + e];
          if (ji >= 0) row.push_back({sl.int2ref[ji], ho[(size_t)I * per + e]});
        }
        std::sort(row.begin(), row.end());
      }
      if (J)
        for (size_t q = 0; q < row.size(); q++) {
          J[pos + q] = row[q].first;
          phase[pos + q] = row[q].second ? -1 : 1;
        }
      pos += (int64_t)row.size();
    }
    off[sl.nstr] = pos;
  });
}

}  // extern "C"
