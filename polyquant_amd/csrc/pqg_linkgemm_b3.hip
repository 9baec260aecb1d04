// BF16x3 link GEMM for the mixed-precision sigma (see pqg_linkgemm.hip for the formulation).
#include <hip/hip_runtime.h>

#include "pqg_linkgemm.hpp"

namespace pqg {

// ---------------------------------------------------------------------------------------
// BF16x3 link GEMM (mixed precision).  Same U as k_lg_u<float>, but each product is formed
// from bf16 halves on the bf16 matrix cores: with x = x_hi + x_lo (x_hi = bf16(x),
// x_lo = bf16(x - x_hi)), a b ~ a_lo b_hi + a_hi b_lo + a_hi b_hi (the dropped a_lo b_lo is
// about 2^-16 of a b), FP32 accumulation, on v_mfma_f32_32x32x16_bf16.
//
// The vector enters as packed pairs (x_hi in the low 16 bits, x_lo in the high 16 bits), so the
// tile build copies 32-bit words and negates by flipping both sign bits.  Block = 4 waves; wave
// w owns the 32-link M block w % MB and the K part w / MB (KS parts of KQ steps of 16); the
// partial sums of the K parts meet in LDS, added in a fixed order.  The tile is 32 columns,
// stored [column][k] so that a lane's 8 consecutive k are two 16-byte LDS reads, double
// buffered: the MFMAs on tile t and the build of tile t + 1 are independent, one barrier per
// tile.  A stays in registers as bf16 halves for all tiles of a row tuple.
// ---------------------------------------------------------------------------------------
typedef __bf16 lg_bf8 __attribute__((ext_vector_type(8)));
typedef float lg_f16v __attribute__((ext_vector_type(16)));
typedef uint32_t lg_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint32_t lg_pair(float x) {
  const __bf16 hi = (__bf16)x;
  const __bf16 lo = (__bf16)(x - (float)hi);
  return (uint32_t)__builtin_bit_cast(uint16_t, hi) | ((uint32_t)__builtin_bit_cast(uint16_t, lo) << 16);
}

__global__ void k_lg_to_pairs(const float* __restrict__ x, uint32_t* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = lg_pair(x[i]);
}

// Source of A_J[ep][k] (k in k_lg_u's row order, -1 = padding) as a pointer into the integral
// tables and a factor (0 for padding, the own row and vanishing entries; the pointer is then any
// valid table entry), so that the loads of many entries can be issued together.
__device__ __forceinline__ const float* lg_a_src(const LgArgs<float>& a, const int2* lq, const int* lab,
                                                 const int* kbf, int Qp, int abp, int sp, int k, float& sc) {
  const float* base = pick3(a.Vrc, sp);
  sc = 0.f;
  if (k < 0 || k >= a.K || k == a.nin - 1) return base;
  if (k < a.nin - 1) {
    sc = 1.f;
    return base + (size_t)Qp * a.np_c + kbf[k];
  }
  const int e = k - a.nin;
  const int abe = lab[e];
  const int se = abe >> 16;
  const int2 l = lq[e];
  if (se == sp) {
    // same-set doubles: the two links share neither hole nor particle
    if (((abe ^ abp) & 0xff00) && ((abe ^ abp) & 0xff)) {
      sc = l.x < 0 ? -0.5f : 0.5f;
      return pick33(a.Vrr, sp, sp) + (size_t)Qp * pick3(a.np, sp) + (l.y >> 16);
    }
  } else if (sp < se) {
    // one excitation in each of two row sets: the path whose first link lies in the lower set
    sc = l.x < 0 ? -1.f : 1.f;
    return pick33(a.Vrr, sp, se) + (size_t)Qp * pick3(a.np, se) + (l.y >> 16);
  }
  return base;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t lg_rsrc(const void* p, uint32_t bytes) {
  // raw buffer (stride 0) of `bytes` bytes: loads past the end return 0, stores there are dropped
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}

// Tile rows of the BF16x3 kernel, in an order that keeps every group 16-byte aligned: the
// excited rows (links e < L, padded to L4 = 4 ceil(L/4)), the column set's off-diagonal pairs
// (nbo, padded to a multiple of 4), then its diagonal pairs and the own row (nin - nbo, padded);
// A maps each tile row back to k_lg_u's order.  Padding rows of A are zero, so the tile may hold
// any finite value there, and a tile column only ever reaches the same output column, so the
// columns past the chunk (never stored) need no masking either.
// HASMAP: the column set has off-diagonal pair rows (pass A); without them (pass B, nbo = 0) the
// map rows and their loads do not exist.  LDS: the staged row at offset 0 (n_c pairs and a zero
// word that the map's "no pair" entries point to), links, pair numbers, the tile buffers (first
// used to stage A), the partial sums.
template <int MB, int KS, int KQ, bool HASMAP>
__global__ __launch_bounds__(512, 1) void k_lg_b3(const LgArgs<float> a) {
  constexpr int NT = LG_B3_NT;
  constexpr int NMAP = (LG_MAXBO / 4 * NT + 255) / 256;  // 4-pair groups of the map rows per thread
  constexpr int NEXC = (LG_MAXL / 4 * NT + 255) / 256;   // 4-link groups of the excited rows
  constexpr int KP = 16 * KS * KQ;                       // tile rows
  static_assert(MB * KS == 4, "four waves per group");
  extern __shared__ __align__(16) char smem[];
  uint32_t* crow = reinterpret_cast<uint32_t*>(smem);
  int2* lq = reinterpret_cast<int2*>(smem + a.lq_off);
  int* lab = reinterpret_cast<int*>(smem + a.lab_off);
  int* kbf = reinterpret_cast<int*>(smem + a.kbf_off);
  // tile planes: [2 groups][hi, lo][NT + 1 columns][S bf16]; column NT takes the stores of idle
  // threads, so no store is conditional
  __bf16* Bt = reinterpret_cast<__bf16*>(smem + a.tile_off);
  float* Pt = reinterpret_cast<float*>(smem + a.part_off);  // [2 groups][(KS - 1) * MB][16][64]
  const int S = a.bstride;
  const int plane = (NT + 1) * S;
  const int64_t J = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int grp = w >> 2, wl = w & 3;  // the two groups take alternate tiles
  const int mw = wl % MB, kh = wl / MB;
  const int gt = tid & 255;            // thread within the group
  const int L = a.L, nin = a.nin, nbo = a.nbo, n_c = a.n_c;
  const int ndg = nin - nbo;           // diagonal-pair rows and the own row
  const int L4 = (L + 3) / 4 * 4, nbo4 = (nbo + 3) / 4 * 4;
  const int kdg = L4 + nbo4;           // first diagonal-pair row of the tile
  lg_row_links(a, J, lq, lab, tid, 512);
  for (int k = tid; k < nin - 1; k += 512) {
    const int j = k - nbo;
    kbf[k] = k < nbo ? k + offpair_hi(k) : j * (j + 1) / 2 + j;
  }
  {
    // the row C(J, :) as pairs, eight loads in flight per thread
    const uint32_t* src = a.Cp + (size_t)J * n_c;
    for (int m0 = 0; m0 < n_c; m0 += 512 * 8) {
      uint32_t v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = src[min(m0 + 512 * u + tid, n_c - 1)];
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (m0 + 512 * u + tid < n_c) crow[m0 + 512 * u + tid] = v[u];
    }
    if (tid == 0) crow[n_c] = 0u;
  }
  __syncthreads();

  // A staged in the tile region as bf16 halves [hi, lo][32 MB rows][S], entries from all threads
  // with their loads issued eight at a time; the own row (several tables) after it
  {
    __bf16* Ah = Bt;
    __bf16* Al = Bt + 32 * MB * S;
    constexpr int NA = 32 * MB * KP;
    for (int i0 = 0; i0 < NA; i0 += 512 * 8) {
      const float* src[8];
      float sc[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = min(i0 + 512 * u + tid, NA - 1);
        const int ep = idx / KP, kk = idx - ep * KP;
        int k = -1;
        if (kk < L4) {
          if (kk < L) k = nin + kk;
        } else if (kk < kdg) {
          if (kk - L4 < nbo) k = kk - L4;
        } else if (kk - kdg < ndg) {
          k = nbo + (kk - kdg);
        }
        if (ep >= L) k = -1;
        const int epc = min(ep, L - 1);
        const int abp = lab[epc];
        src[u] = lg_a_src(a, lq, lab, kbf, lq[epc].y >> 16, abp, abp >> 16, k, sc[u]);
      }
      float v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = sc[u] * *src[u];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int idx = i0 + 512 * u + tid;
        if (idx < NA) {
          const int ep = idx / KP, kk = idx - ep * KP;
          const __bf16 hi = (__bf16)v[u];
          Ah[ep * S + kk] = hi;
          Al[ep * S + kk] = (__bf16)(v[u] - (float)hi);
        }
      }
    }
    __syncthreads();
    if (tid < L) {
      const int abp = lab[tid];
      const float own = lg_own_entry(a, J, tid, lq[tid].y >> 16, abp >> 16);
      const __bf16 hi = (__bf16)own;
      Ah[tid * S + kdg + ndg - 1] = hi;
      Al[tid * S + kdg + ndg - 1] = (__bf16)(own - (float)hi);
    }
  }
  __syncthreads();
  // A fragments: lane holds A[32 mw + lane % 32][16 (kh KQ + q) + 8 (lane / 32) + e]
  lg_bf8 ahi[KQ], alo[KQ];
  {
    const __bf16* Ah = Bt + (32 * mw + (lane & 31)) * S + 8 * (lane >> 5);
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      ahi[q] = *reinterpret_cast<const lg_bf8*>(Ah + 16 * (kh * KQ + q));
      alo[q] = *reinterpret_cast<const lg_bf8*>(Ah + 32 * MB * S + 16 * (kh * KQ + q));
    }
  }
  __syncthreads();
  {
    uint32_t* z = reinterpret_cast<uint32_t*>(Bt);
    for (int i = tid; i < 2 * plane; i += 512) z[i] = 0u;  // both groups, both planes
  }
  __syncthreads();

  // per-thread work of a tile, fixed for the block: byte offsets of its loads (relative to the
  // tile's first column) and element offsets of its LDS stores (idle threads: column NT)
  //   map groups: 4 consecutive off-diagonal pairs of one column (16-byte loads of mapT)
  //   excited groups: one column of 4 consecutive links (4 loads of 4 bytes)
  //   diagonal groups: 4 consecutive diagonal-pair/own rows of one column (from crow)
  const int nmg = nbo4 / 4, nmgt = nmg * NT;
  const int nxg = L4 / 4, nxgt = nxg * NT;
  const int ndg4 = (ndg + 3) / 4, ndgt = ndg4 * NT;
  uint32_t moff[NMAP];
  int mst[NMAP];
#pragma unroll
  for (int i = 0; i < NMAP; i++) {
    const int g = min(gt + 256 * i, max(nmgt - 1, 0));
    const int col = g / max(nmg, 1), kb4 = g - col * max(nmg, 1);
    moff[i] = (uint32_t)((col * nbo4 + 4 * kb4) * 4);
    mst[i] = (gt + 256 * i < nmgt) ? col * S + L4 + 4 * kb4 : NT * S;
  }
  uint32_t xoff[NEXC][4];
  int xst[NEXC];
#pragma unroll
  for (int i = 0; i < NEXC; i++) {
    const int g = min(gt + 256 * i, nxgt - 1);
    const int col = g & (NT - 1), e4 = g / NT;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = min(4 * e4 + j, L - 1);  // rows past L: any finite value (A is zero there)
      xoff[i][j] = (uint32_t)(((lq[e].x & 0x7fffffff) * (int64_t)n_c + col) * 4);
    }
    xst[i] = (gt + 256 * i < nxgt) ? col * S + 4 * e4 : NT * S;
  }
  const bool dg = gt < ndgt;
  const int dcol = dg ? gt / ndg4 : NT, dj4 = dg ? gt - dcol * ndg4 : 0;
  const int dst = dcol * S + kdg + 4 * dj4;
  // occupancy bits of the diagonal rows of this thread's group, the own row and padding set
  const uint64_t dpad = ~0ull << (ndg - 1);

  const int64_t gbeg = (int64_t)a.c0 + (int64_t)blockIdx.y * a.gcols;
  const int64_t gend = std::min<int64_t>((int64_t)a.c0 + a.cw, gbeg + a.gcols);
  const int ntiles = gend > gbeg ? (int)((gend - gbeg + NT - 1) / NT) : 0;
  const __amdgpu_buffer_rsrc_t rC = lg_rsrc(a.Cp, a.cp_bytes);
  const __amdgpu_buffer_rsrc_t rM = lg_rsrc(a.mapT, a.mapt_bytes);
  const uint64_t* cbits = a.cbits;
  lg_u4 tm[NMAP];
  uint32_t tx[NEXC][4];
  uint64_t cbd = 0;
  auto prefetch = [&](int64_t m0) {
    if constexpr (HASMAP) {
      const int so = (int)(m0 * nbo4 * 4);
#pragma unroll
      for (int i = 0; i < NMAP; i++)
        tm[i] = __builtin_bit_cast(lg_u4, __builtin_amdgcn_raw_buffer_load_b128(rM, (int)moff[i], so, 0));
    }
    const int so = (int)(m0 * 4);
#pragma unroll
    for (int i = 0; i < NEXC; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) tx[i][j] = __builtin_amdgcn_raw_buffer_load_b32(rC, (int)xoff[i][j], so, 0);
    cbd = cbits[min<int64_t>(m0 + min(dcol, NT - 1), (int64_t)n_c - 1)];
  };
  // four pairs -> hi and lo planes (8 bytes each)
  auto put4 = [&](__bf16* T, int off, lg_u4 v) {
    uint2 h, l;
    h.x = __builtin_amdgcn_perm(v[1], v[0], 0x05040100u);
    h.y = __builtin_amdgcn_perm(v[3], v[2], 0x05040100u);
    l.x = __builtin_amdgcn_perm(v[1], v[0], 0x07060302u);
    l.y = __builtin_amdgcn_perm(v[3], v[2], 0x07060302u);
    *reinterpret_cast<uint2*>(T + off) = h;
    *reinterpret_cast<uint2*>(T + plane + off) = l;
  };
  // branch-free tile build: map words are (byte offset into the staged row | odd << 31), the
  // "no pair" word points at the zero word
  auto store = [&](int64_t m0, __bf16* T) {
    if constexpr (HASMAP) {
#pragma unroll
      for (int i = 0; i < NMAP; i++) {
        lg_u4 v;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const uint32_t x = tm[i][j];
          const uint32_t c = *reinterpret_cast<const uint32_t*>(smem + (x & 0x7fffffffu));
          v[j] = c ^ ((uint32_t)((int32_t)x >> 31) & 0x80008000u);
        }
        put4(T, mst[i], v);
      }
    }
#pragma unroll
    for (int i = 0; i < NEXC; i++) {
      lg_u4 v;
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] = tx[i][j];
      put4(T, xst[i], v);
    }
    {
      const uint32_t c = crow[(int)std::min<int64_t>(m0 + dcol, (int64_t)n_c)];
      const uint64_t bits = (cbd | dpad) >> (4 * dj4);
      lg_u4 v;
#pragma unroll
      for (int j = 0; j < 4; j++) v[j] = c & (0u - (uint32_t)((bits >> j) & 1));
      put4(T, dst, v);
    }
  };

  const bool active = 32 * mw < L;
  const bool full_rows = 32 * mw + 32 <= L;
  // U rows of this wave as a buffer; the lane's byte offset, rows 8 (r / 4) + r % 4 and the tile
  // column as a uniform offset
  const __amdgpu_buffer_rsrc_t rU =
      lg_rsrc(a.U + ((size_t)J * L + 32 * mw) * a.ldu, (uint32_t)(32 * a.ldu * 4));
  const int uoffb = ((4 * (lane >> 5)) * a.ldu + (lane & 31)) * 4;
  __bf16* Tg = Bt + 2 * grp * plane;  // this group's buffer
  float* P = Pt + grp * (KS - 1) * MB * 1024;
  // Tile i belongs to group i % 2.  Phase p: the owner of tile p builds it (then stores the output
  // of its tile p - 2); the other group runs the MFMAs of tile p - 1.  One barrier per phase.
  if (grp < ntiles) prefetch(gbeg + (int64_t)NT * grp);
  lg_f16v acc = {};
  for (int p = 0; p <= ntiles + 1; p++) {
    if ((p & 1) == grp) {
      // the tile build first, then the stores of tile p - 2: the loads a build waits for were
      // issued in this group's previous build phase, before that phase's stores, so the wait
      // never covers stores issued just before it
      if (p < ntiles) {
        store(gbeg + (int64_t)NT * p, Tg);
        if (p + 2 < ntiles) prefetch(gbeg + (int64_t)NT * (p + 2));
      }
      if (p >= 2 && active && kh == 0) {
        const int64_t m0 = gbeg + (int64_t)NT * (p - 2);
        const int cofs = (int)(m0 - a.c0) * 4;
        float v[16];
#pragma unroll
        for (int r = 0; r < 16; r++) {
          v[r] = acc[r];
#pragma unroll
          for (int h2 = 1; h2 < KS; h2++) v[r] += P[(((h2 - 1) * MB + mw) * 16 + r) * 64 + lane];
        }
        if (full_rows && gend - m0 >= NT) {
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int ro = 8 * (r >> 2) + (r & 3);
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v[r]), rU, uoffb,
                                                  ro * a.ldu * 4 + cofs, 0);
          }
        } else {
          const bool colok = m0 + (lane & 31) < gend;
#pragma unroll
          for (int r = 0; r < 16; r++) {
            const int ro = 8 * (r >> 2) + (r & 3);
            if (colok && 32 * mw + ro + 4 * (lane >> 5) < L)
              __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v[r]), rU, uoffb,
                                                    ro * a.ldu * 4 + cofs, 0);
          }
        }
      }
    } else if (p >= 1 && p <= ntiles) {
      acc = lg_f16v{};
      if (active) {
        const __bf16* T = Tg + (lane & 31) * S + 8 * (lane >> 5);
#pragma unroll
        for (int q = 0; q < KQ; q++) {
          const lg_bf8 bh = *reinterpret_cast<const lg_bf8*>(T + 16 * (kh * KQ + q));
          const lg_bf8 bl = *reinterpret_cast<const lg_bf8*>(T + plane + 16 * (kh * KQ + q));
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[q], bh, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[q], bl, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[q], bh, acc, 0, 0, 0);
        }
        if (KS > 1 && kh > 0) {
#pragma unroll
          for (int r = 0; r < 16; r++) P[(((kh - 1) * MB + mw) * 16 + r) * 64 + lane] = acc[r];
        }
      }
    }
    __syncthreads();
  }
}

template <int MB, int KS, int KQ>
static void lg_b3_go(const LgArgs<float>& a, dim3 grid, size_t lds, hipStream_t st) {
  auto kern = a.nbo > 0 ? k_lg_b3<MB, KS, KQ, true> : k_lg_b3<MB, KS, KQ, false>;
  PQG_REQUIRE(a.nbo == 0 || a.mapT, PQG_ERR_ARG, "BF16x3 link GEMM: pair rows without their map");
  PQG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid, 512, lds, st>>>(a);
}
void lg_b3_launch(const LgArgs<float>& a, const LgB3Shape& sh, dim3 grid, hipStream_t st) {
  if (sh.mb == 2) {
    switch (sh.kq) {
      case 1: lg_b3_go<2, 2, 1>(a, grid, sh.lds, st); break;
      case 2: lg_b3_go<2, 2, 2>(a, grid, sh.lds, st); break;
      case 3: lg_b3_go<2, 2, 3>(a, grid, sh.lds, st); break;
      case 4: lg_b3_go<2, 2, 4>(a, grid, sh.lds, st); break;
      case 5: lg_b3_go<2, 2, 5>(a, grid, sh.lds, st); break;
      case 6: lg_b3_go<2, 2, 6>(a, grid, sh.lds, st); break;
      default: lg_b3_go<2, 2, 7>(a, grid, sh.lds, st); break;
    }
  } else {
    switch (sh.kq) {
      case 1: lg_b3_go<1, 4, 1>(a, grid, sh.lds, st); break;
      case 2: lg_b3_go<1, 4, 2>(a, grid, sh.lds, st); break;
      case 3: lg_b3_go<1, 4, 3>(a, grid, sh.lds, st); break;
      default: lg_b3_go<1, 4, 4>(a, grid, sh.lds, st); break;
    }
  }
}


void lg_to_pairs(const float* x, uint32_t* y, int64_t n, hipStream_t st) {
  k_lg_to_pairs<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(x, y, n);
  PQG_HIP(hipGetLastError());
}

}  // namespace pqg
