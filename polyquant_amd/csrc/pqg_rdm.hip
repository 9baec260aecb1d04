// pqg_rdm.hip — one-particle reduced density matrix of CI vectors.
//
// create_1rdm (src/ci/detset/rdm1.cpp:6-79) for the string set t = (species, spin).  With the
// internal vector viewed as X[s][rho] (s: a string of set t, rho: the tuple of the other strings),
//   D[o][o] = sum_s [o in s] sum_rho X[s][rho]^2
//   D[h][p] = sum_s [h in s, p not in s] (-1)^(occupied orbitals of s between h and p)
//                   sum_rho X[s][rho] X[s'][rho],        s' = s - h + p.
// Every ordered pair (s -> s') is visited once; that is the reference's addition of each
// unordered determinant pair to both (hole, part) and (part, hole) (rdm1.cpp:65-66), because the
// phase of a single excitation is the same in both directions.  Determinants outside a
// truncated space hold zero in the internal vector, and strings outside a truncated string set
// have no index (str_index = -1), so both drop out as the reference's dets.find does (rdm1.cpp:44).
//
// Grouped form (multi-particle sets with at most 16 strings per group): every single excitation
// pair (s, s') shares nocc - 1 orbitals, the core K = s & s'.  The strings K + {o} (o not in K)
// form the group of K, m = n_orb - nocc + 1 members, and the group's Gram matrix
// G_K[i][j] = sum_rho X[K+o_i][rho] X[K+o_j][rho] holds every pair of the group at once:
//   D[o_i][o_j] += (-1)^(orbitals of K between o_i and o_j) G_K[i][j]      (i != j)
//   D[o_i][o_i] += G_K[i][i]
// Each pair lies in exactly one group and each (string, occupied orbital) in exactly one
// diagonal entry, so the sums are those above.  Each row is read nocc times (once per group that
// holds it) instead of L + 1 times.
//
// Kernels (HBM-bound):
//   k_slot_outer   moves set t to the outermost index (X[A][n_t][B] -> X[n_t][A][B]);
//   k_rdm_group    one workgroup per (core K, rho chunk, vector): the m(m+1)/2 Gram entries of
//                  the group, accumulated in registers;
//   k_rdm_group_reduce  one workgroup per (h, p, vector): the signed sum over the cores without
//                  h and p, in a fixed order (deterministic);
//   k_rdm_links / k_rdm_reduce  the per-string form (single-particle sets, more than 16 members
//                  per group): one workgroup per (string s, rho chunk, vector) takes the dot
//                  products of row s with itself and with the rows of its L = nocc * nvirt single
//                  excitations, then one workgroup per (h, p, vector) sums them with their phases.
#include <algorithm>
#include <cstdlib>
#include <unordered_map>
#include <vector>

#include "pqg_ci.hpp"
#include "pqg_strings.hpp"

namespace pqg {

constexpr int RDM_THREADS = 256;
constexpr int RDM_G = 8;        // excitations per pass over row s
constexpr int RDM_MAXL = 1024;  // nocc * nvirt <= 31 * 32 for one-word strings

// out[(s * A + a) * B + b] = in[(a * n + s) * B + b] for each vector (blockIdx.z)
__global__ __launch_bounds__(256) void k_slot_outer(const double* __restrict__ in, int64_t ldi, double* __restrict__ out,
                                                    int64_t A, int64_t n, int64_t B, int64_t ld) {
  const double* src = in + (int64_t)blockIdx.z * ldi;
  double* dst = out + (int64_t)blockIdx.z * ld;
  const int64_t total = A * n * B;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (int64_t)gridDim.x * 256) {
    const int64_t o = i / B, b = i - o * B;
    const int64_t s = o / A, a = o - s * A;
    dst[i] = src[(a * n + s) * B + b];
  }
}

// in[A][n] -> out[n][A] (the B = 1 case) through 64 x 64 LDS tiles
__global__ __launch_bounds__(256) void k_slot_outer_t(const double* __restrict__ in, int64_t ldi, double* __restrict__ out,
                                                      int64_t A, int64_t n, int64_t ld) {
  __shared__ double tile[64][65];
  const int64_t a0 = (int64_t)blockIdx.y * 64, s0 = (int64_t)blockIdx.x * 64;
  const double* src = in + (int64_t)blockIdx.z * ldi;
  double* dst = out + (int64_t)blockIdx.z * ld;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int i = ty; i < 64; i += 4) {
    const int64_t a = a0 + i, s = s0 + tx;
    if (a < A && s < n) tile[i][tx] = src[a * n + s];
  }
  __syncthreads();
  for (int i = ty; i < 64; i += 4) {
    const int64_t s = s0 + i, a = a0 + tx;
    if (a < A && s < n) dst[s * A + a] = tile[tx][i];
  }
}

struct RdmArgs {
  SlotDev sl;
  const double* X;  // [k][n_t][R], leading dimension ld per vector
  int64_t ld, R, chunk;
  int nchunk, L, nv;
  double* P;        // [k][nchunk][n_t][L]
  double* Pd;       // [k][nchunk][n_t]
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(RDM_THREADS) void k_rdm_links(RdmArgs a) {
  __shared__ int32_t lj[RDM_MAXL];
  __shared__ double red[RDM_THREADS / 64][RDM_G + 1];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int s = blockIdx.x, c = blockIdx.y, v = blockIdx.z;
  const int n_t = a.sl.nstr, L = a.L, nv = a.nv;
  const uint64_t I = a.sl.bits[s];
  const uint64_t vm = ~I & low_mask(a.sl.n_orb);
  // excitation e: the (e / nv)-th occupied orbital to the (e % nv)-th virtual one, the order of
  // the string's single-excitation enumeration
  for (int e = tid; e < L; e += RDM_THREADS) {
    const int h = nth_bit(I, e / nv), p = nth_bit(vm, e % nv);
    lj[e] = str_index(a.sl, a.sl.rtab, I ^ ((1ull << h) | (1ull << p)));
  }
  __syncthreads();
  const double* X = a.X + (int64_t)v * a.ld;
  const double* xs = X + (int64_t)s * a.R;
  const int64_t r0 = (int64_t)c * a.chunk, r1 = min(a.R, r0 + a.chunk);
  const int64_t slot_base = ((int64_t)v * a.nchunk + c) * n_t + s;
  const int npass = max(1, (L + RDM_G - 1) / RDM_G);
  for (int q = 0; q < npass; q++) {
    const int e0 = q * RDM_G;
    const double* y[RDM_G];
#pragma unroll
    for (int g = 0; g < RDM_G; g++) {
      const int e = e0 + g;
      const int j = e < L ? lj[e] : -1;
      y[g] = X + (int64_t)(j >= 0 ? j : s) * a.R;  // stays in bounds when the excitation is absent
    }
    double acc[RDM_G], dg = 0.0;
#pragma unroll
    for (int g = 0; g < RDM_G; g++) acc[g] = 0.0;
    for (int64_t r = r0 + tid; r < r1; r += RDM_THREADS) {
      const double x = xs[r];
      dg += x * x;
#pragma unroll
      for (int g = 0; g < RDM_G; g++) acc[g] += x * y[g][r];
    }
#pragma unroll
    for (int g = 0; g < RDM_G; g++) acc[g] = wave_sum(acc[g]);
    dg = wave_sum(dg);
    if (lane == 0) {
#pragma unroll
      for (int g = 0; g < RDM_G; g++) red[wave][g] = acc[g];
      red[wave][RDM_G] = dg;
    }
    __syncthreads();
    if (tid <= RDM_G) {
      double t = 0.0;
#pragma unroll
      for (int w = 0; w < RDM_THREADS / 64; w++) t += red[w][tid];
      if (tid < RDM_G) {
        if (e0 + tid < L) a.P[slot_base * L + e0 + tid] = lj[e0 + tid] >= 0 ? t : 0.0;
      } else if (q == 0) {
        a.Pd[slot_base] = t;
      }
    }
    __syncthreads();
  }
}

// D[v][h + p * n] for every (h, p) of set t: blockIdx.x = h + p * n, blockIdx.y = vector
__global__ __launch_bounds__(256) void k_rdm_reduce(RdmArgs a, double* __restrict__ D) {
  __shared__ double red[4];
  const int n = a.sl.n_orb, n_t = a.sl.nstr, L = a.L, nv = a.nv;
  const int h = blockIdx.x % n, p = blockIdx.x / n, v = blockIdx.y;
  const uint64_t mask = low_mask(n);
  const int lo = min(h, p), hi = max(h, p);
  double acc = 0.0;
  for (int s = threadIdx.x; s < n_t; s += 256) {
    const uint64_t I = a.sl.bits[s];
    if (!((I >> h) & 1)) continue;
    if (h == p) {
      for (int c = 0; c < a.nchunk; c++) acc += a.Pd[((int64_t)v * a.nchunk + c) * n_t + s];
    } else if (!((I >> p) & 1)) {
      const int e = __popcll(I & low_mask(h)) * nv + __popcll(~I & mask & low_mask(p));
      double t = 0.0;
      for (int c = 0; c < a.nchunk; c++) t += a.P[(((int64_t)v * a.nchunk + c) * n_t + s) * L + e];
      acc += (__popcll(I & between_mask(lo, hi)) & 1) ? -t : t;
    }
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) D[(int64_t)v * n * n + h + (int64_t)p * n] = red[0] + red[1] + red[2] + red[3];
}

constexpr int RDM_GMAX = 16;  // members per group in the grouped form

struct GroupArgs {
  const double* X;      // [k][n_t][R], leading dimension ld per vector
  const int32_t* mem;   // [ngroup][m] member string indices, -1 when absent
  const uint64_t* kb;   // [ngroup] core bits
  int64_t ld, R, chunk, ngroup;
  int nchunk, m, n;
  double* G;            // [k][nchunk][ngroup][m(m+1)/2], (i <= j) packed row by row
};

template <int MG>
__global__ __launch_bounds__(256) void k_rdm_group(GroupArgs a) {
  constexpr int T = MG * (MG + 1) / 2;
  __shared__ double red[4][T];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t K = blockIdx.x;
  const int c = blockIdx.y, v = blockIdx.z, m = a.m;
  const double* X = a.X + (int64_t)v * a.ld;
  const double* rows[MG];
  double w[MG];
#pragma unroll
  for (int i = 0; i < MG; i++) {
    const int j = i < m ? a.mem[K * m + i] : -1;
    rows[i] = X + (int64_t)(j >= 0 ? j : 0) * a.R;  // absent members read row 0, weighted 0
    w[i] = j >= 0 ? 1.0 : 0.0;
  }
  double acc[T];
#pragma unroll
  for (int q = 0; q < T; q++) acc[q] = 0.0;
  const int64_t r0 = (int64_t)c * a.chunk, r1 = min(a.R, r0 + a.chunk);
  for (int64_t r = r0 + tid; r < r1; r += 256) {
    double y[MG];
#pragma unroll
    for (int i = 0; i < MG; i++) y[i] = w[i] * rows[i][r];
    int q = 0;
#pragma unroll
    for (int i = 0; i < MG; i++)
#pragma unroll
      for (int j = i; j < MG; j++) acc[q++] += y[i] * y[j];
  }
  if constexpr (T <= 64) {
    // reduce-scatter over the wave: at offset o, a lane keeps the half of its entries selected
    // by its bit o and adds its partner's copy of them; after six steps lane q holds the wave's
    // sum of entry q (entries T..63 are zero padding)
    double v[64];
#pragma unroll
    for (int q = 0; q < 64; q++) v[q] = q < T ? acc[q] : 0.0;
#pragma unroll
    for (int s = 0; s < 6; s++) {
      const int o = 32 >> s, len = 64 >> s, h = len / 2;
      const bool hi = (lane & o) != 0;
#pragma unroll
      for (int i = 0; i < h; i++) {
        const double send = hi ? v[i] : v[i + h];
        const double keep = hi ? v[i + h] : v[i];
        v[i] = keep + __shfl_xor(send, o, 64);
      }
    }
    if (lane < T) red[wave][lane] = v[0];
  } else {
#pragma unroll
    for (int q = 0; q < T; q++) {
      const double t = wave_sum(acc[q]);
      if (lane == 0) red[wave][q] = t;
    }
  }
  __syncthreads();
  // entry (i, j) of the MG-wide packing -> (i, j) of the m-wide packing
  const int64_t base = (((int64_t)v * a.nchunk + c) * a.ngroup + K) * (m * (m + 1) / 2);
  for (int q = tid; q < T; q += 256) {
    int i = 0, rem = q;
    while (rem >= MG - i) {
      rem -= MG - i;
      i++;
    }
    const int j = i + rem;
    if (j < m) a.G[base + i * m - i * (i - 1) / 2 + (j - i)] = red[0][q] + red[1][q] + red[2][q] + red[3][q];
  }
}

// D[v][h + p * n]: blockIdx.x = h + p * n, blockIdx.y = vector
__global__ __launch_bounds__(256) void k_rdm_group_reduce(GroupArgs a, double* __restrict__ D) {
  __shared__ double red[4];
  const int n = a.n, m = a.m;
  const int h = blockIdx.x % n, p = blockIdx.x / n, v = blockIdx.y;
  const uint64_t mask = low_mask(n), hp = (1ull << h) | (1ull << p);
  const int lo = min(h, p), hi = max(h, p);
  const int T = m * (m + 1) / 2;
  double acc = 0.0;
  for (int64_t K = threadIdx.x; K < a.ngroup; K += 256) {
    const uint64_t kb = a.kb[K];
    if (kb & hp) continue;
    const int i = __popcll(~kb & mask & low_mask(lo)), j = __popcll(~kb & mask & low_mask(hi));
    const int q = i * m - i * (i - 1) / 2 + (j - i);
    double t = 0.0;
    for (int c = 0; c < a.nchunk; c++) t += a.G[(((int64_t)v * a.nchunk + c) * a.ngroup + K) * T + q];
    acc += (h != p && (__popcll(kb & between_mask(lo, hi)) & 1)) ? -t : t;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) D[(int64_t)v * n * n + h + (int64_t)p * n] = red[0] + red[1] + red[2] + red[3];
}

// Cores K (every (nocc - 1)-subset of the orbitals, ascending value) and their members; false
// when the grouped form does not apply (single-particle or empty sets, more than RDM_GMAX members,
// or tables above 2^26 entries).
static bool build_groups(pqg_ci* ci, Slot& sl) {
  if (sl.gm > 0) return true;
  if (sl.gm < 0) return false;
  const int n = sl.n_orb, kc = sl.nocc - 1, m = n - kc;
  double ng = 1.0;
  for (int i = 0; i < kc; i++) ng = ng * (n - i) / (i + 1);
  if (sl.nocc < 1 || m > RDM_GMAX || ng * m > double(1 << 26) || std::getenv("PQG_RDM_LINKS")) {
    sl.gm = -1;
    return false;
  }
  // full CI works on reference-order vectors (the reference address is then a product of the
  // reference string indices), other spaces on internal-order ones
  const std::vector<uint64_t>& order = ci->fci ? sl.ref : sl.sorted;
  std::unordered_map<uint64_t, int32_t> pos;
  pos.reserve(order.size() * 2);
  for (size_t i = 0; i < order.size(); i++) pos[order[i]] = (int32_t)i;
  std::vector<uint64_t> kb;
  std::vector<int32_t> mem;
  kb.reserve((size_t)ng);
  mem.reserve((size_t)ng * m);
  const uint64_t top = 1ull << n;
  for (uint64_t K = (1ull << kc) - 1; K < top;) {
    kb.push_back(K);
    for (int o = 0; o < n; o++) {
      if ((K >> o) & 1) continue;
      auto it = pos.find(K | (1ull << o));
      mem.push_back(it == pos.end() ? -1 : it->second);
    }
    if (K == 0) break;
    const uint64_t c = K & (~K + 1), r = K + c;  // next subset of the same size (Gosper)
    K = (((r ^ K) >> 2) / c) | r;
  }
  sl.ngroup = (int64_t)kb.size();
  sl.gbits.reset(ci->ctx, kb.size());
  upload(ci->stream, sl.gbits.get(), kb.data(), kb.size() * 8);
  sl.gmem.reset(ci->ctx, mem.size());
  upload(ci->stream, sl.gmem.get(), mem.data(), mem.size() * 4);
  sl.gm = m;
  return true;
}

template <int MG>
static void launch_group(const GroupArgs& a, int k, hipStream_t st) {
  k_rdm_group<MG><<<dim3((unsigned)a.ngroup, (unsigned)a.nchunk, (unsigned)k), 256, 0, st>>>(a);
}

static void rdm1_grouped(pqg_ci* ci, Slot& sl, const double* X, int64_t ldx, int k, double* D, hipStream_t st) {
  const int64_t N = ci->n_int, R = N / sl.nstr;
  // rho chunks: enough workgroups for the chip, and up to 8 chunks of at least 1600 columns
  // (measured at 16 orbitals 8a8b, 4 vectors: 1 chunk 5.60 ms, 4: 4.99, 8: 4.77, 12: 4.95,
  // 16: 5.71 -- more chunks put more of the repeated row reads in the MALL until the per-
  // workgroup reduction of the m(m+1)/2 entries dominates)
  int64_t nchunk = std::max<int64_t>(1, (4096 + sl.ngroup * k - 1) / (sl.ngroup * k));
  nchunk = std::max<int64_t>(nchunk, std::min<int64_t>(8, R / 1600));
  nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, R / 2048));
  if (const char* e = std::getenv("PQG_RDM_NCHUNK")) nchunk = std::max<int64_t>(1, std::min<int64_t>(std::atoll(e), R));
  const int64_t chunk = (R + nchunk - 1) / nchunk;
  nchunk = (R + chunk - 1) / chunk;
  const int m = sl.gm;
  PQG_REQUIRE(sl.ngroup < (1ll << 31), PQG_ERR_UNSUPPORTED, "too many string groups");
  DevBuf<double> G(ci->ctx, (size_t)k * nchunk * sl.ngroup * (m * (m + 1) / 2));
  GroupArgs a{};
  a.X = X;
  a.mem = sl.gmem.get();
  a.kb = sl.gbits.get();
  a.ld = ldx;
  a.R = R;
  a.chunk = chunk;
  a.ngroup = sl.ngroup;
  a.nchunk = (int)nchunk;
  a.m = m;
  a.n = sl.n_orb;
  a.G = G.get();
  PQG_REQUIRE(a.X && a.mem && a.kb && a.G, PQG_ERR_HIP, "1-RDM buffers not allocated");
  {
    ProfScope prof(ci->ctx, st, "rdm_group", k);
    if (m <= 4)
      launch_group<4>(a, k, st);
    else if (m <= 6)
      launch_group<6>(a, k, st);
    else if (m <= 8)
      launch_group<8>(a, k, st);
    else if (m <= 9)
      launch_group<9>(a, k, st);
    else if (m <= 10)
      launch_group<10>(a, k, st);
    else if (m <= 12)
      launch_group<12>(a, k, st);
    else
      launch_group<16>(a, k, st);
    PQG_HIP(hipGetLastError());
  }
  k_rdm_group_reduce<<<dim3((unsigned)(sl.n_orb * sl.n_orb), (unsigned)k), 256, 0, st>>>(a, D);
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipStreamSynchronize(st));
}

// Moves set t of k internal-order vectors (leading dimension ldc) to the outermost index:
// out (leading dimension N) = X[n_t][A][B].
void slot_outer(pqg_ci* ci, int t, const double* C, int64_t ldc, double* out, int k, hipStream_t st) {
  const int64_t N = ci->n_int, n_t = ci->slot[t].nstr, B = ci->stride[t], A = N / (n_t * B);
  if (B == 1 && (A + 63) / 64 <= 65535) {
    dim3 g((unsigned)((n_t + 63) / 64), (unsigned)((A + 63) / 64), (unsigned)k);
    k_slot_outer_t<<<g, 256, 0, st>>>(C, ldc, out, A, n_t, N);
  } else {
    const int64_t blocks = std::min<int64_t>((N + 255) / 256, 1 << 16);
    k_slot_outer<<<dim3((unsigned)blocks, 1, (unsigned)k), 256, 0, st>>>(C, ldc, out, A, n_t, B, N);
  }
  PQG_HIP(hipGetLastError());
}

// Full CI with the grouped form: the vectors stay in reference order (no boundary permutation).
static bool rdm1_reference_order(pqg_ci* ci, int t) { return ci->fci && build_groups(ci, ci->slot[t]); }

// 1-RDMs of set t for k vectors C (leading dimension ldc), written to D (device, k consecutive
// n x n column-major blocks).  C is in reference order when rdm1_reference_order(ci, t) holds,
// else in internal order.
void rdm1_vectors(pqg_ci* ci, int t, const double* C, int64_t ldc, int k, double* D, hipStream_t st) {
  Slot& sl = ci->slot[t];
  const int64_t N = ci->n_int, n_t = sl.nstr, R = N / n_t;
  const int64_t B = ci->stride[t], A = N / (n_t * B);
  const int L = sl.nocc * sl.nvirt;
  PQG_REQUIRE(L <= RDM_MAXL, PQG_ERR_UNSUPPORTED, "too many single excitations per string");
  PQG_REQUIRE(n_t < (1ll << 31), PQG_ERR_UNSUPPORTED, "string set too large");
  PQG_REQUIRE(k <= 65535, PQG_ERR_ARG, "at most 65535 vectors per call");
  DevBuf<double> xt, P, Pd;
  const double* X = C;
  int64_t ldx = ldc;
  if (A > 1) {
    xt.reset(ci->ctx, (size_t)N * k);
    if (B == 1 && (A + 63) / 64 <= 65535) {
      dim3 g((unsigned)((n_t + 63) / 64), (unsigned)((A + 63) / 64), (unsigned)k);
      k_slot_outer_t<<<g, 256, 0, st>>>(C, ldc, xt.get(), A, n_t, N);
    } else {
      const int64_t blocks = std::min<int64_t>((N + 255) / 256, 1 << 16);
      k_slot_outer<<<dim3((unsigned)blocks, 1, (unsigned)k), 256, 0, st>>>(C, ldc, xt.get(), A, n_t, B, N);
    }
    PQG_HIP(hipGetLastError());
    X = xt.get();
    ldx = N;
  }
  if (build_groups(ci, sl)) {
    rdm1_grouped(ci, sl, X, ldx, k, D, st);
    return;
  }
  // rho chunks: enough workgroups for the chip when the set has few strings
  int64_t nchunk = std::max<int64_t>(1, (4096 + n_t * k - 1) / (n_t * k));
  nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, R / 2048));
  const int64_t chunk = (R + nchunk - 1) / nchunk;
  nchunk = (R + chunk - 1) / chunk;
  P.reset(ci->ctx, (size_t)std::max(1, L) * nchunk * n_t * k);
  Pd.reset(ci->ctx, (size_t)nchunk * n_t * k);
  RdmArgs a{};
  a.sl = sl.dev();
  a.X = X;
  a.ld = ldx;
  a.R = R;
  a.chunk = chunk;
  a.nchunk = (int)nchunk;
  a.L = L;
  a.nv = std::max(1, sl.nvirt);
  a.P = P.get();
  a.Pd = Pd.get();
  PQG_REQUIRE(a.X && a.P && a.Pd && a.sl.bits && a.sl.rtab, PQG_ERR_HIP, "1-RDM buffers not allocated");
  {
    ProfScope prof(ci->ctx, st, "rdm_links", k);
    k_rdm_links<<<dim3((unsigned)n_t, (unsigned)nchunk, (unsigned)k), RDM_THREADS, 0, st>>>(a);
    PQG_HIP(hipGetLastError());
  }
  k_rdm_reduce<<<dim3((unsigned)(sl.n_orb * sl.n_orb), (unsigned)k), 256, 0, st>>>(a, D);
  PQG_HIP(hipGetLastError());
  // the scratch buffers go back to the arena when this returns; the work on them must be done
  PQG_HIP(hipStreamSynchronize(st));
}

}  // namespace pqg

using namespace pqg;

extern "C" {

pqg_status pqg_ci_rdm1(pqg_ci* ci, const double* C, int64_t ldc, int k, int part, int spin, double* D,
                       int on_device) {
  return guard([&] {
    PQG_REQUIRE(ci && C && D && k >= 1, PQG_ERR_ARG, "invalid argument");
    PQG_REQUIRE(part >= 0 && part < ci->nsp && (spin == 0 || spin == 1), PQG_ERR_ARG,
                "particle or spin index out of range");
    PQG_REQUIRE(ldc >= ci->n_ref, PQG_ERR_ARG, "leading dimension smaller than N");
    hipStream_t st = ci->stream;
    const int t = 2 * part + spin;
    const int n = ci->slot[t].n_orb;
    const int64_t N = ci->n_int, Nr = ci->n_ref;
    DevBuf<double> dd(ci->ctx, (size_t)n * n * k);
    const double* Cd = C;
    DevBuf<double> cref, cin;
    if (!on_device) {
      cref.reset(ci->ctx, (size_t)Nr * k);
      PQG_HIP(hipMemcpy2DAsync(cref.get(), Nr * 8, C, ldc * 8, Nr * 8, k, hipMemcpyHostToDevice, st));
      Cd = cref.get();
      ldc = Nr;
    }
    if (rdm1_reference_order(ci, t)) {
      rdm1_vectors(ci, t, Cd, ldc, k, dd.get(), st);
    } else {
      cin.reset(ci->ctx, (size_t)N * k);
      to_internal(ci, Cd, ldc, cin.get(), N, k, st);
      rdm1_vectors(ci, t, cin.get(), N, k, dd.get(), st);
    }
    PQG_HIP(hipMemcpyAsync(D, dd.get(), (size_t)n * n * k * 8,
                           on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  });
}

}  // extern "C"
