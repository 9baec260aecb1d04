// Full-CI sigma on the matrix cores ("link GEMM").
//
// C is viewed as an n_r x n_c matrix: its rows are indexed by tuples of "row" strings, its
// columns by the strings of one "column" set.  Pass A works on the natural layout (rows =
// electron alpha x positron strings, columns = electron beta; with positrons on C permuted per
// alpha row to that order), pass B on the transposed vector (rows = electron beta x positron
// strings, columns = electron alpha).  A link of a string S is one single excitation
// S -> S - h + p, numbered e = a * nv + b with h the a-th occupied and p the b-th virtual
// orbital of S (ascending); Q(S, e) = idx2(h, p), s(S, e) its phase, J(S, e) the excited string
// and rev(S, e) the number of the link of J(S, e) that leads back to S.
//
// Everything but the diagonal is
//   sigma(I, m) += sum_e s(I,e) U[J(I,e)][rev(I,e)][m],   U[J][e'][m] = sum_k A_J[e'][k] T_J[k][m]
// with one GEMM per row tuple J (M = its links e', on MFMA) and K made of
//   in-row rows (gathered from the staged row C(J, :) through a map of the column set):
//     off-diagonal pairs kb of the column set (pass A):  T = s * C(J, m ^ pair kb) when exactly
//       one orbital of kb is occupied in m;  A = (Q(J,e')|kb)          -> single x single couplings
//     diagonal pairs jj of the column set:  T = n_j(m) C(J, m);  A = (Q(J,e')|jj)
//                                                      -> spectator sums of the column strings
//     the own row:  T = C(J, m);  A = single-excitation element of link e' without the column
//       spectators (one-string part + spectator sums of the other strings of the row)
//   row rows (the excited rows C(J(J,e), :) of the links e of J):
//     same set:  A = 1/2 s(J,e) (Q(J,e')|Q(J,e)) when the links share neither hole nor particle
//       -> same-set doubles (the four paths of each double add up to
//          phase * ((h0p0|h1p1) - (h0p1|h1p0)), slater_condon_singleparticle.cpp:107-175)
//     different row sets, e' in the lower one:  A = s(J,e) (Q(J,e')|Q(J,e))  -> couplings
// Each single excitation is symmetric in its two strings, so the element seen from J's link
// rev equals the reference's element of I -> J (slater_condon_singleparticle.cpp:53-105).
// The GEMM writes U for a chunk of columns; a second kernel gathers it back (pull form, fixed
// summation order, so results are bit-reproducible).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "pqg_ci.hpp"
#include "pqg_linkgemm.hpp"
#include "pqg_mfma.hpp"

namespace pqg {

// Block = 4 x KSPLIT waves.  Wave w owns "slot" w & 3 (an M block, or with NSPLIT = 2 half the
// columns of one); with KSPLIT = 2 waves w and w + 4 split the K steps in two halves whose
// partial sums meet in LDS.  The A fragments of a row tuple stay in registers for all of its
// tiles; the global loads of the next tile are issued before the MFMAs of the current one.
template <typename T, int NSPLIT, int KSPLIT, int KQH, int NT_>
__global__ __launch_bounds__(256 * KSPLIT, 3 - KSPLIT) void k_lg_u(const LgArgs<T> a) {
  constexpr int LG_THREADS = 256 * KSPLIT;
  // prefetch registers per thread: in-row map entries and excited-row values of one tile
  constexpr int NT = NT_, NF = NT / 16, NFW = NF / NSPLIT;
  constexpr int LG_TBMAX = LG_MAXBO * NT / LG_THREADS, LG_TAMAX = LG_MAXL * NT / LG_THREADS;
  constexpr int NTP = NT + 1;
  constexpr int KP = 4 * KSPLIT * KQH;  // rows of the LDS tile; rows K..KP-1 stay zero
  typedef typename Mfma<T>::acc_t acc_t;
  extern __shared__ __align__(16) char smem[];
  int2* lq = reinterpret_cast<int2*>(smem);
  int* lab = reinterpret_cast<int*>(smem + a.lab_off);
  int* kbf = reinterpret_cast<int*>(smem + a.kbf_off);
  T* crow = reinterpret_cast<T*>(smem + a.crow_off);
  T* Ts = reinterpret_cast<T*>(smem + a.tile_off);
  T* Xs = Ts;  // partial sums of the second K half, after the MFMAs of a tile
  const int64_t J = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int kh = w >> 2, slot = w & 3;
  const int L = a.L, K = a.K, nin = a.nin;
  lg_row_links(a, J, lq, lab, tid, LG_THREADS);
  for (int k = tid; k < nin - 1; k += LG_THREADS) {
    const int j = k - a.nbo;
    kbf[k] = k < a.nbo ? k + offpair_hi(k) : j * (j + 1) / 2 + j;
  }
  for (int m = tid; m < a.n_c; m += LG_THREADS) crow[m] = a.C[(size_t)J * a.n_c + m];
  for (int i = tid; i < (KP - K) * NTP; i += LG_THREADS) Ts[K * NTP + i] = T(0);
  __syncthreads();

  const int mb0 = slot / NSPLIT;
  const int fb = (slot % NSPLIT) * NFW;
  const bool active = 16 * mb0 < L;
  T af[KQH];
  {
    const int ep = 16 * mb0 + (lane & 15);
    const bool rowok = ep < L;
    const int Qp = rowok ? (lq[ep].y >> 16) : 0;
    const int abp = rowok ? lab[ep] : 0;
    const int sp = abp >> 16;
    // own row: one-string part of the single plus the other row strings' spectator sums
    const T own = rowok ? lg_own_entry(a, J, ep, Qp, sp) : T(0);
#pragma unroll
    for (int q = 0; q < KQH; q++) {
      const int k = 4 * (kh * KQH + q) + (lane >> 4);
      af[q] = rowok ? lg_a_entry(a, lq, lab, kbf, Qp, abp, sp, own, k) : T(0);
    }
  }

  const int64_t gbeg = (int64_t)a.c0 + (int64_t)blockIdx.y * a.gcols;
  const int64_t gend = std::min<int64_t>((int64_t)a.c0 + a.cw, gbeg + a.gcols);
  // Thread t holds entries t + LG_THREADS i of the off-diagonal pair rows and of the excited
  // rows, all in column t % NT; the diagonal-pair and own rows come from its column string.
  const int nbt = a.nbo * NT, nat = L * NT;
  const int ncol = tid % NT;
  int tbv[LG_TBMAX];
  T tav[LG_TAMAX];
  uint64_t cb = 0;
  // Branch-free: every slot loads from a clamped in-bounds address; which slots hold real
  // entries (and which columns lie inside the chunk) is decided when the tile is stored.
  const int64_t clast = (int64_t)a.n_c - 1;
  auto prefetch = [&](int64_t m0) {
    if (nbt > 0) {
#pragma unroll
      for (int i = 0; i < LG_TBMAX; i++) {
        const int idx = min(tid + LG_THREADS * i, nbt - 1);
        const int64_t m = min(m0 + (idx % NT), clast);
        tbv[i] = a.map[(size_t)(idx / NT) * a.n_c + m];
      }
    }
#pragma unroll
    for (int i = 0; i < LG_TAMAX; i++) {
      const int idx = min(tid + LG_THREADS * i, nat - 1);
      const int64_t m = min(m0 + (idx % NT), clast);
      tav[i] = a.C[(size_t)(lq[idx / NT].x & 0x7fffffff) * a.n_c + m];
    }
    cb = a.cbits[min(m0 + ncol, clast)];
  };
  if (gbeg < gend) prefetch(gbeg);
  for (int64_t m0 = gbeg; m0 < gend; m0 += NT) {
#pragma unroll
    for (int i = 0; i < LG_TBMAX; i++) {
      const int idx = tid + LG_THREADS * i;
      if (idx < nbt) {
        const int x = (m0 + idx % NT < gend) ? tbv[i] : -1;
        T v = T(0);
        if (x >= 0) {
          v = crow[x >> 1];
          if (x & 1) v = -v;
        }
        Ts[(idx / NT) * NTP + idx % NT] = v;
      }
    }
    {
      // diagonal pairs jj: n_j(m) C(J, m); the own row: C(J, m)
      const bool in = m0 + ncol < gend;
      const T cv = in ? crow[m0 + ncol] : T(0);
      for (int r = a.nbo + tid / NT; r < nin; r += LG_THREADS / NT) {
        const int j = r - a.nbo;
        Ts[r * NTP + ncol] = (r == nin - 1 || ((cb >> j) & 1)) ? cv : T(0);
      }
    }
#pragma unroll
    for (int i = 0; i < LG_TAMAX; i++) {
      const int idx = tid + LG_THREADS * i;
      if (idx < nat) Ts[(nin + idx / NT) * NTP + idx % NT] = (m0 + idx % NT < gend) ? tav[i] : T(0);
    }
    __syncthreads();
    if (m0 + NT < gend) prefetch(m0 + NT);
    acc_t acc[NFW];
#pragma unroll
    for (int f = 0; f < NFW; f++) acc[f] = acc_t{0, 0, 0, 0};
    if (active) {
      // every K step runs: the padding rows of the tile and of A are zero
#pragma unroll
      for (int q = 0; q < KQH; q++) {
        const T* brow = Ts + (4 * (kh * KQH + q) + (lane >> 4)) * NTP + 16 * fb + (lane & 15);
#pragma unroll
        for (int f = 0; f < NFW; f++) acc[f] = Mfma<T>::op(af[q], brow[16 * f], acc[f]);
      }
    }
    if constexpr (KSPLIT == 2) {
      __syncthreads();  // the tile is no longer read: its space takes the partial sums
      if (active && kh == 1) {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
          for (int f = 0; f < NFW; f++)
            Xs[(16 * mb0 + Mfma<T>::row(r, lane)) * NT + 16 * (fb + f) + (lane & 15)] = acc[f][r];
      }
      __syncthreads();
    }
    if (active && kh == 0) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int ep = 16 * mb0 + Mfma<T>::row(r, lane);
        if (ep >= L) continue;
        T* urow = a.U + ((size_t)J * L + ep) * a.ldu;
#pragma unroll
        for (int f = 0; f < NFW; f++) {
          const int col = 16 * (fb + f) + (lane & 15);
          const int64_t m = m0 + col;
          if (m < gend)
            urow[m - a.c0] = KSPLIT == 2 ? acc[f][r] + Xs[(16 * mb0 + Mfma<T>::row(r, lane)) * NT + col] : acc[f][r];
        }
      }
    }
    __syncthreads();
  }
}

// sigma(I, c0 + j) (+)= sum_e s(I,e) U[J(I,e)][rev(I,e)][j] for the columns j of a chunk
// sigma(I, c0 + m) (+)= sum_e s(I, e) U[J(I, e)][rev(I, e)][m]: each thread takes CPT runs of
// VEC = 16 / sizeof(T) consecutive columns, read as one 16-byte vector per U row (rows are
// 16-byte aligned); the sum over links runs in a fixed order (bit-reproducible).
template <typename T, int CPT>
__global__ __launch_bounds__(256) void k_lg_gather(const LgArgs<T> a, T* __restrict__ S) {
  constexpr int VEC = 16 / sizeof(T);
  typedef typename std::conditional<sizeof(T) == 8, double2, float4>::type vec_t;
  __shared__ int2 lq[LG_MAXL];
  __shared__ int lab[LG_MAXL];
  const int64_t I = blockIdx.x;
  lg_row_links(a, I, lq, lab, threadIdx.x, 256);
  __syncthreads();
  const int L = a.L, cw = a.cw;
  const T* __restrict__ U = a.U;
  const int v0 = blockIdx.y * 256 * CPT + threadIdx.x;  // first vector index of this thread
  T acc[CPT][VEC];
#pragma unroll
  for (int j = 0; j < CPT; j++)
#pragma unroll
    for (int q = 0; q < VEC; q++) acc[j][q] = T(0);
  for (int e = 0; e < L; e++) {
    const int2 l = lq[e];
    const T* src = U + ((size_t)(l.x & 0x7fffffff) * L + (l.y & 0xffff)) * a.ldu;
    const bool neg = l.x < 0;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      const int col = (v0 + 256 * j) * VEC;
      if (col < cw) {  // the vector may run past cw into the row padding: those lanes are dropped
        const vec_t u = *reinterpret_cast<const vec_t*>(src + col);
        const T* uv = reinterpret_cast<const T*>(&u);
#pragma unroll
        for (int q = 0; q < VEC; q++) acc[j][q] += neg ? -uv[q] : uv[q];
      }
    }
  }
  T* dst = S + (size_t)I * a.n_c + a.c0;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int col = (v0 + 256 * j) * VEC;
#pragma unroll
    for (int q = 0; q < VEC; q++)
      if (col + q < cw) dst[col + q] = a.accumulate ? dst[col + q] + acc[j][q] : acc[j][q];
  }
}

// S[v][i] += Hd[i] * C[v][i]  (the diagonal, diagonal_Hii of slater_condon.cpp:102-153)
template <typename T>
__global__ void k_lg_diag(const double* __restrict__ Hd, const T* __restrict__ C, int64_t ldc, T* __restrict__ S,
                          int64_t lds, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int v = blockIdx.y;
  S[v * lds + i] += (T)Hd[i] * C[v * ldc + i];
}

// ---------------------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------------------
static int lg_grid(int64_t n, int b) { return (int)((n + b - 1) / b); }

static uint64_t lg_binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  k = std::min(k, n - k);
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
  return (uint64_t)r;
}

// index of a string in a complete set sorted by value = its colex rank
static int64_t lg_rank(uint64_t s, const std::vector<std::vector<uint64_t>>& bn) {
  int64_t r = 0;
  int j = 1;
  for (uint64_t b = s; b; b &= b - 1, j++) r += (int64_t)bn[__builtin_ctzll(b)][j];
  return r;
}

// U scratch per chunk: about eight vectors' worth (a chunk re-stages each row of C once),
// at least 1 GiB (one chunk for small spaces), at most 16 GiB
// Bytes of the link-GEMM output block U (the columns are processed in chunks that fit it).
// Fewer, wider chunks amortise each workgroup's set-up (staged row, A): at 16 orbitals 8a8b a
// 4-vector sigma took 1022 / 910 / 860 ms (FP64) and 327 / 320 / 295 ms (BF16x3) with 5, 10.6
// and 24 GB.  PQG_LG_UBUDGET_GB overrides the size.
size_t lg_budget_bytes(int64_t N, size_t w) {
  if (const char* e = std::getenv("PQG_LG_UBUDGET_GB")) return (size_t)(std::atof(e) * (1ull << 30));
  return std::min<size_t>(24ull << 30, std::max<size_t>(1ull << 30, 32ull * (size_t)N * w));
}

// K steps per wave group: a compile-time bound so the MFMA loop has no branches.  With the
// K range split over two wave groups: 4, 8, .., 24, 26 steps each; unsplit: 8, 16, .., 48, 52.
static int lg_kqh_bound(int kq, int ksplit) {
  const int kqh = (kq + ksplit - 1) / ksplit;
  const int g = 4 * (3 - ksplit);  // step granularity
  if (kqh <= 12 * g) return (kqh + g - 1) / g * g;
  return kqh <= 13 * g ? 13 * g : 0;
}

// Wave groups splitting K: two (eight waves sharing one row tuple), except for FP32 on small
// spaces, where one group per workgroup measured faster (water/6-31G FP32 sigma: 5.2 ms vs
// 5.8 ms; 16 orbitals 8a8b: 730 ms with two groups vs 841 ms with one).
static int lg_ksplit(const pqg_ci* ci, size_t w) {
  if (const char* e = std::getenv("PQG_LG_KSPLIT")) return std::atoi(e) == 2 ? 2 : 1;
  return (w == 4 && ci->n_int < 20000000) ? 1 : 2;
}

// Pass A (natural layout; with positrons, on C permuted to (e-alpha, positron strings,
// e-beta)): row tuple = (e-alpha, positron alpha, positron beta), active = e-alpha and every
// positron set with links; columns = e-beta.  Pass B (transposed vector): row tuple =
// (e-beta, positron alpha, positron beta), active = e-beta; columns = e-alpha.
struct LgPass {
  int nrow = 0, rslot[LG_MAXSETS] = {0, 0, 0};
  int64_t rstride[LG_MAXSETS] = {1, 1, 1};
  int nsets = 0, act[LG_MAXSETS] = {0, 0, 0};
  int col = 1;
  int nbo = 0, ncd = 0, nin = 0;
  int64_t n_r = 0;
  int L = 0;
};

static int64_t positron_strings(const pqg_ci* ci) {
  return ci->nslot == 4 ? (int64_t)ci->slot[2].nstr * ci->slot[3].nstr : 1;
}

static LgPass lg_pass_desc(const pqg_ci* ci, int pass) {
  LgPass d;
  const int64_t nP = positron_strings(ci);
  const int lead = pass == 0 ? 0 : 1;
  d.col = 1 - lead;
  d.n_r = (int64_t)ci->slot[lead].nstr * nP;
  d.rslot[d.nrow] = lead;
  d.rstride[d.nrow++] = nP;
  if (ci->nslot == 4) {
    d.rslot[d.nrow] = 2;
    d.rstride[d.nrow++] = ci->slot[3].nstr;
    d.rslot[d.nrow] = 3;
    d.rstride[d.nrow++] = 1;
  }
  for (int u = 0; u < d.nrow; u++) {
    const int t = d.rslot[u];
    const int L = ci->slot[t].nocc * ci->slot[t].nvirt;
    if (L > 0 && (u == 0 || pass == 0)) {
      d.act[d.nsets++] = u;
      d.L += L;
    }
  }
  const int nc = ci->slot[d.col].n_orb;
  d.nbo = pass == 0 ? nc * (nc - 1) / 2 : 0;
  d.ncd = nc;
  d.nin = d.nbo + d.ncd + 1;
  return d;
}

static size_t lg_lds_bytes(const LgPass& d, int n_c, int kqh, int ksplit, size_t w, int nt = LG_NT) {
  auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
  const size_t ntp = nt + 1;
  size_t off = r16((size_t)d.L * 8);      // links
  off = r16(off + (size_t)d.L * 4);       // hole/particle numbers
  off = r16(off + (size_t)d.nin * 4);     // pair indices
  off = r16(off + (size_t)n_c * w);       // staged row
  // tile (also holds the 64 x nt partial sums of the second K half)
  return off + std::max<size_t>((size_t)4 * ksplit * kqh * ntp, (size_t)64 * nt) * w;
}

bool link_gemm_eligible(const pqg_ci* ci) {
  if (!ci->fci || ci->nslot < 2) return false;
  if (std::getenv("PQG_NO_LINK_GEMM")) return false;
  for (int t = 0; t < ci->nslot; t++)
    if (!ci->slot[t].complete) return false;
  for (int t = 0; t < 2; t++)
    if (ci->slot[t].nocc < 1 || ci->slot[t].nvirt < 1) return false;
  for (int pass = 0; pass < 2; pass++) {
    const LgPass d = lg_pass_desc(ci, pass);
    if (d.L > LG_MAXL || d.nin > LG_MAXIN || d.nbo > LG_MAXBO) return false;
    for (int ks = 1; ks <= 2; ks++) {
      const int kqh = lg_kqh_bound((d.nin + d.L + 3) / 4, ks);
      if (!kqh) return false;
      if (lg_lds_bytes(d, ci->slot[d.col].nstr, kqh, ks, 8) > 160 * 1024) return false;
    }
    if (d.n_r >= (1ll << 31) - 1) return false;
  }
  return true;
}

void link_gemm_setup(pqg_ci* ci) {
  LinkGemm& lg = ci->lg;
  lg.enabled = link_gemm_eligible(ci);
  if (!lg.enabled) return;
  int nmax = 0;
  for (int t = 0; t < ci->nslot; t++) nmax = std::max(nmax, ci->slot[t].n_orb);
  std::vector<std::vector<uint64_t>> bn(nmax + 1, std::vector<uint64_t>(nmax + 2, 0));
  for (int n = 0; n <= nmax; n++)
    for (int k = 0; k <= nmax + 1; k++) bn[n][k] = lg_binom(n, k);
  for (int t = 0; t < ci->nslot; t++) {
    const Slot& s = ci->slot[t];
    LinkTab& lt = lg.tab[t];
    const int n = s.n_orb, nv = s.nvirt;
    lt.L = s.nocc * nv;
    lt.nv = nv;
    lt.nb = n * (n - 1) / 2;
    if (lt.L > 0) {
      std::vector<int2> links((size_t)s.nstr * lt.L);
      for (int i = 0; i < s.nstr; i++) {
        const uint64_t S = s.sorted[i];
        int occ[64], vir[64], no = 0, nvv = 0;
        for (int o = 0; o < n; o++) {
          if ((S >> o) & 1) occ[no++] = o;
          else vir[nvv++] = o;
        }
        for (int a = 0; a < no; a++)
          for (int b = 0; b < nvv; b++) {
            const int h = occ[a], p = vir[b];
            const uint64_t Jb = S ^ ((1ull << h) | (1ull << p));
            const int lo = std::min(h, p), hi = std::max(h, p);
            const uint64_t btw = ((1ull << hi) - 1) & ~((2ull << lo) - 1);
            const int odd = __builtin_popcountll(S & btw) & 1;
            const int Q = hi * (hi + 1) / 2 + lo;
            const int ar = __builtin_popcountll(Jb & ((1ull << p) - 1));
            const int br = h - __builtin_popcountll(Jb & ((1ull << h) - 1));
            const int64_t tgt = lg_rank(Jb, bn);
            links[(size_t)i * lt.L + a * nv + b] =
                make_int2((int)((uint32_t)tgt | ((uint32_t)odd << 31)), (Q << 16) | (ar * nv + br));
          }
      }
      lt.links.reset(ci->ctx, links.size());
      upload(ci->stream, lt.links.get(), links.data(), links.size() * sizeof(int2));
    }
    if (t < 2) {
      // column-set map: e-beta (pass A) with its off-diagonal pairs, e-alpha (pass B) without;
      // then the diagonal pairs (occupied orbitals) and the own row
      const int nbo = t == 1 ? lt.nb : 0;
      const int nin = nbo + n + 1;
      lt.nin = nin;
      std::vector<int32_t> map((size_t)nin * s.nstr, -1);
      for (int p = 1; p < n && t == 1; p++)
        for (int q = 0; q < p; q++) {
          const int kb = p * (p - 1) / 2 + q;
          const uint64_t btw = ((1ull << p) - 1) & ~((2ull << q) - 1);
          for (int m = 0; m < s.nstr; m++) {
            const uint64_t S = s.sorted[m];
            if ((((S >> p) ^ (S >> q)) & 1) == 0) continue;
            const uint64_t Jb = S ^ ((1ull << p) | (1ull << q));
            const int odd = __builtin_popcountll(S & btw) & 1;
            map[(size_t)kb * s.nstr + m] = (int32_t)((lg_rank(Jb, bn) << 1) | odd);
          }
        }
      for (int m = 0; m < s.nstr; m++) {
        const uint64_t S = s.sorted[m];
        for (int j = 0; j < n; j++)
          if ((S >> j) & 1) map[(size_t)(nbo + j) * s.nstr + m] = m << 1;
        map[(size_t)(nin - 1) * s.nstr + m] = m << 1;
      }
      lt.tbm.reset(ci->ctx, map.size());
      upload(ci->stream, lt.tbm.get(), map.data(), map.size() * 4);
      if (nbo > 0) {
        // rows padded to a multiple of 4 pairs, and one tile of -1 rows past the last string
        // words: byte offset of the target pair in the kernel's staged row | odd << 31, and for
        // "no pair" the offset of the zero word after the row
        const int nbo4 = (nbo + 3) / 4 * 4;
        const uint32_t zero = (uint32_t)s.nstr * 4u;
        std::vector<uint32_t> mt((size_t)(s.nstr + LG_B3_NT) * nbo4, zero);
        for (int m = 0; m < s.nstr; m++)
          for (int kb = 0; kb < nbo; kb++) {
            const int32_t x = map[(size_t)kb * s.nstr + m];
            if (x >= 0) mt[(size_t)m * nbo4 + kb] = ((uint32_t)(x >> 1) * 4u) | ((uint32_t)(x & 1) << 31);
          }
        lt.tbmT.reset(ci->ctx, mt.size());
        upload(ci->stream, lt.tbmT.get(), mt.data(), mt.size() * 4);
      }
    }
  }
}

template <typename T, int NSPLIT, int KSPLIT, int KQH, int NT>
static void lg_go(const LgArgs<T>& a, dim3 grid, size_t lds, hipStream_t st) {
  auto kern = k_lg_u<T, NSPLIT, KSPLIT, KQH, NT>;
  PQG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid, 256 * KSPLIT, lds, st>>>(a);
}

template <typename T, int NSPLIT, int NT>
static void lg_go_k2(const LgArgs<T>& a, int kqh, dim3 grid, size_t lds, hipStream_t st) {
  switch (kqh) {
    case 4: lg_go<T, NSPLIT, 2, 4, NT>(a, grid, lds, st); break;
    case 8: lg_go<T, NSPLIT, 2, 8, NT>(a, grid, lds, st); break;
    case 12: lg_go<T, NSPLIT, 2, 12, NT>(a, grid, lds, st); break;
    case 16: lg_go<T, NSPLIT, 2, 16, NT>(a, grid, lds, st); break;
    case 20: lg_go<T, NSPLIT, 2, 20, NT>(a, grid, lds, st); break;
    case 24: lg_go<T, NSPLIT, 2, 24, NT>(a, grid, lds, st); break;
    default: lg_go<T, NSPLIT, 2, 26, NT>(a, grid, lds, st); break;
  }
}

template <typename T, int NSPLIT>
static void lg_go_k1(const LgArgs<T>& a, int kqh, dim3 grid, size_t lds, hipStream_t st) {
  switch (kqh) {
    case 8: lg_go<T, NSPLIT, 1, 8, LG_NT>(a, grid, lds, st); break;
    case 16: lg_go<T, NSPLIT, 1, 16, LG_NT>(a, grid, lds, st); break;
    case 24: lg_go<T, NSPLIT, 1, 24, LG_NT>(a, grid, lds, st); break;
    case 32: lg_go<T, NSPLIT, 1, 32, LG_NT>(a, grid, lds, st); break;
    case 40: lg_go<T, NSPLIT, 1, 40, LG_NT>(a, grid, lds, st); break;
    case 48: lg_go<T, NSPLIT, 1, 48, LG_NT>(a, grid, lds, st); break;
    default: lg_go<T, NSPLIT, 1, 52, LG_NT>(a, grid, lds, st); break;
  }
}

// BF16x3 kernel shape for a pass: MB 32-link M blocks, KS K parts of KQ steps of 16 (0: none)
static LgB3Shape lg_b3_shape(const LgPass& d, int n_c) {
  LgB3Shape sh;
  if (d.L <= 0 || d.L > 64) return sh;
  sh.mb = d.L > 32 ? 2 : 1;
  sh.ks = 4 / sh.mb;
  // tile rows: excited rows, off-diagonal pairs, diagonal pairs + own row, each padded to 4
  const int ndg = d.nin - d.nbo;
  if (ndg > 32 || d.nbo > LG_MAXBO) return LgB3Shape{};  // one diagonal group per thread
  const int K = (d.L + 3) / 4 * 4 + (d.nbo + 3) / 4 * 4 + (ndg + 3) / 4 * 4;
  sh.kq = (K + 16 * sh.ks - 1) / (16 * sh.ks);
  if (sh.kq > (sh.ks == 2 ? 7 : 4)) return LgB3Shape{};
  sh.stride = 16 * sh.ks * sh.kq + 8;  // bf16 elements per tile column (16-byte rows)
  auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
  // the staged row (n_c pairs and a zero word) at offset 0, then links, pair numbers, tiles
  size_t off = r16((size_t)(n_c + 1) * 4);
  sh.lq_off = (int)off;
  off = r16(off + (size_t)d.L * 8);
  sh.lab_off = (int)off;
  off = r16(off + (size_t)d.L * 4);
  sh.kbf_off = (int)off;
  off = r16(off + (size_t)d.nin * 4);
  sh.tile_off = (int)off;
  off = r16(off + (size_t)2 * 2 * (LG_B3_NT + 1) * sh.stride * 2);  // 2 groups x (hi, lo) planes
  sh.part_off = (int)off;
  off += (size_t)2 * (sh.ks - 1) * sh.mb * 1024 * 4;
  sh.lds = off;
  if (sh.lds > 160 * 1024) return LgB3Shape{};
  return sh;
}

// mixed precision takes the BF16x3 kernel where its LDS fits and its 32-bit load offsets reach
// the whole vector (PQG_LG_FP32=1: the FP32 MFMA kernel)
static bool lg_use_b3(const LgPass& d, int n_c) {
  static const bool fp32 = std::getenv("PQG_LG_FP32") != nullptr;
  return !fp32 && lg_b3_shape(d, n_c).mb > 0 && ((int64_t)d.n_r * n_c + LG_B3_NT) * 4 < (int64_t(1) << 32);
}

// the 64-column tile exists only for the kernel with the K split over two wave groups
template <typename T, int NSPLIT, int NT>
static void lg_go_kq(const LgArgs<T>& a, int ksplit, int kqh, dim3 grid, size_t lds, hipStream_t st) {
  if constexpr (NT != LG_NT) {
    PQG_REQUIRE(ksplit == 2, PQG_ERR_UNSUPPORTED, "64-column link-GEMM tile without the K split");
    lg_go_k2<T, NSPLIT, NT>(a, kqh, grid, lds, st);
  } else {
    if (ksplit == 2)
      lg_go_k2<T, NSPLIT, NT>(a, kqh, grid, lds, st);
    else
      lg_go_k1<T, NSPLIT>(a, kqh, grid, lds, st);
  }
}

static const double* vblock(pqg_ci* ci, int t, int u, double) { return ci->V64[t][u].get(); }
static const float* vblock(pqg_ci* ci, int t, int u, float) {
  return t == u ? ci->lg.V32s[t].get() : ci->V32[t][u].get();
}
static const double* aone(pqg_ci* ci, int t, double) { return ci->slot[t].A.get(); }
static const float* aone(pqg_ci* ci, int t, float) { return ci->slot[t].A32.get(); }
static const double* xblock(pqg_ci* ci, int t, int u, double) { return ci->X64[t][u].get(); }
static const float* xblock(pqg_ci* ci, int t, int u, float) { return ci->X32[t][u].get(); }

// U bytes per column of a pass, and the chunk width that keeps U within its budget
static size_t lg_per_col(const LgPass& d, size_t w) { return (size_t)d.n_r * d.L * w; }
// share: the part of the budget one output block takes (1, or 1/2 with two blocks)
static int lg_chunk(const pqg_ci* ci, const LgPass& d, int n_c, size_t w, int nt, double share = 1.0) {
  const size_t per_col = lg_per_col(d, w);
  const size_t want = std::min<size_t>((size_t)(share * (double)lg_budget_bytes(ci->n_int, w)),
                                       per_col * (size_t)n_c);
  const size_t budget = std::max<size_t>(per_col * nt, want);
  int cw = (int)std::min<size_t>((size_t)n_c, budget / per_col);
  if (cw < n_c) {
    // as many chunks as the budget needs, of balanced width (no narrow last chunk)
    cw = std::max(nt, cw / nt * nt);
    const int nch = (n_c + cw - 1) / cw;
    cw = std::min(cw, ((n_c + nch - 1) / nch + nt - 1) / nt * nt);
  }
  return cw;
}

// tile width of a pass: 64 columns with the K split over two wave groups when the tile and the
// staged row fit LDS (FP32 always, FP64 on spaces of 2e7 determinants and more), else 32.
// Measured, FP32 sigma of 4 vectors: 16 orbitals 8a8b 675 ms with 64 columns, 726 ms with 32;
// water 6-31G, which takes the 256-thread kernel, 8.3 ms with 64 and 5.1 ms with 32.  FP64: at
// 16 orbitals 8a8b only pass B fits (K = 81), 831 ms against 862 ms; water 6-31G 8.2 ms against
// 8.0 ms.  PQG_LG_NT=32|64 overrides (64 only where it fits).
static int lg_nt(const LgPass& d, int n_c, size_t w, int ksplit, int kqh) {
  if (ksplit != 2) return LG_NT;
  const bool fits = lg_lds_bytes(d, n_c, kqh, ksplit, w, LG_NT32) <= 160 * 1024;
  if (const char* e = std::getenv("PQG_LG_NT")) return std::atoi(e) == 64 && fits ? LG_NT32 : LG_NT;
  return fits && (w == 4 || d.n_r * n_c >= 20000000) ? LG_NT32 : LG_NT;
}

static bool lg_overlap_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("PQG_LG_OVERLAP");
    return !e || std::atoi(e) != 0;
  }();
  return on;
}

static bool lg_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  PQG_HIP(hipStreamIsCapturing(st, &cs));
  return cs != hipStreamCaptureStatusNone;
}

// the side stream of the overlapped gathers and its events, once per handle
static void lg_side_init(LinkGemm& lg) {
  if (lg.side) return;
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
  for (auto& e : lg.ev)
    if (!e) PQG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  PQG_HIP(hipStreamCreateWithFlags(&lg.side, hipStreamNonBlocking));
}

LinkGemm::~LinkGemm() {
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
  if (side) {
    (void)hipStreamSynchronize(side);
    (void)hipStreamDestroy(side);
  }
  for (auto e : ev)
    if (e) (void)hipEventDestroy(e);
}

// One pass of the link GEMM for one vector; accumulate: add to S or overwrite it.
template <typename T>
static void lg_pass(pqg_ci* ci, int pass, const T* C, T* S, bool accumulate, hipStream_t st) {
  LinkGemm& lg = ci->lg;
  const LgPass d = lg_pass_desc(ci, pass);
  PQG_REQUIRE(d.L > 0, PQG_ERR_UNSUPPORTED, "link GEMM pass without links");
  const Slot& sc = ci->slot[d.col];
  LgArgs<T> a{};
  a.C = C;
  a.nrow = d.nrow;
  for (int u = 0; u < d.nrow; u++) {
    a.rstride[u] = d.rstride[u];
    a.rns[u] = ci->slot[d.rslot[u]].nstr;
  }
  a.nsets = d.nsets;
  int off = 0;
  for (int q = 0; q < d.nsets; q++) {
    const int t = d.rslot[d.act[q]];
    a.act[q] = d.act[q];
    a.links[q] = lg.tab[t].links.get();
    a.Ls[q] = lg.tab[t].L;
    a.nvs[q] = lg.tab[t].nv;
    a.loff[q] = off;
    a.np[q] = ci->slot[t].npair;
    a.Vrc[q] = vblock(ci, t, d.col, T{});
    a.Aone[q] = aone(ci, t, T{});
    for (int q2 = 0; q2 < d.nsets; q2++) a.Vrr[q][q2] = vblock(ci, t, d.rslot[d.act[q2]], T{});
    for (int u = 0; u < d.nrow; u++) a.Xrow[q][u] = d.rslot[u] == t ? nullptr : xblock(ci, t, d.rslot[u], T{});
    off += lg.tab[t].L;
  }
  a.map = lg.tab[d.col].tbm.get();
  a.cbits = sc.bits.get();
  a.nbo = d.nbo;
  a.ncd = d.ncd;
  a.nin = d.nin;
  a.np_c = sc.npair;
  a.n_r = (int)d.n_r;
  a.n_c = sc.nstr;
  a.L = d.L;
  a.K = d.nin + d.L;
  a.accumulate = accumulate ? 1 : 0;
  const int ksplit = lg_ksplit(ci, sizeof(T));
  const int kqh = lg_kqh_bound((a.K + 3) / 4, ksplit);
  auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
  a.lab_off = (int)r16((size_t)a.L * 8);
  a.kbf_off = (int)r16(a.lab_off + (size_t)a.L * 4);
  a.crow_off = (int)r16(a.kbf_off + (size_t)a.nin * 4);
  a.tile_off = (int)r16(a.crow_off + (size_t)a.n_c * sizeof(T));
  const bool b3 = std::is_same_v<T, float> && lg_use_b3(d, a.n_c);
  const int nt = b3 ? LG_B3_NT : lg_nt(d, a.n_c, sizeof(T), ksplit, kqh);
  const size_t lds = lg_lds_bytes(d, a.n_c, kqh, ksplit, sizeof(T), nt);
  LgB3Shape sh;
  if (b3) {
    // the vector as bf16 pairs, and the BF16x3 kernel's LDS layout
    sh = lg_b3_shape(d, a.n_c);
    const int64_t nv = (int64_t)d.n_r * a.n_c;
    grow(lg.P, ci->ctx, (size_t)nv + LG_B3_NT, st);  // tile loads may run one tile past the end
    lg_to_pairs(reinterpret_cast<const float*>(C), lg.P.get(), nv, st);
    a.Cp = lg.P.get();
    a.cp_bytes = (uint32_t)(((size_t)nv + LG_B3_NT) * 4);
    a.mapT = lg.tab[d.col].tbmT.get();
    a.mapt_bytes = (uint32_t)(lg.tab[d.col].tbmT.n * 4);
    a.lq_off = sh.lq_off;
    a.lab_off = sh.lab_off;
    a.kbf_off = sh.kbf_off;
    a.tile_off = sh.tile_off;
    a.part_off = sh.part_off;
    a.bstride = sh.stride;
  }
  // FP32: two output blocks when the columns take more than one chunk; the gather of chunk i
  // (HBM-bound) runs on a side stream while the GEMM of chunk i + 1 writes the other block.  Both
  // blocks are halves of the one U buffer (half the budget each), so the overlap needs no memory
  // beyond what one block took (the serving loop's admission estimate counts one).  Not under
  // graph capture or in the allocation-free call; PQG_LG_OVERLAP=0 turns it off.  Measured at 16
  // orbitals 8a8b, sigma of 4 vectors: BF16x3 288.9 ms against 296.6 ms; FP64 858 ms against
  // 864 ms, the FP64 GEMM slowing by what the hidden gathers save (five chunks instead of four,
  // shared bandwidth), so FP64 keeps one stream.
  bool overlap = sizeof(T) == 4 && !fixed_scratch() && lg_overlap_enabled() &&
                 lg_chunk(ci, d, a.n_c, sizeof(T), nt) < a.n_c && !lg_capturing(st);
  const double share = overlap ? 0.5 : 1.0;
  // U holds n_r * L * ldu values (ldu = cw rounded up to 4); sized for the larger of the two
  // passes so it grows once
  int cw = lg_chunk(ci, d, a.n_c, sizeof(T), nt, share);
  size_t need = lg_per_col(d, sizeof(T)) * (size_t)(cw + 3);
  {
    const LgPass o = lg_pass_desc(ci, 1 - pass);
    const int n_co = ci->slot[o.col].nstr;
    const int kqo = lg_kqh_bound((o.nin + o.L + 3) / 4, ksplit);
    need = std::max(need, lg_per_col(o, sizeof(T)) * (size_t)(lg_chunk(ci, o, n_co, sizeof(T),
                                                                        lg_nt(o, n_co, sizeof(T), ksplit, kqo),
                                                                        share) + 3));
  }
  size_t half = 0;  // byte offset of the second block in U
  if (overlap) {
    need = (need + 255) / 256 * 256;
    grow(lg.U, ci->ctx, 2 * need, st);
    half = lg.U.n / 2 / 256 * 256;  // >= need: U holds at least two blocks
    // when U is larger (an FP64 pass grew it), wider chunks: as wide as half of U holds, at most
    // the one-block width, balanced as lg_chunk balances them
    const int64_t fit = (int64_t)(half / lg_per_col(d, sizeof(T))) - 3;
    const int cw1 = lg_chunk(ci, d, a.n_c, sizeof(T), nt);
    if (fit > cw) {
      int w = (int)std::min<int64_t>(fit, cw1);
      w = std::max(nt, w / nt * nt);
      const int nch = (a.n_c + w - 1) / w;
      cw = std::max(cw, std::min(w, ((a.n_c + nch - 1) / nch + nt - 1) / nt * nt));
    }
    lg_side_init(lg);
  } else if (FixedScratch* fs = fixed_scratch()) {
    // allocation-free call: U is a view of what the fixed workspace has left, and the column
    // chunk shrinks to fit it (the per-element sums do not depend on the chunking)
    if (!lg.U.get()) lg.U.view(fs, std::min(need, fs->left() > 256 ? fs->left() - 256 : 0));
    const int64_t fit = (int64_t)(lg.U.n / lg_per_col(d, sizeof(T))) - 3;
    if (fit < cw) cw = (int)(fit / nt * nt);
    PQG_REQUIRE(cw >= nt || cw >= a.n_c, PQG_ERR_ARG,
                "the fixed scratch of the allocation-free entry point cannot hold one link-GEMM column tile");
  } else {
    grow(lg.U, ci->ctx, need, st);
  }
  a.U = reinterpret_cast<T*>(lg.U.get());
  const int mb = (a.L + 15) / 16;
  // algorithmic flops of the GEMM as posed (M = L links, K = in-row rows + links, no padding)
  const int64_t flops_per_col = 2ll * a.n_r * a.L * a.K;
  // with two blocks the handle's stream waits for the side stream's last gather when the pass
  // ends (work that follows on it sees every gather); on an error the side stream is drained
  int chunk = 0;
  try {
  for (int c0 = 0; c0 < a.n_c; c0 += cw, chunk++) {
    const int blk = chunk & 1;
    hipStream_t gst = st;
    if (overlap) {
      a.U = reinterpret_cast<T*>(lg.U.get() + (blk ? half : 0));
      // the block is free once the gather of chunk - 2 has read it
      if (chunk >= 2) PQG_HIP(hipStreamWaitEvent(st, lg.ev[2 + blk], 0));
      gst = lg.side;
    }
    a.c0 = c0;
    a.cw = std::min(cw, a.n_c - c0);
    a.ldu = (a.cw + 3) / 4 * 4;
    const int tiles = lg_grid(a.cw, nt);
    const int gy = std::max(1, std::min(tiles, lg_grid(2048, a.n_r)));
    a.gcols = lg_grid(tiles, gy) * nt;
    dim3 grid((unsigned)a.n_r, (unsigned)lg_grid(a.cw, a.gcols));
    {
      ProfScope prof(ci->ctx, st, sizeof(T) == 8 ? "lg_gemm_f64" : "lg_gemm_f32", flops_per_col * a.cw);
      if constexpr (sizeof(T) == 4) {
        if (b3) {
          lg_b3_launch(reinterpret_cast<const LgArgs<float>&>(a), sh, grid, st);
        } else if (nt == LG_NT32) {
          if (mb <= 2) lg_go_kq<T, 2, LG_NT32>(a, ksplit, kqh, grid, lds, st);
          else lg_go_kq<T, 1, LG_NT32>(a, ksplit, kqh, grid, lds, st);
        } else {
          if (mb <= 2) lg_go_kq<T, 2, LG_NT>(a, ksplit, kqh, grid, lds, st);
          else lg_go_kq<T, 1, LG_NT>(a, ksplit, kqh, grid, lds, st);
        }
      } else if (nt == LG_NT32) {
        if (mb <= 2) lg_go_kq<T, 2, LG_NT32>(a, ksplit, kqh, grid, lds, st);
        else lg_go_kq<T, 1, LG_NT32>(a, ksplit, kqh, grid, lds, st);
      } else {
        if (mb <= 2) lg_go_kq<T, 2, LG_NT>(a, ksplit, kqh, grid, lds, st);
        else lg_go_kq<T, 1, LG_NT>(a, ksplit, kqh, grid, lds, st);
      }
      PQG_HIP(hipGetLastError());
    }
    if (overlap) {
      PQG_HIP(hipEventRecord(lg.ev[blk], st));
      PQG_HIP(hipStreamWaitEvent(gst, lg.ev[blk], 0));
    }
    {
      ProfScope prof(ci->ctx, gst, sizeof(T) == 8 ? "lg_gather_f64" : "lg_gather_f32",
                     (int64_t)a.n_r * a.cw * (a.L + 2) * (int64_t)sizeof(T));
      constexpr int CPT = 4;
      dim3 g2((unsigned)a.n_r, (unsigned)lg_grid(a.cw, 256 * CPT * (16 / (int)sizeof(T))));
      k_lg_gather<T, CPT><<<g2, 256, 0, gst>>>(a, S);
      PQG_HIP(hipGetLastError());
    }
    if (overlap) PQG_HIP(hipEventRecord(lg.ev[2 + blk], gst));
  }
  if (overlap) PQG_HIP(hipStreamWaitEvent(st, lg.ev[2 + ((chunk - 1) & 1)], 0));
  } catch (...) {
    if (overlap) (void)hipStreamSynchronize(lg.side);
    throw;
  }
}

void link_gemm_pass(pqg_ci* ci, int pass, const double* C, double* S, bool accumulate, hipStream_t st) {
  lg_pass<double>(ci, pass, C, S, accumulate, st);
}
void link_gemm_pass(pqg_ci* ci, int pass, const float* C, float* S, bool accumulate, hipStream_t st) {
  lg_pass<float>(ci, pass, C, S, accumulate, st);
}

template <typename T>
static void lg_diag(pqg_ci* ci, const T* C, int64_t ldc, T* S, int64_t lds, int k, hipStream_t st) {
  dim3 g((unsigned)lg_grid(ci->n_int, 256), (unsigned)k);
  k_lg_diag<T><<<g, 256, 0, st>>>(ci->hdiag.get(), C, ldc, S, lds, ci->n_int);
  PQG_HIP(hipGetLastError());
}
void link_gemm_diag(pqg_ci* ci, const double* C, int64_t ldc, double* S, int64_t lds, int k, hipStream_t st) {
  lg_diag<double>(ci, C, ldc, S, lds, k, st);
}
void link_gemm_diag(pqg_ci* ci, const float* C, int64_t ldc, float* S, int64_t lds, int k, hipStream_t st) {
  lg_diag<float>(ci, C, ldc, S, lds, k, st);
}

}  // namespace pqg
