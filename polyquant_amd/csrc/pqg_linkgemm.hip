// Sigma terms of one-species full CI on the matrix cores ("link GEMM").
//
// C is viewed as an n_r x n_c matrix whose rows are indexed by the strings of set r and
// whose columns by those of set c (r = alpha on the natural layout, r = beta on the
// transposed vector).  A link of a string S is one single excitation S -> S - h + p,
// numbered e = a * nv + b with h the a-th occupied and p the b-th virtual orbital of S
// (ascending); Q(S, e) = idx2(h, p), s(S, e) = the excitation's phase, J(S, e) the excited
// string and rev(S, e) the number of the link of J(S, e) that leads back to S.
//
// Two pieces of H are formed here (the diagonal and the singles with their spectator sums
// stay in the enumeration kernel of pqg_ci.hip):
//   r x c single-single couplings  sum_{Qr,Qc} (Qr|Qc) E^r_Qr E^c_Qc C     (r = alpha only)
//   same-set doubles of r          1/2 sum_{Q,Q'} (Q|Q') E^r_Q E^r_Q' C over the paths
//                                  I -> J -> K whose two links share neither hole nor particle
//                                  (exactly the double excitations I <-> K; the four paths of
//                                  each double add up to phase * ((h0p0|h1p1) - (h0p1|h1p0)),
//                                  slater_condon_singleparticle.cpp:107-175)
// Both are   sigma(I, m) += sum_e s(I,e) U[J(I,e)][rev(I,e)][m]   with, per row string J,
//   U[J][e'][m] = sum_k A_J[e'][k] T_J[k][m]         (one GEMM per row string, on MFMA)
//   k < nb:  T_J[k][m] = s * C(J, m ^ pair k)  when exactly one orbital of the off-diagonal
//            pair k of set c is occupied in m (0 otherwise);  A_J[e'][k] = (Q(J,e')|pair k)
//   k >= nb: T_J[nb+e][m] = C(J(J,e), m);  A_J[e'][nb+e] = 1/2 s(J,e) (Q(J,e')|Q(J,e)) when the
//            links e', e of J have different holes and different particles (0 otherwise).
// The GEMM writes U for a chunk of columns; a second kernel gathers it back (pull form, fixed
// summation order, so results are bit-reproducible).
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "pqg_ci.hpp"
#include "pqg_mfma.hpp"

namespace pqg {

constexpr int LG_NT = 32;  // columns per tile
constexpr int LG_MAXL = 128;
// prefetch registers per thread: pair-map entries (nb * LG_NT / 256) and excited-row values
// (L * LG_NT / 256); the kernel needs nb <= 128 and L <= 64
constexpr int LG_TBMAX = 16, LG_TAMAX = 8;

constexpr int LG_MAXSETS = 3;

// Row strings are tuples over up to three string sets (e.g. electron alpha x positron alpha x
// positron beta); the row index is sum_s i_s * stride_s.  Links of a row are the links of its
// strings in set order, numbered from loff[s].
template <typename T>
struct LgArgs {
  const T* C;          // this vector, n_r x n_c row-major
  T* U;                // chunk scratch [n_r][L][cw]
  int nsets;
  const int2* links[LG_MAXSETS];  // per set [n_s][L_s]: {J | odd << 31, Q << 16 | rev}
  int Ls[LG_MAXSETS], nvs[LG_MAXSETS], ns[LG_MAXSETS], loff[LG_MAXSETS], np[LG_MAXSETS];
  int64_t strides[LG_MAXSETS];
  const T* Vrc[LG_MAXSETS];              // (P|Q): P a pair of row set s, Q a pair of the column set
  const T* Vrr[LG_MAXSETS][LG_MAXSETS];  // (P|Q): P of row set s, Q of row set t (charges folded in)
  const int32_t* tbm;  // column-set pair map [nb][n_c]: (index << 1 | odd) or -1
  int np_c;
  int n_r, n_c, L, nb, K;
  int c0, cw, gcols;
  int lab_off, kbf_off, crow_off, tile_off;  // LDS byte offsets (links at 0)
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

// Links of row J into LDS: lq[e] = {target row | odd << 31, Q << 16 | rev}, lab[e] =
// set << 16 | hole number << 8 | particle number.
template <typename T>
__device__ void lg_row_links(const LgArgs<T>& a, int64_t J, int2* lq, int* lab, int tid) {
  for (int e = tid; e < a.L; e += 256) {
    int s = 0;
    if (a.nsets > 1 && e >= a.loff[1]) s = 1;
    if (a.nsets > 2 && e >= a.loff[2]) s = 2;
    const int el = e - pick3(a.loff, s);
    const int64_t stride = pick3(a.strides, s);
    const int64_t i = (J / stride) % pick3(a.ns, s);
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

// Block = 4 waves.  With NSPLIT = 1 wave w owns M blocks w, w + 4, ... (MPW of them) and
// every column of the tile; with NSPLIT = 2 waves (2j, 2j+1) share M block j and split the
// columns.  The A fragments of a row string stay in registers for all of its tiles.
template <typename T, int MPW, int NSPLIT, int KQ>
__global__ __launch_bounds__(256) void k_lg_u(const LgArgs<T> a) {
  constexpr int NT = LG_NT, NF = NT / 16, NFW = NF / NSPLIT;
  constexpr int NTP = NT + 16 / (int)sizeof(T);
  typedef typename Mfma<T>::acc_t acc_t;
  extern __shared__ __align__(16) char smem[];
  int2* lq = reinterpret_cast<int2*>(smem);
  T* crow = reinterpret_cast<T*>(smem + a.crow_off);
  T* Ts = reinterpret_cast<T*>(smem + a.tile_off);
  const int J = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int L = a.L, K = a.K, nb = a.nb;
  constexpr int KP = 4 * KQ;  // rows of the LDS tile; rows K..KP-1 stay zero
  // LDS: links of J, set/hole/particle numbers of each link, full pair index of each
  // off-diagonal pair of the column set, the staged row C(J, :), the B tile
  int* lab = reinterpret_cast<int*>(smem + a.lab_off);
  int* kbf = reinterpret_cast<int*>(smem + a.kbf_off);
  lg_row_links(a, J, lq, lab, tid);
  for (int k = tid; k < nb; k += 256) kbf[k] = k + offpair_hi(k);
  if (nb > 0)
    for (int m = tid; m < a.n_c; m += 256) crow[m] = a.C[(size_t)J * a.n_c + m];
  for (int i = tid; i < (KP - K) * NTP; i += 256) Ts[K * NTP + i] = T(0);
  __syncthreads();

  const int mb0 = w / NSPLIT;
  const int fb = (w % NSPLIT) * NFW;
  const bool active = 16 * mb0 < L;
  T af[MPW][KQ];
#pragma unroll
  for (int i = 0; i < MPW; i++) {
    const int ep = 16 * (mb0 + i * (4 / NSPLIT)) + (lane & 15);
    const bool rowok = ep < L;
    const int Qp = rowok ? (lq[ep].y >> 16) : 0;
    const int abp = rowok ? lab[ep] : 0;
    const int sp = abp >> 16;
#pragma unroll
    for (int q = 0; q < KQ; q++) {
      const int k = 4 * q + (lane >> 4);
      T v = T(0);
      if (rowok && k < K) {
        if (k < nb) {
          v = pick3(a.Vrc, sp)[(size_t)Qp * a.np_c + kbf[k]];
        } else {
          const int e = k - nb;
          const int abe = lab[e];
          const int se = abe >> 16;
          const int2 l = lq[e];
          if (se == sp) {
            // same-set doubles: the two links share neither hole nor particle
            if (((abe ^ abp) & 0xff00) && ((abe ^ abp) & 0xff))
              v = (l.x < 0 ? T(-0.5) : T(0.5)) * pick33(a.Vrr, sp, sp)[(size_t)Qp * pick3(a.np, sp) + (l.y >> 16)];
          } else if (sp < se) {
            // one excitation in each of two row sets: only the path whose first link lies in
            // the lower set carries the coupling
            v = (l.x < 0 ? T(-1) : T(1)) * pick33(a.Vrr, sp, se)[(size_t)Qp * pick3(a.np, se) + (l.y >> 16)];
          }
        }
      }
      af[i][q] = v;
    }
  }

  const int64_t gbeg = (int64_t)a.c0 + (int64_t)blockIdx.y * a.gcols;
  const int64_t gend = std::min<int64_t>((int64_t)a.c0 + a.cw, gbeg + a.gcols);
  // The global loads of the next tile (pair-map entries and excited rows) are issued into
  // registers before the MFMAs of the current one; thread t holds entries t + 256 i.
  const int nbt = nb * NT, nat = L * NT;
  int tbv[LG_TBMAX];
  T tav[LG_TAMAX];
  // Branch-free: every slot loads from a clamped in-bounds address; which slots hold real
  // entries (and which columns lie inside the chunk) is decided when the tile is stored.
  const int64_t clast = (int64_t)a.n_c - 1;
  auto prefetch = [&](int64_t m0) {
    if (nbt > 0) {
#pragma unroll
      for (int i = 0; i < LG_TBMAX; i++) {
        const int idx = min(tid + 256 * i, nbt - 1);
        const int64_t m = min(m0 + (idx % NT), clast);
        tbv[i] = a.tbm[(size_t)(idx / NT) * a.n_c + m];
      }
    }
#pragma unroll
    for (int i = 0; i < LG_TAMAX; i++) {
      const int idx = min(tid + 256 * i, nat - 1);
      const int64_t m = min(m0 + (idx % NT), clast);
      tav[i] = a.C[(size_t)(lq[idx / NT].x & 0x7fffffff) * a.n_c + m];
    }
  };
  if (gbeg < gend) prefetch(gbeg);
  for (int64_t m0 = gbeg; m0 < gend; m0 += NT) {
#pragma unroll
    for (int i = 0; i < LG_TBMAX; i++) {
      const int idx = tid + 256 * i;
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
#pragma unroll
    for (int i = 0; i < LG_TAMAX; i++) {
      const int idx = tid + 256 * i;
      if (idx < nat) Ts[(nb + idx / NT) * NTP + idx % NT] = (m0 + idx % NT < gend) ? tav[i] : T(0);
    }
    __syncthreads();
    if (m0 + NT < gend) prefetch(m0 + NT);
    if (active) {
      acc_t acc[MPW][NFW];
#pragma unroll
      for (int i = 0; i < MPW; i++)
#pragma unroll
        for (int f = 0; f < NFW; f++) acc[i][f] = acc_t{0, 0, 0, 0};
      // every K step runs: the padding rows of the tile and of A are zero
#pragma unroll
      for (int q = 0; q < KQ; q++) {
        const T* brow = Ts + (4 * q + (lane >> 4)) * NTP + 16 * fb + (lane & 15);
#pragma unroll
        for (int f = 0; f < NFW; f++) {
          const T b = brow[16 * f];
#pragma unroll
          for (int i = 0; i < MPW; i++) acc[i][f] = Mfma<T>::op(af[i][q], b, acc[i][f]);
        }
      }
#pragma unroll
      for (int i = 0; i < MPW; i++) {
        const int mb = mb0 + i * (4 / NSPLIT);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int ep = 16 * mb + Mfma<T>::row(r, lane);
          if (ep >= L) continue;
          T* urow = a.U + ((size_t)J * L + ep) * a.cw;
#pragma unroll
          for (int f = 0; f < NFW; f++) {
            const int64_t m = m0 + 16 * (fb + f) + (lane & 15);
            if (m < gend) urow[m - a.c0] = acc[i][f][r];
          }
        }
      }
    }
    __syncthreads();
  }
}

// sigma(I, c0 + j) += sum_e s(I,e) U[J(I,e)][rev(I,e)][j] for the columns j of a chunk
template <typename T, int CPT>
__global__ __launch_bounds__(256) void k_lg_gather(const LgArgs<T> a, T* __restrict__ S) {
  __shared__ int2 lq[LG_MAXL];
  __shared__ int lab[LG_MAXL];
  const int64_t I = blockIdx.x;
  lg_row_links(a, I, lq, lab, threadIdx.x);
  __syncthreads();
  const int L = a.L, cw = a.cw;
  const T* __restrict__ U = a.U;
  const int j0 = blockIdx.y * 256 * CPT + threadIdx.x;
  T acc[CPT];
#pragma unroll
  for (int j = 0; j < CPT; j++) acc[j] = T(0);
  for (int e = 0; e < L; e++) {
    const int2 l = lq[e];
    const T* src = U + ((size_t)(l.x & 0x7fffffff) * L + (l.y & 0xffff)) * cw;
    const bool neg = l.x < 0;
#pragma unroll
    for (int j = 0; j < CPT; j++) {
      const int col = j0 + 256 * j;
      if (col < cw) {
        const T u = src[col];
        acc[j] += neg ? -u : u;
      }
    }
  }
  T* dst = S + (size_t)I * a.n_c + a.c0;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    const int col = j0 + 256 * j;
    if (col < cw) dst[col] += acc[j];
  }
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
// at most 16 GiB
size_t lg_budget_bytes(int64_t N, size_t w) {
  return std::min<size_t>(16ull << 30, std::max<size_t>(64ull << 20, 8ull * (size_t)N * w));
}

// K steps are a compile-time bound (multiple of 8) so the MFMA loop has no branches
static int lg_kq_bound(int kq) { return kq <= 48 ? (kq + 7) / 8 * 8 : 0; }

// Row sets of the two passes.  Pass A (natural layout; with positrons, on C permuted to
// (e-alpha, positron strings, e-beta)): rows = e-alpha x positron strings, with the links of
// e-alpha and of every positron set that holds a particle; columns = e-beta.  Pass B
// (transposed vector): rows = e-beta x positron strings with the e-beta links only; columns =
// e-alpha.
struct LgPass {
  int nsets = 0;
  int slot[LG_MAXSETS] = {0, 0, 0};
  int64_t stride[LG_MAXSETS] = {1, 1, 1};
  int col = 1;        // column slot
  bool pairs = true;  // in-row block over the column set's off-diagonal pairs
  int64_t n_r = 0;
  int L = 0;
};

static int64_t positron_strings(const pqg_ci* ci) {
  return ci->nslot == 4 ? (int64_t)ci->slot[2].nstr * ci->slot[3].nstr : 1;
}

static LgPass lg_pass_desc(const pqg_ci* ci, int pass) {
  LgPass d;
  const int64_t nP = positron_strings(ci);
  if (pass == 0) {
    d.col = 1;
    d.pairs = true;
    d.n_r = (int64_t)ci->slot[0].nstr * nP;
    d.slot[d.nsets] = 0;
    d.stride[d.nsets++] = nP;
    if (ci->nslot == 4) {
      if (ci->lg.tab[2].L > 0) {
        d.slot[d.nsets] = 2;
        d.stride[d.nsets++] = ci->slot[3].nstr;
      }
      if (ci->lg.tab[3].L > 0) {
        d.slot[d.nsets] = 3;
        d.stride[d.nsets++] = 1;
      }
    }
  } else {
    d.col = 0;
    d.pairs = false;
    d.n_r = (int64_t)ci->slot[1].nstr * nP;
    d.slot[d.nsets] = 1;
    d.stride[d.nsets++] = nP;
  }
  for (int q = 0; q < d.nsets; q++) d.L += ci->lg.tab[d.slot[q]].L;
  return d;
}

bool link_gemm_eligible(const pqg_ci* ci) {
  if (!ci->fci || ci->nslot < 2) return false;
  if (std::getenv("PQG_NO_LINK_GEMM")) return false;
  for (int t = 0; t < ci->nslot; t++)
    if (!ci->slot[t].complete) return false;
  for (int t = 0; t < 2; t++)
    if (ci->slot[t].nocc < 1 || ci->slot[t].nvirt < 1) return false;
  const Slot& sb = ci->slot[1];
  if (sb.n_orb * (sb.n_orb - 1) / 2 * LG_NT > 256 * LG_TBMAX) return false;
  for (int pass = 0; pass < 2; pass++) {
    // link counts as the pass would see them
    int L = 0;
    for (int t = 0; t < ci->nslot; t++) {
      const bool in = pass == 0 ? (t != 1) : (t == 1);
      if (in) L += ci->slot[t].nocc * ci->slot[t].nvirt;
    }
    if (L * LG_NT > 256 * LG_TAMAX) return false;
    const Slot& sc = ci->slot[pass == 0 ? 1 : 0];
    const int nb = pass == 0 ? sc.n_orb * (sc.n_orb - 1) / 2 : 0;
    const int KQ = lg_kq_bound((nb + L + 3) / 4);
    if (!KQ) return false;
    const size_t lds = 2048 + (nb > 0 ? (size_t)sc.nstr * 8 : 0) + (size_t)4 * KQ * (LG_NT + 2) * 8;
    if (lds > 160 * 1024 - 1024) return false;
    const int64_t nr = (int64_t)ci->slot[pass == 0 ? 0 : 1].nstr * positron_strings(ci);
    if (nr >= (1ll << 31) - 1) return false;
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
    if (lt.L == 0) continue;
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
    PQG_HIP(hipMemcpy(lt.links.get(), links.data(), links.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (t == 1) {  // only the electron beta set serves as the column set of the pair block
      std::vector<int32_t> map((size_t)lt.nb * s.nstr, -1);
      for (int p = 1; p < n; p++)
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
      lt.tbm.reset(ci->ctx, map.size());
      PQG_HIP(hipMemcpy(lt.tbm.get(), map.data(), map.size() * 4, hipMemcpyHostToDevice));
    }
  }
}

template <typename T, int MPW, int NSPLIT, int KQ>
static void lg_go(const LgArgs<T>& a, dim3 grid, size_t lds, hipStream_t st) {
  auto kern = k_lg_u<T, MPW, NSPLIT, KQ>;
  PQG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  kern<<<grid, 256, lds, st>>>(a);
}

template <typename T, int MPW, int NSPLIT>
static void lg_go_kq(const LgArgs<T>& a, int KQ, dim3 grid, size_t lds, hipStream_t st) {
  switch (KQ) {
    case 8: lg_go<T, MPW, NSPLIT, 8>(a, grid, lds, st); break;
    case 16: lg_go<T, MPW, NSPLIT, 16>(a, grid, lds, st); break;
    case 24: lg_go<T, MPW, NSPLIT, 24>(a, grid, lds, st); break;
    case 32: lg_go<T, MPW, NSPLIT, 32>(a, grid, lds, st); break;
    case 40: lg_go<T, MPW, NSPLIT, 40>(a, grid, lds, st); break;
    default: lg_go<T, MPW, NSPLIT, 48>(a, grid, lds, st); break;
  }
}

static const double* vblock(pqg_ci* ci, int t, int u, double) { return ci->V64[t][u].get(); }
static const float* vblock(pqg_ci* ci, int t, int u, float) {
  return t == u ? ci->lg.V32s[t].get() : ci->V32[t][u].get();
}

// U bytes per column of a pass, and the chunk width that keeps U within its budget
static size_t lg_per_col(const pqg_ci* ci, const LgPass& d, size_t w) { return (size_t)d.n_r * d.L * w; }
static int lg_chunk(const pqg_ci* ci, const LgPass& d, int n_c, size_t w) {
  const size_t per_col = lg_per_col(ci, d, w);
  const size_t want = std::min<size_t>(lg_budget_bytes(ci->n_int, w), per_col * (size_t)n_c);
  const size_t budget = std::max<size_t>(per_col * LG_NT, want);
  int cw = (int)std::min<size_t>((size_t)n_c, budget / per_col);
  if (cw < n_c) cw = std::max(LG_NT, cw / LG_NT * LG_NT);
  return cw;
}

// One pass of the link GEMM for one vector.
template <typename T>
static void lg_pass(pqg_ci* ci, int pass, const T* C, T* S, hipStream_t st) {
  LinkGemm& lg = ci->lg;
  const LgPass d = lg_pass_desc(ci, pass);
  if (d.L == 0) return;
  const Slot& sc = ci->slot[d.col];
  LgArgs<T> a{};
  a.C = C;
  a.nsets = d.nsets;
  int off = 0;
  for (int q = 0; q < d.nsets; q++) {
    const int t = d.slot[q];
    a.links[q] = lg.tab[t].links.get();
    a.Ls[q] = lg.tab[t].L;
    a.nvs[q] = lg.tab[t].nv;
    a.ns[q] = ci->slot[t].nstr;
    a.loff[q] = off;
    a.np[q] = ci->slot[t].npair;
    a.strides[q] = d.stride[q];
    a.Vrc[q] = vblock(ci, t, d.col, T{});
    for (int q2 = 0; q2 < d.nsets; q2++) a.Vrr[q][q2] = vblock(ci, t, d.slot[q2], T{});
    off += lg.tab[t].L;
  }
  a.tbm = d.pairs ? lg.tab[d.col].tbm.get() : nullptr;
  a.np_c = sc.npair;
  a.n_r = (int)d.n_r;
  a.n_c = sc.nstr;
  a.L = d.L;
  a.nb = d.pairs ? lg.tab[d.col].nb : 0;
  a.K = a.nb + a.L;
  const int KQ = lg_kq_bound((a.K + 3) / 4);
  constexpr int NTP = LG_NT + 16 / (int)sizeof(T);
  auto r16 = [](size_t x) { return (x + 15) / 16 * 16; };
  a.lab_off = (int)r16((size_t)a.L * sizeof(int2));
  a.kbf_off = (int)r16(a.lab_off + (size_t)a.L * 4);
  a.crow_off = (int)r16(a.kbf_off + (size_t)a.nb * 4);
  a.tile_off = (int)r16(a.crow_off + (a.nb > 0 ? (size_t)a.n_c * sizeof(T) : 0));
  const size_t lds = a.tile_off + (size_t)4 * KQ * NTP * sizeof(T);
  // U holds n_r * L * cw values; sized for the larger of the two passes so it grows once
  const int cw = lg_chunk(ci, d, a.n_c, sizeof(T));
  size_t need = lg_per_col(ci, d, sizeof(T)) * (size_t)cw;
  {
    const LgPass o = lg_pass_desc(ci, 1 - pass);
    const int n_co = ci->slot[o.col].nstr;
    need = std::max(need, lg_per_col(ci, o, sizeof(T)) * (size_t)lg_chunk(ci, o, n_co, sizeof(T)));
  }
  grow(lg.U, ci->ctx, need, st);
  a.U = reinterpret_cast<T*>(lg.U.get());
  const int mb = (a.L + 15) / 16;
  const int64_t flops_per_col = 2ll * a.n_r * (16ll * mb) * (4ll * KQ);
  for (int c0 = 0; c0 < a.n_c; c0 += cw) {
    a.c0 = c0;
    a.cw = std::min(cw, a.n_c - c0);
    const int tiles = lg_grid(a.cw, LG_NT);
    const int gy = std::max(1, std::min(tiles, lg_grid(2048, a.n_r)));
    a.gcols = lg_grid(tiles, gy) * LG_NT;
    dim3 grid((unsigned)a.n_r, (unsigned)lg_grid(a.cw, a.gcols));
    {
      ProfScope prof(ci->ctx, st, sizeof(T) == 8 ? "lg_gemm_f64" : "lg_gemm_f32", flops_per_col * a.cw);
      if (mb <= 2) lg_go_kq<T, 1, 2>(a, KQ, grid, lds, st);
      else lg_go_kq<T, 1, 1>(a, KQ, grid, lds, st);
      PQG_HIP(hipGetLastError());
    }
    {
      ProfScope prof(ci->ctx, st, sizeof(T) == 8 ? "lg_gather_f64" : "lg_gather_f32",
                     (int64_t)a.n_r * a.cw * (a.L + 2) * (int64_t)sizeof(T));
      constexpr int CPT = 4;
      dim3 g2((unsigned)a.n_r, (unsigned)lg_grid(a.cw, 256 * CPT));
      k_lg_gather<T, CPT><<<g2, 256, 0, st>>>(a, S);
      PQG_HIP(hipGetLastError());
    }
  }
}

void link_gemm_pass(pqg_ci* ci, int pass, const double* C, double* S, hipStream_t st) { lg_pass<double>(ci, pass, C, S, st); }
void link_gemm_pass(pqg_ci* ci, int pass, const float* C, float* S, hipStream_t st) { lg_pass<float>(ci, pass, C, S, st); }

}  // namespace pqg
