// Sigma for CI spaces whose strings all hold at most one particle (full CI), on MFMA.
//
// With one particle per string a determinant is a tuple of orbitals (x_d) over the "active"
// string sets d (those with one particle), and H is a sum of Kronecker-structured terms: a
// one-body term h_d on each index and, for each pair of active sets (d, e), the coupling
// f_de (x' x | y' y)_de acting on the index pair.  Every Slater-Condon case of the reference
// (slater_condon*.cpp) reduces to these: singles carry the spectator Coulomb sums, doubles
// one coupling element, phases are +1 (no occupied orbital lies between hole and particle
// of a one-particle string), and same-string exchange terms vanish.
//
// So sigma = sum over pairs (d, e) of  W_de  applied to C viewed as a (n_d n_e) x rest matrix,
// with W_de[(x'y'),(xy)] = f_de (x'x|y'y)_de and the one-body matrices folded into the first
// pair that contains each index.  Each term is one dense GEMM on the matrix cores (FP64
// v_mfma_f64_16x16x4f64, or FP32 v_mfma_f32_16x16x4f32 for the mixed-precision bulk) that
// addresses C and sigma in place through the strides of the pair's indices.
#include <algorithm>
#include <cstring>

#include "pqg_ci.hpp"
#include "pqg_mfma.hpp"

namespace pqg {

// One launch computes every pair term of sigma.  Pair p's matrix W_p (M x M, row-major) is
// applied to C viewed as an M x (rest * k) matrix without moving C: row m = (x1, x2) of the
// pair sits at offset x1*s1 + x2*s2 of the natural layout, column n = (root, r1, r2) at
// root*ldc + r1*sr1 + r2*sr2.  K is split into `split` chunks for parallelism; each
// (pair, chunk) writes its own partial sigma in the natural layout and a second kernel sums
// the partials in a fixed order, so results are bit-reproducible (no atomics).
struct PairGemm {
  const void* W;
  int M, n2;         // m = x1 * n2 + x2
  int64_t s1, s2;
  int rest, nr2;     // j = r1 * nr2 + r2
  int64_t sr1, sr2;
  int mtiles, ntiles, kchunk, block0;
};

constexpr int kMaxPairs = 6;

struct PairGemmArgs {
  PairGemm p[kMaxPairs];
  int npairs, split, k;
  int64_t N, ldc;
};

constexpr int PG_BM = 64, PG_BK = 16;

// Block = 4 waves; wave w owns rows [16w, 16w+16) of the 64-row tile and all BN = 16*NF
// columns.  Global loads of the next K tile are issued before the MFMAs of the current one.
template <typename T, int NF>
__global__ __launch_bounds__(256) void k_pair_gemm(PairGemmArgs a, const T* __restrict__ C, T* __restrict__ P) {
  constexpr int BN = 16 * NF;
  constexpr int PAD = 16 / sizeof(T);
  constexpr int NA = PG_BM * PG_BK / 256, NB = BN * PG_BK / 256;
  __shared__ T As[PG_BK][PG_BM + PAD];
  __shared__ T Bs[PG_BK][BN + PAD];
  __shared__ int64_t colIn[BN], colOut[BN];
  __shared__ int colOk[BN];
  const int bid = blockIdx.x;
  int pi = 0;
#pragma unroll
  for (int q = 1; q < kMaxPairs; q++)
    if (q < a.npairs && bid >= a.p[q].block0) pi = q;
  PairGemm g = a.p[0];
#pragma unroll
  for (int q = 1; q < kMaxPairs; q++)
    if (q == pi) g = a.p[q];
  const T* __restrict__ W = reinterpret_cast<const T*>(g.W);
  const int local = bid - g.block0;
  const int sp = local % a.split, t = local / a.split;
  const int mt = t % g.mtiles, nt = t / g.mtiles;
  const int m0 = mt * PG_BM, n0 = nt * BN;
  const int M = g.M, ntot = g.rest * a.k;
  const int kbeg = sp * g.kchunk, kend = min(M, kbeg + g.kchunk);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < BN; i += 256) {
    const int gn = n0 + i;
    const int root = gn / g.rest, j = gn % g.rest;
    const int64_t off = (int64_t)(j / g.nr2) * g.sr1 + (int64_t)(j % g.nr2) * g.sr2;
    colOk[i] = gn < ntot;
    colIn[i] = (int64_t)root * a.ldc + off;
    colOut[i] = (int64_t)root * a.N + off;
  }
  __syncthreads();
  T ra[NA], rb[NB];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int e = tid + 256 * i, mm = e / PG_BK, kk = e % PG_BK;
      const int gm = m0 + mm, gk = k0 + kk;
      ra[i] = (gm < M && gk < kend) ? W[(int64_t)gm * M + gk] : T(0);
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int e = tid + 256 * i, kk = e / BN, nn = e % BN;
      const int gk = k0 + kk;
      T v = T(0);
      if (gk < kend && colOk[nn]) {
        const int64_t rk = (int64_t)(gk / g.n2) * g.s1 + (int64_t)(gk % g.n2) * g.s2;
        v = C[rk + colIn[nn]];
      }
      rb[i] = v;
    }
  };
  typedef typename Mfma<T>::acc_t acc_t;
  acc_t acc[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) acc[f] = acc_t{0, 0, 0, 0};
  if (kbeg < kend) load(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += PG_BK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NA; i++) {
      const int e = tid + 256 * i;
      As[e % PG_BK][e / PG_BK] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      const int e = tid + 256 * i;
      Bs[e / BN][e % BN] = rb[i];
    }
    __syncthreads();
    if (k0 + PG_BK < kend) load(k0 + PG_BK);
    // all fragments of the tile first, so one LDS wait covers the 4 * NF MFMAs that follow
    T av[PG_BK / 4], bv[PG_BK / 4][NF];
#pragma unroll
    for (int q = 0; q < PG_BK / 4; q++) {
      const int kk = 4 * q + (lane >> 4);
      av[q] = As[kk][wv * 16 + (lane & 15)];
#pragma unroll
      for (int f = 0; f < NF; f++) bv[q][f] = Bs[kk][f * 16 + (lane & 15)];
    }
#pragma unroll
    for (int q = 0; q < PG_BK / 4; q++)
#pragma unroll
      for (int f = 0; f < NF; f++) acc[f] = Mfma<T>::op(av[q], bv[q][f], acc[f]);
  }
  T* __restrict__ out = P + (int64_t)(pi * a.split + sp) * a.N * a.k;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int gm = m0 + wv * 16 + Mfma<T>::row(r, lane);
    if (gm >= M) continue;
    const int64_t rowoff = (int64_t)(gm / g.n2) * g.s1 + (int64_t)(gm % g.n2) * g.s2;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const int nn = f * 16 + (lane & 15);
      if (colOk[nn]) out[rowoff + colOut[nn]] = acc[f][r];
    }
  }
}

// S[root][i] = sum over the partials, in a fixed order
template <typename T>
__global__ void k_pair_reduce(const T* __restrict__ P, int nparts, int64_t N, int k, T* __restrict__ S, int64_t lds) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int root = blockIdx.y;
  if (i >= N) return;
  const int64_t stride = N * k;
  const T* p = P + (int64_t)root * N + i;
  T v = p[0];
  for (int q = 1; q < nparts; q++) v += p[q * stride];
  S[(int64_t)root * lds + i] = v;
}

__global__ void k_build_w(const double* __restrict__ V, int npair_e, const double* __restrict__ h1, const double* __restrict__ h2,
                          int n1, int n2, double f, double* __restrict__ W) {
  // W[(x'y'),(xy)] = f * V[idx2(x',x)][idx2(y',y)] + h1[x'][x] d(y',y) + h2[y'][y] d(x',x)
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nn = (int64_t)n1 * n2;
  if (idx >= nn * nn) return;
  const int64_t r = idx / nn, c = idx % nn;
  const int xp = (int)(r / n2), yp = (int)(r % n2), x = (int)(c / n2), y = (int)(c % n2);
  const int P = xp > x ? xp * (xp + 1) / 2 + x : x * (x + 1) / 2 + xp;
  const int Q = yp > y ? yp * (yp + 1) / 2 + y : y * (y + 1) / 2 + yp;
  double v = f * V[(size_t)P * npair_e + Q];
  if (h1 && yp == y) v += h1[(size_t)xp * n1 + x];
  if (h2 && xp == x) v += h2[(size_t)yp * n2 + y];
  W[idx] = v;
}

__global__ void k_w_to_f32(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

static int tg(int64_t n, int b) { return (int)((n + b - 1) / b); }
static thread_local pqg_ctx* g_tensor_ctx = nullptr;  // context of the running sigma (profiling)

}  // namespace pqg

// ---------------------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------------------
namespace pqg {

bool tensor_eligible(const pqg_ci* ci) {
  if (!ci->fci) return false;
  int active = 0;
  for (int t = 0; t < ci->nslot; t++) {
    if (ci->slot[t].nocc > 1) return false;
    if (ci->slot[t].nocc == 1) active++;
  }
  return active >= 1;
}

// Builds the pair matrices.  `h1` holds each slot's one-body matrix, row-major, on the device.
void tensor_setup(pqg_ci* ci, const std::vector<const double*>& h1) {
  auto& tp = ci->tensor;
  tp.enabled = tensor_eligible(ci);
  if (!tp.enabled) return;
  pqg_ctx* ctx = ci->ctx;
  hipStream_t st = ci->stream;
  tp.dims.clear();
  for (int t = 0; t < ci->nslot; t++)
    if (ci->slot[t].nocc == 1) tp.dims.push_back(t);
  const int nd = (int)tp.dims.size();
  std::vector<bool> h_done(nd, false);
  tp.pairs.clear();
  if (nd == 1) {
    TensorPair pr;
    pr.d1 = 0;
    pr.d2 = -1;
    const int t = tp.dims[0];
    const int n = ci->slot[t].n_orb;
    pr.W.reset(ctx, (size_t)n * n);
    PQG_HIP(hipMemcpyAsync(pr.W.get(), h1[t], (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
    tp.pairs.push_back(std::move(pr));
  }
  for (int i = 0; i < nd; i++)
    for (int j = i + 1; j < nd; j++) {
      TensorPair pr;
      pr.d1 = i;
      pr.d2 = j;
      const int t = tp.dims[i], u = tp.dims[j];
      const int n1 = ci->slot[t].n_orb, n2 = ci->slot[u].n_orb;
      const int64_t nn = (int64_t)n1 * n2;
      pr.W.reset(ctx, (size_t)(nn * nn));
      const double* h_t = h_done[i] ? nullptr : h1[t];
      const double* h_u = h_done[j] ? nullptr : h1[u];
      h_done[i] = h_done[j] = true;
      // V64[t][u] already carries the charge factor
      k_build_w<<<tg(nn * nn, 256), 256, 0, st>>>(ci->V64[t][u].get(), ci->slot[u].npair, h_t, h_u, n1, n2, 1.0,
                                                  pr.W.get());
      PQG_HIP(hipGetLastError());
      tp.pairs.push_back(std::move(pr));
    }
  PQG_HIP(hipStreamSynchronize(st));
}

static void ensure_tensor_f32(pqg_ci* ci) {
  auto& tp = ci->tensor;
  if (tp.f32_ready) return;
  for (auto& pr : tp.pairs) {
    pr.W32.reset(ci->ctx, pr.W.n);
    k_w_to_f32<<<tg(pr.W.n, 256), 256, 0, ci->stream>>>(pr.W.get(), pr.W32.get(), pr.W.n);
    PQG_HIP(hipGetLastError());
  }
  tp.f32_ready = true;
}

template <typename T, int NF>
static void launch_pair_gemm(const PairGemmArgs& a, int blocks, const T* C, T* P, hipStream_t st) {
  k_pair_gemm<T, NF><<<blocks, 256, 0, st>>>(a, C, P);
}

template <typename T>
static void tensor_sigma_impl(pqg_ci* ci, const T* C, T* S, int k, int64_t ld, hipStream_t st) {
  auto& tp = ci->tensor;
  g_tensor_ctx = ci->ctx;
  const int64_t N = ci->n_int;
  const int nd = (int)tp.dims.size();
  std::vector<int64_t> n(nd), stride(nd);
  for (int d = 0; d < nd; d++) n[d] = ci->slot[tp.dims[d]].n_orb;
  int64_t sacc = 1;
  for (int d = nd - 1; d >= 0; --d) {
    stride[d] = sacc;
    sacc *= n[d];
  }
  const int npairs = (int)tp.pairs.size();
  PQG_REQUIRE(npairs >= 1 && npairs <= kMaxPairs, PQG_ERR_UNSUPPORTED, "tensor sigma: unsupported number of active string sets");
  // column tile width: the NF in {1, 3, 4, 9} with the least padding of rest*k (ties: wider)
  int64_t ntot_max = 0;
  for (auto& pr : tp.pairs) {
    const int64_t lead = pr.d2 < 0 ? n[pr.d1] : n[pr.d1] * n[pr.d2];
    ntot_max = std::max(ntot_max, (N / lead) * k);
  }
  int NF = 1;
  int64_t best = INT64_MAX;
  for (int f : {1, 3, 4, 9}) {
    const int64_t padded = (ntot_max + 16 * f - 1) / (16 * f) * (16 * f);
    if (padded <= best) {
      best = padded;
      NF = f;
    }
  }
  PairGemmArgs a{};
  a.npairs = npairs;
  a.k = k;
  a.N = N;
  a.ldc = ld;
  int base_blocks = 0;
  for (int q = 0; q < npairs; q++) {
    const auto& pr = tp.pairs[q];
    PairGemm& g = a.p[q];
    g.W = sizeof(T) == 8 ? static_cast<const void*>(pr.W.get()) : static_cast<const void*>(pr.W32.get());
    std::vector<int> restd;
    for (int d = 0; d < nd; d++)
      if (d != pr.d1 && d != pr.d2) restd.push_back(d);
    PQG_REQUIRE(restd.size() <= 2, PQG_ERR_UNSUPPORTED, "tensor sigma: more than two spectator string sets");
    if (pr.d2 < 0) {
      g.M = (int)n[pr.d1];
      g.n2 = 1;
      g.s1 = stride[pr.d1];
      g.s2 = 0;
    } else {
      g.M = (int)(n[pr.d1] * n[pr.d2]);
      g.n2 = (int)n[pr.d2];
      g.s1 = stride[pr.d1];
      g.s2 = stride[pr.d2];
    }
    g.rest = (int)(N / g.M);
    g.nr2 = restd.size() == 2 ? (int)n[restd[1]] : 1;
    g.sr1 = restd.empty() ? 0 : stride[restd[0]];
    g.sr2 = restd.size() == 2 ? stride[restd[1]] : 0;
    g.mtiles = (g.M + PG_BM - 1) / PG_BM;
    g.ntiles = (int)((g.rest * (int64_t)k + 16 * NF - 1) / (16 * NF));
    base_blocks += g.mtiles * g.ntiles;
  }
  // split K until the chip holds about two blocks per CU, within a partial-buffer budget
  const int target = 2 * std::max(1, ci->ctx->num_cu);
  int split = 1;
  int minM = INT32_MAX;
  for (int q = 0; q < npairs; q++) minM = std::min(minM, a.p[q].M);
  while (base_blocks * split * 2 <= target && split * 2 * PG_BK <= minM && split < 16) split *= 2;
  a.split = split;
  int blocks = 0;
  for (int q = 0; q < npairs; q++) {
    PairGemm& g = a.p[q];
    g.kchunk = ((g.M + split - 1) / split + PG_BK - 1) / PG_BK * PG_BK;
    g.block0 = blocks;
    blocks += g.mtiles * g.ntiles * split;
  }
  const int nparts = npairs * split;
  const size_t need = (size_t)nparts * N * k * sizeof(T);
  grow(ci->scratch, ci->ctx, need, st);
  T* P = reinterpret_cast<T*>(ci->scratch.get());
  int64_t flops = 0;
  for (int q = 0; q < npairs; q++) flops += 2ll * a.p[q].M * (int64_t)a.p[q].M * a.p[q].rest * k;
  {
    ProfScope prof(g_tensor_ctx, st, sizeof(T) == 8 ? "tensor_gemm_f64" : "tensor_gemm_f32", flops);
    switch (NF) {
      case 1: launch_pair_gemm<T, 1>(a, blocks, C, P, st); break;
      case 3: launch_pair_gemm<T, 3>(a, blocks, C, P, st); break;
      case 4: launch_pair_gemm<T, 4>(a, blocks, C, P, st); break;
      default: launch_pair_gemm<T, 9>(a, blocks, C, P, st); break;
    }
    PQG_HIP(hipGetLastError());
  }
  {
    ProfScope prof(g_tensor_ctx, st, sizeof(T) == 8 ? "tensor_reduce_f64" : "tensor_reduce_f32",
                   (int64_t)(nparts + 1) * N * k * (int64_t)sizeof(T));
    dim3 gr(tg(N, 256), k);
    k_pair_reduce<T><<<gr, 256, 0, st>>>(P, nparts, N, k, S, ld);
    PQG_HIP(hipGetLastError());
  }
}

void tensor_sigma(pqg_ci* ci, const double* C, double* S, int k, int64_t ld, hipStream_t st) {
  ProfScope prof(ci->ctx, st, "sigma_tensor_f64", k);
  tensor_sigma_impl<double>(ci, C, S, k, ld, st);
}

void tensor_sigma(pqg_ci* ci, const float* C, float* S, int k, int64_t ld, hipStream_t st) {
  ensure_tensor_f32(ci);
  ProfScope prof(ci->ctx, st, "sigma_tensor_f32", k);
  tensor_sigma_impl<float>(ci, C, S, k, ld, st);
}

}  // namespace pqg
