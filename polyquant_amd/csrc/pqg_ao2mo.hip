// AO -> MO four-index transform on MFMA (gfx950).
//
// (ij|kl) = sum_{mu nu la si} Ca[mu,i] Ca[nu,j] Cb[la,k] Cb[si,l] (mu nu|la si), the contraction
// of transform_mo_2_body_integrals (src/integral/integral.cpp:207-477), with mu nu, i j in the
// basis and orbitals of species a and la si, k l in those of species b (the same basis for one
// species; two bases for the electron-positron block), organised as two half-transforms over
// pair indices:
//   half 1, for every AO pair P = (mu nu) of basis a:  M_P = unpack((P | ..)),  T_P = Cb^T M_P Cb,
//            B[P][(kl)] = T_P[k][l] (k <= l);
//   half 2, for every MO pair (kl) of basis b:         M = unpack(B[.][(kl)]),  T = Ca^T M Ca,
//            out[(ij) + (kl) * npair_a] = T[i][j]  (output column (kl), written contiguously).
// Each C^T M C is two dense products on the matrix cores: U = M C (nao x nao x nmo, M
// symmetric), then T = U^T C (nmo x nao x nmo).  The coefficient matrices may have fewer
// columns than basis functions (frozen core and deleted virtuals removed by the caller).
// Batches of pairs run as one launch (batched GEMM, 64 x 64 output tile per workgroup, K
// staged through LDS).
//   PQG_AO2MO_FP64:   v_mfma_f64_16x16x4f64, FP64 accumulation.
//   PQG_AO2MO_BF16X3: each FP32 operand split into bf16 hi + lo, three products
//                     hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_bf16, FP32 accumulation.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "pqg_internal.hpp"

namespace pqg {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ int64_t pidx(int64_t i, int64_t j) { return i > j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// ---------------------------------------------------------------------------------------
// unpack / pack
// ---------------------------------------------------------------------------------------
// X[b][nu][mu] = eri8[idx2(idx2(mu, nu), q0 + b)] for b < nb (symmetric n x n per batch)
template <typename T>
__global__ void k_unpack8(const double* __restrict__ eri8, int n, int64_t q0, int nb, T* __restrict__ X) {
  const int64_t n2 = (int64_t)n * n;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n2 * nb) return;
  const int b = (int)(idx / n2);
  const int64_t r = idx % n2;
  const int nu = (int)(r / n), mu = (int)(r % n);
  const int64_t P = pidx(mu, nu), Q = q0 + b;
  X[idx] = (T)eri8[pidx(P, Q)];
}

// X[b][la][si] = B[idx2(la, si)][c0 + b]   (B row-major npair x ldb; also row c0 + b of a
// column-major matrix with leading dimension ldb whose columns are pairs (la si))
template <typename S, typename T>
__global__ void k_unpack_half(const S* __restrict__ B, int64_t ldb, int n, int64_t c0, int nb, T* __restrict__ X) {
  const int64_t n2 = (int64_t)n * n;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n2 * nb) return;
  // consecutive threads take consecutive batch entries for one (la, si): coalesced reads
  const int b = (int)(idx % nb);
  const int64_t r = idx / nb;
  const int la = (int)(r / n), si = (int)(r % n);
  X[b * n2 + r] = (T)B[pidx(la, si) * ldb + c0 + b];
}

// out[b][pidx(i, j)] = T[b][i][j] for i <= j, with row stride ldo (batch b -> row b0 + b)
template <typename T, typename O>
__global__ void k_pack(const T* __restrict__ Tm, int n, int nb, O* __restrict__ out, int64_t ldo) {
  const int64_t np = (int64_t)n * (n + 1) / 2;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= np * nb) return;
  const int b = (int)(idx / np);
  const int64_t p = idx % np;
  // invert p -> (i, j) with i >= j : i = floor((sqrt(8p+1)-1)/2)
  int64_t i = (int64_t)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (i * (i + 1) / 2 > p) i--;
  while ((i + 1) * (i + 2) / 2 <= p) i++;
  const int64_t j = p - i * (i + 1) / 2;
  out[(int64_t)b * ldo + p] = (O)Tm[(int64_t)b * n * n + j * n + i];
}

// ---------------------------------------------------------------------------------------
// batched GEMM: out_b (m x nn) = op(A_b) * Cm,  op(A) = A (TA = 0) or A^T (TA = 1),
// A_b row-major m x kd (TA = 0) or kd x m (TA = 1), Cm row-major kd x nn (shared), out row-major.
// ---------------------------------------------------------------------------------------
constexpr int BM = 64, BN = 64, BK = 16;

// FP64: 4 waves, each a 32 x 32 quadrant = 2 x 2 tiles of 16 x 16 (v_mfma_f64_16x16x4f64).
// Operand layouts: A 16x4 -> lane l holds A[l%16][l/16]; B 4x16 -> lane l holds B[l/16][l%16];
// D 16x16 -> lane l, register r holds D[4*r + l/16][l%16], r = 0..3 (rows interleaved).
template <int TA>
__global__ __launch_bounds__(256) void k_bgemm_f64(const double* __restrict__ A, const double* __restrict__ Cm,
                                                   double* __restrict__ out, int m, int nn, int kd) {
  __shared__ double As[BK][BM + 2];  // As[k][m]
  __shared__ double Bs[BK][BN + 2];  // Bs[k][nn]
  const double* Ab = A + (int64_t)blockIdx.z * m * kd;
  double* Ob = out + (int64_t)blockIdx.z * m * nn;
  const int m0 = blockIdx.y * BM, c0 = blockIdx.x * BN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = d4{0, 0, 0, 0};
  for (int k0 = 0; k0 < kd; k0 += BK) {
    for (int i = tid; i < BK * BM; i += 256) {
      const int kk = i / BM, mm = i % BM;
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < m && gk < kd) v = TA ? Ab[(int64_t)gk * m + gm] : Ab[(int64_t)gm * kd + gk];
      As[kk][mm] = v;
    }
    for (int i = tid; i < BK * BN; i += 256) {
      const int kk = i / BN, c = i % BN;
      const int gn = c0 + c, gk = k0 + kk;
      Bs[kk][c] = (gn < nn && gk < kd) ? Cm[(int64_t)gk * nn + gn] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < BK; ks += 4) {
      const int kk = ks + (lane >> 4);
#pragma unroll
      for (int a = 0; a < 2; a++) {
        const double av = As[kk][wm + a * 16 + (lane & 15)];
#pragma unroll
        for (int b = 0; b < 2; b++) {
          const double bv = Bs[kk][wn + b * 16 + (lane & 15)];
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[a][b], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm + a * 16 + 4 * r + (lane >> 4);
        const int gn = c0 + wn + b * 16 + (lane & 15);
        if (gm < m && gn < nn) Ob[(int64_t)gm * nn + gn] = acc[a][b][r];
      }
}

// BF16x3: 4 waves, each a 32 x 32 quadrant = one v_mfma_f32_32x32x16_bf16 tile, K step 16.
// Operand layouts: A 32x16 -> lane l holds A[l%32][8*(l/32) + e], e = 0..7;
// B 16x32 -> lane l holds B[8*(l/32) + e][l%32]; D 32x32 -> lane l, register r holds
// D[8*(r/4) + 4*(l/32) + r%4][l%32].
__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

template <int TA>
__global__ __launch_bounds__(256) void k_bgemm_bf16x3(const float* __restrict__ A, const float* __restrict__ Cm,
                                                      float* __restrict__ out, int m, int nn, int kd) {
  __shared__ __bf16 Ah[BM][BK + 8], Al[BM][BK + 8];  // [m][k]
  __shared__ __bf16 Bh[BN][BK + 8], Bl[BN][BK + 8];  // [n][k]
  const float* Ab = A + (int64_t)blockIdx.z * m * kd;
  float* Ob = out + (int64_t)blockIdx.z * m * nn;
  const int m0 = blockIdx.y * BM, c0 = blockIdx.x * BN;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
  f16v acc = {};
  for (int k0 = 0; k0 < kd; k0 += BK) {
    for (int i = tid; i < BK * BM; i += 256) {
      const int mm = i / BK, kk = i % BK;
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < m && gk < kd) v = TA ? Ab[(int64_t)gk * m + gm] : Ab[(int64_t)gm * kd + gk];
      split_bf16(v, Ah[mm][kk], Al[mm][kk]);
    }
    for (int i = tid; i < BK * BN; i += 256) {
      const int kk = i / BN, c = i % BN;
      const int gn = c0 + c, gk = k0 + kk;
      const float v = (gn < nn && gk < kd) ? Cm[(int64_t)gk * nn + gn] : 0.f;
      split_bf16(v, Bh[c][kk], Bl[c][kk]);
    }
    __syncthreads();
    bf8 ah, al, bh, bl;
    const int ar = wm + (lane & 31), bc = wn + (lane & 31), kb = 8 * (lane >> 5);
#pragma unroll
    for (int e = 0; e < 8; e++) {
      ah[e] = Ah[ar][kb + e];
      al[e] = Al[ar][kb + e];
      bh[e] = Bh[bc][kb + e];
      bl[e] = Bl[bc][kb + e];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; r++) {
    const int gm = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int gn = c0 + wn + (lane & 31);
    if (gm < m && gn < nn) Ob[(int64_t)gm * nn + gn] = acc[r];
  }
}

__global__ void k_cvt_d2f(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}

static int gr(int64_t n, int b) { return (int)((n + b - 1) / b); }

static thread_local pqg_ctx* g_prof_ctx = nullptr;  // context of the running transform

template <typename T>
static void bgemm(const T* A, const T* Cm, T* out, int m, int nn, int kd, int nb, bool transA, hipStream_t st) {
  dim3 g(gr(nn, BN), gr(m, BM), nb);
  ProfScope prof(g_prof_ctx, st, std::is_same<T, double>::value ? "ao2mo_gemm_f64" : "ao2mo_gemm_bf16x3",
                 2ll * m * nn * kd * nb);  // units: flops of the batch
  if constexpr (std::is_same<T, double>::value) {
    if (transA)
      k_bgemm_f64<1><<<g, 256, 0, st>>>(A, Cm, out, m, nn, kd);
    else
      k_bgemm_f64<0><<<g, 256, 0, st>>>(A, Cm, out, m, nn, kd);
  } else {
    if (transA)
      k_bgemm_bf16x3<1><<<g, 256, 0, st>>>(A, Cm, out, m, nn, kd);
    else
      k_bgemm_bf16x3<0><<<g, 256, 0, st>>>(A, Cm, out, m, nn, kd);
  }
  PQG_HIP(hipGetLastError());
}

// C^T M_b C for nb symmetric nao x nao matrices M_b (row-major) and C nao x nmo (row-major),
// through U = M C (nao x nmo), T = U^T C (nmo x nmo)
template <typename T>
static void ctmc(const T* M, const T* Cm, T* U, T* Tout, int nao, int nmo, int nb, hipStream_t st) {
  bgemm<T>(M, Cm, U, nao, nmo, nao, nb, false, st);
  bgemm<T>(U, Cm, Tout, nmo, nmo, nao, nb, true, st);
}

// The AO integrals a transform reads: the 8-fold packed tensor of one basis (mixed == 0), or
// the npair(nao_a) x npair(nao_b) column-major block (mu nu|la si) of two bases, mu nu in basis
// a and la si in basis b (mixed == 1).
struct AoSource {
  const double* eri;
  int mixed;
};

// out4(idx2(i,j), idx2(k,l)) = (ij|kl), npair(nmo_a) x npair(nmo_b) column-major, with i, j over
// the columns of Ca (basis a) and k, l over those of Cb (basis b).  Half 1 transforms the pair of
// basis b of every AO pair P of basis a (one AO row of the source), B[P][(kl)]; half 2 transforms
// the pair of basis a of every column (kl) of B and writes output column (kl) contiguously.
template <typename T>
static void ao2mo_impl(pqg_ctx* ctx, AoSource src, int nao_a, int nmo_a, int nao_b, int nmo_b, const double* Ca,
                       const double* Cb, double* out4, hipStream_t st) {
  const int64_t pa_ao = (int64_t)nao_a * (nao_a + 1) / 2, pa_mo = (int64_t)nmo_a * (nmo_a + 1) / 2;
  const int64_t pb_mo = (int64_t)nmo_b * (nmo_b + 1) / 2;
  const int64_t n2 = (int64_t)std::max(nao_a, nao_b) * std::max(nao_a, nao_b);
  // batch size: keep the three work arrays (at most n2 elements per batch entry each) within
  // about 1/4 of what the budget leaves after the half-transformed B
  const uint64_t avail = ctx->arena->budget() > ctx->arena->used() ? ctx->arena->budget() - ctx->arena->used() : 0;
  const uint64_t interm = (uint64_t)pa_ao * pb_mo * sizeof(T);  // half-transformed B
  PQG_REQUIRE(avail > interm + (64ull << 20), PQG_ERR_OOM, "AO->MO intermediate does not fit the HBM budget");
  const int64_t most = std::max(pa_ao, pb_mo);
  int nb = (int)std::min<int64_t>(most, std::max<int64_t>(1, (int64_t)((avail - interm) / 4 / (3 * n2 * sizeof(T)))));
  nb = std::min(nb, 4096);
  DevBuf<T> Car(ctx, (size_t)nao_a * nmo_a), Cbr(ctx, (size_t)nao_b * nmo_b), X(ctx, n2 * nb), U(ctx, n2 * nb),
      Tm(ctx, n2 * nb), B(ctx, pa_ao * pb_mo);
  // C in row-major [mu][i] (input is column-major nao x nmo)
  for (int w = 0; w < 2; w++) {
    const int nao = w ? nao_b : nao_a, nmo = w ? nmo_b : nmo_a;
    std::vector<double> h((size_t)nao * nmo);
    PQG_HIP(hipMemcpyAsync(h.data(), w ? Cb : Ca, h.size() * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    std::vector<T> r(h.size());
    for (int mu = 0; mu < nao; mu++)
      for (int i = 0; i < nmo; i++) r[(size_t)mu * nmo + i] = (T)h[mu + (size_t)i * nao];
    PQG_HIP(hipMemcpyAsync((w ? Cbr : Car).get(), r.data(), r.size() * sizeof(T), hipMemcpyHostToDevice, st));
    PQG_HIP(hipStreamSynchronize(st));
  }
  const int64_t nb2 = (int64_t)nao_b * nao_b, na2 = (int64_t)nao_a * nao_a;
  // half 1: row P of B from AO row P of the source
  for (int64_t p0 = 0; p0 < pa_ao; p0 += nb) {
    const int b = (int)std::min<int64_t>(nb, pa_ao - p0);
    if (src.mixed)
      k_unpack_half<double, T><<<gr(nb2 * b, 256), 256, 0, st>>>(src.eri, pa_ao, nao_b, p0, b, X.get());
    else
      k_unpack8<T><<<gr(nb2 * b, 256), 256, 0, st>>>(src.eri, nao_b, p0, b, X.get());
    PQG_HIP(hipGetLastError());
    ctmc<T>(X.get(), Cbr.get(), U.get(), Tm.get(), nao_b, nmo_b, b, st);
    k_pack<T, T><<<gr(pb_mo * b, 256), 256, 0, st>>>(Tm.get(), nmo_b, b, B.get() + p0 * pb_mo, pb_mo);
    PQG_HIP(hipGetLastError());
  }
  // half 2: column (kl) of B holds (mu nu|kl); its transform over the pair of basis a is output
  // column (kl)
  for (int64_t c0 = 0; c0 < pb_mo; c0 += nb) {
    const int b = (int)std::min<int64_t>(nb, pb_mo - c0);
    k_unpack_half<T, T><<<gr(na2 * b, 256), 256, 0, st>>>(B.get(), pb_mo, nao_a, c0, b, X.get());
    PQG_HIP(hipGetLastError());
    ctmc<T>(X.get(), Car.get(), U.get(), Tm.get(), nao_a, nmo_a, b, st);
    k_pack<T, double><<<gr(pa_mo * b, 256), 256, 0, st>>>(Tm.get(), nmo_a, b, out4 + c0 * pa_mo, pa_mo);
    PQG_HIP(hipGetLastError());
  }
}

// Frozen-core operator (compute_frozen_core_ints, integral.cpp:565-698) of species a from the core
// density of species b, one output element (i, j) per thread:
//   out(i, j) += cj * sum_kl (ij|kl) Dc(k, l) - cx * sum_kl (ik|jl) Dx(k, l).
// eri: same species, the 8-fold packed AO integrals (pair index idx2 of the two pair indices);
// two species, the npair_a x npair_b column-major block (ij|kl) at idx2(i,j) + npair_a * idx2(k,l).
struct FcArgs {
  const double* eri;
  const double *dc, *dx;  // dx: exchange density, nullptr without exchange
  int na, nb, same;
  double cj, cx;
};

__device__ __forceinline__ int64_t fidx2(int64_t p, int64_t q) { return p >= q ? p * (p + 1) / 2 + q : q * (q + 1) / 2 + p; }

__device__ __forceinline__ double fc_eri(const FcArgs& a, int64_t pij, int64_t pkl) {
  return a.same ? a.eri[fidx2(pij, pkl)] : a.eri[pij + (int64_t)a.na * (a.na + 1) / 2 * pkl];
}

__global__ void k_fc_jk(FcArgs a, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= a.na) return;
  const int64_t pij = fidx2(i, j);
  double J = 0.0, K = 0.0;
  for (int l = 0; l < a.nb; l++)
    for (int k = 0; k < a.nb; k++) {
      J += a.dc[k + (int64_t)l * a.nb] * fc_eri(a, pij, fidx2(k, l));
      if (a.dx) K += a.dx[k + (int64_t)l * a.nb] * fc_eri(a, fidx2(i, k), fidx2(j, l));
    }
  out[i + (int64_t)j * a.na] += a.cj * J - a.cx * K;
}

}  // namespace pqg

using namespace pqg;

extern "C" pqg_status pqg_fc_ints(pqg_ctx* ctx, int nao_a, int nao_b, const double* eri, int same_particle, int a_spin,
                                  int b_spin, int restricted_b, double charge_prod, const double* dm_b, double* out) {
  return guard([&] {
    PQG_REQUIRE(ctx && eri && dm_b && out, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(nao_a >= 1 && nao_b >= 1, PQG_ERR_ARG, "empty basis");
    PQG_REQUIRE(!same_particle || nao_a == nao_b, PQG_ERR_ARG, "same species needs one basis");
    PQG_REQUIRE(a_spin >= 0 && a_spin < 2 && b_spin >= 0 && b_spin < 2, PQG_ERR_ARG, "spin index out of range");
    PQG_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t na = nao_a, nb = nao_b;
    const int64_t pa = na * (na + 1) / 2, pb = nb * (nb + 1) / 2;
    const int64_t neri = same_particle ? pa * (pa + 1) / 2 : pa * pb;
    const int64_t nd = nb * nb;
    // host densities -> the Coulomb density (2 D restricted, D_alpha + D_beta unrestricted) and
    // the exchange density of spin b
    std::vector<double> dc(nd), dx;
    for (int64_t q = 0; q < nd; q++) dc[q] = restricted_b ? 2.0 * dm_b[q] : dm_b[q] + dm_b[nd + q];
    const bool exchange = same_particle && a_spin == b_spin;
    if (exchange) dx.assign(dm_b + (restricted_b ? 0 : b_spin * nd), dm_b + (restricted_b ? 0 : b_spin * nd) + nd);
    DevBuf<double> de(ctx, neri), ddc(ctx, nd), ddx(ctx, exchange ? nd : 1), dout(ctx, na * na);
    PQG_HIP(hipMemcpyAsync(de.get(), eri, neri * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(ddc.get(), dc.data(), nd * 8, hipMemcpyHostToDevice, st));
    if (exchange) PQG_HIP(hipMemcpyAsync(ddx.get(), dx.data(), nd * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(dout.get(), out, na * na * 8, hipMemcpyHostToDevice, st));
    FcArgs a{};
    a.eri = de.get();
    a.dc = ddc.get();
    a.dx = exchange ? ddx.get() : nullptr;
    a.na = nao_a;
    a.nb = nao_b;
    a.same = same_particle ? 1 : 0;
    // same species: spinscale (1/2 when species b is unrestricted, whose two spin blocks each
    // add one call); two species: the charge product
    a.cj = same_particle ? (restricted_b ? 1.0 : 0.5) : charge_prod;
    a.cx = 1.0;
    k_fc_jk<<<dim3(gr(na, 64), (unsigned)na), 64, 0, st>>>(a, dout.get());
    PQG_HIP(hipGetLastError());
    PQG_HIP(hipMemcpyAsync(out, dout.get(), na * na * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  });
}

// nao_a, nmo_a / nao_b, nmo_b: the bases and MO counts of the row pair (Ca) and the column pair
// (Cb); mixed == 0: eri is the 8-fold packed tensor of one basis (nao_a == nao_b), mixed == 1: the
// npair(nao_a) x npair(nao_b) column-major block of two bases.
static void ao2mo_call(pqg_ctx* ctx, int nao_a, int nmo_a, int nao_b, int nmo_b, int mixed, const double* eri,
                       const double* Ca, const double* Cb, double* out4, int on_device, int mode, double* seconds) {
  PQG_REQUIRE(ctx && eri && Ca && Cb && out4, PQG_ERR_ARG, "NULL argument");
  PQG_REQUIRE(nao_a >= 1 && nao_b >= 1, PQG_ERR_ARG, "empty basis");
  PQG_REQUIRE(nmo_a >= 1 && nmo_a <= nao_a && nmo_b >= 1 && nmo_b <= nao_b, PQG_ERR_ARG,
              "need 1 <= nmo <= nao (the coefficient matrix is nao x nmo)");
  PQG_REQUIRE(mixed || nao_a == nao_b, PQG_ERR_ARG, "the 8-fold packed tensor holds one basis");
  PQG_REQUIRE(mode == PQG_AO2MO_FP64 || mode == PQG_AO2MO_BF16X3, PQG_ERR_ARG, "unknown mode");
  PQG_HIP(hipSetDevice(ctx->device));
  hipStream_t st = ctx->stream;
  const int64_t pa = (int64_t)nao_a * (nao_a + 1) / 2, pb = (int64_t)nao_b * (nao_b + 1) / 2;
  const int64_t neri = mixed ? pa * pb : pa * (pa + 1) / 2;
  const int64_t nout = (int64_t)nmo_a * (nmo_a + 1) / 2 * ((int64_t)nmo_b * (nmo_b + 1) / 2);
  const size_t nca = (size_t)nao_a * nmo_a, ncb = (size_t)nao_b * nmo_b;
  DevBuf<double> e8, ca, cb, o4;
  const double *deri = eri, *dA = Ca, *dB = Cb;
  double* dout = out4;
  if (!on_device) {
    e8.reset(ctx, neri);
    ca.reset(ctx, nca);
    cb.reset(ctx, ncb);
    o4.reset(ctx, nout);
    PQG_HIP(hipMemcpyAsync(e8.get(), eri, neri * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(ca.get(), Ca, nca * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(cb.get(), Cb, ncb * 8, hipMemcpyHostToDevice, st));
    deri = e8.get();
    dA = ca.get();
    dB = cb.get();
    dout = o4.get();
  }
  hipEvent_t t0, t1;
  PQG_HIP(hipEventCreate(&t0));
  PQG_HIP(hipEventCreate(&t1));
  PQG_HIP(hipEventRecord(t0, st));
  g_prof_ctx = ctx;
  const AoSource src{deri, mixed};
  if (mode == PQG_AO2MO_FP64)
    ao2mo_impl<double>(ctx, src, nao_a, nmo_a, nao_b, nmo_b, dA, dB, dout, st);
  else
    ao2mo_impl<float>(ctx, src, nao_a, nmo_a, nao_b, nmo_b, dA, dB, dout, st);
  PQG_HIP(hipEventRecord(t1, st));
  PQG_HIP(hipEventSynchronize(t1));
  float ms = 0.f;
  PQG_HIP(hipEventElapsedTime(&ms, t0, t1));
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (seconds) *seconds = ms * 1e-3;
  if (!on_device) {
    PQG_HIP(hipMemcpyAsync(out4, dout, nout * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  }
}

extern "C" pqg_status pqg_ao2mo(pqg_ctx* ctx, int nao, int nmo, const double* eri8, const double* C, double* out4,
                                int on_device, int mode, double* seconds) {
  return guard([&] { ao2mo_call(ctx, nao, nmo, nao, nmo, 0, eri8, C, C, out4, on_device, mode, seconds); });
}

extern "C" pqg_status pqg_ao2mo_pair(pqg_ctx* ctx, int nao, int nmo, const double* eri8, const double* Ca,
                                     const double* Cb, double* out4, int on_device, int mode, double* seconds) {
  return guard([&] { ao2mo_call(ctx, nao, nmo, nao, nmo, 0, eri8, Ca, Cb, out4, on_device, mode, seconds); });
}

extern "C" pqg_status pqg_ao2mo_mixed(pqg_ctx* ctx, int nao_a, int nmo_a, int nao_b, int nmo_b, const double* eri_ab,
                                      const double* Ca, const double* Cb, double* out4, int on_device, int mode,
                                      double* seconds) {
  return guard([&] {
    ao2mo_call(ctx, nao_a, nmo_a, nao_b, nmo_b, 1, eri_ab, Ca, Cb, out4, on_device, mode, seconds);
  });
}
