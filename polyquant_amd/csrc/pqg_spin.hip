// pqg_spin.hip — <S^2> of CI vectors per species.
//
// evaluate_s2 (src/ci/detset/s2.cpp:5-130).  For determinant I with alpha string A and beta
// string B of a species, x = A ^ B are the unpaired orbitals, ua = x & A and ub = x & B:
//   diagonal      C_I^2 (|ua| + sz (sz - 1)),   sz = |ua| - |x| / 2              (s2.cpp:44-77)
//   off-diagonal  for i in ua, j in ub: J = (A - i + j, B - j + i, other strings of I),
//                 -phase_a phase_b C_I C_J                                       (s2.cpp:80-112)
// The reference enumerates every alpha single of A against every beta single of B and keeps
// the pairs with hole_a = part_b and hole_b = part_a; those are exactly the (i, j) above.
// phase_a counts the occupied orbitals of A strictly between i and j, phase_b those of B.
//
// Spin-raising form (used when its tables fit): with S+ = sum_i a+_{i alpha} a_{i beta},
// S^2 = S- S+ + Sz (Sz + 1), so <C|S^2|C> = |S+ C|^2 + Sz (Sz + 1) |C|^2 for any C (zero outside a
// truncated space), Sz = (N_alpha - N_beta) / 2.  An element of S+ C is indexed by an alpha
// string A+ with N_alpha + 1 orbitals, a beta core K with N_beta - 1 orbitals and the other
// strings rho:
//   (S+ C)[A+][K][rho] = sum_{i in A+, i not in K} (-1)^(|A+ below i| + |K below i|)
//                        C[A+ - i][K + i][rho]
// (the sign of a+_{i alpha} a_{i beta} on alpha-then-beta ordered strings, up to the constant
// (-1)^N_alpha).  About |A+ \ K| ~ 5 terms per element instead of |alpha \ beta| x |beta \ alpha|
// ~ 16 pairs per determinant, and the beta set is moved outermost so the gathers of one core
// fall in its own member rows.
//   k_s2p   one workgroup per (beta core K, vector), threads over (A+, rho)
//   k_norm2 |C|^2 per vector
// Direct form (k_s2, the fallback): one thread per determinant (internal order; entries outside
// a truncated space are zero and are skipped), string indices of the flipped strings from the
// colex rank tables, C_J gathered.  Block sums of both forms are added on the host in a fixed
// order.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "pqg_ci.hpp"
#include "pqg_strings.hpp"

namespace pqg {

struct S2Args {
  SlotDev sa, sb;            // alpha and beta string sets of the species
  const double* X;           // internal order, leading dimension ld per vector
  int64_t ld, n;             // n = internal product-space size
  int64_t stride_a, stride_b;
  double* part;              // [k][gridDim.x]
};

__device__ __forceinline__ double s2_wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// W: integer type of the address arithmetic (32-bit when the space has fewer than 2^32 entries)
template <typename W>
__global__ __launch_bounds__(256) void k_s2(S2Args a) {
  __shared__ double red[4];
  const int v = blockIdx.y;
  const double* X = a.X + (int64_t)v * a.ld;
  const W sta = (W)a.stride_a, stb = (W)a.stride_b, nsa = (W)a.sa.nstr, nsb = (W)a.sb.nstr;
  double acc = 0.0;
  for (int64_t I = (int64_t)blockIdx.x * 256 + threadIdx.x; I < a.n; I += (int64_t)gridDim.x * 256) {
    const double c = X[I];
    if (c == 0.0) continue;
    const int ia = (int)(((W)I / sta) % nsa), ib = (int)(((W)I / stb) % nsb);
    const uint64_t A = a.sa.bits[ia], B = a.sb.bits[ib];
    const uint64_t x = A ^ B, ua = x & A, ub = x & B;
    const int na = __popcll(ua), nt = __popcll(x);
    const double sz = na - 0.5 * nt;
    double t = c * (na + sz * (sz - 1.0));
    for (uint64_t qa = ua; qa; qa &= qa - 1) {
      const int i = ctz64(qa);
      for (uint64_t qb = ub; qb; qb &= qb - 1) {
        const int j = ctz64(qb);
        const uint64_t f = (1ull << i) | (1ull << j);
        const int ja = str_index(a.sa, a.sa.rtab, A ^ f), jb = str_index(a.sb, a.sb.rtab, B ^ f);
        if (ja < 0 || jb < 0) continue;
        const int64_t J = I + (int64_t)(ja - ia) * a.stride_a + (int64_t)(jb - ib) * a.stride_b;
        const uint64_t bm = between_mask(min(i, j), max(i, j));
        const int odd = (__popcll(A & bm) + __popcll(B & bm)) & 1;
        t += odd ? X[J] : -X[J];  // -phase_a phase_b C_J
      }
    }
    acc += c * t;
  }
  acc = s2_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) a.part[(int64_t)v * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

struct S2pArgs {
  SlotDev sa, sb;
  const double* X;          // beta set outermost, leading dimension ld per vector
  int64_t ld;
  int64_t stride_a, stride_b;  // strides in that layout
  int nother;
  int64_t ostr[2];          // strides and string counts of the other sets (slot order)
  int64_t onstr[2];
  int64_t R;                // product of the other sets' string counts
  const uint64_t* kb;       // beta cores (N_beta - 1 orbitals)
  const uint64_t* ap;       // alpha strings with N_alpha + 1 orbitals
  int64_t nap;
  int n;
  double* part;             // [k][number of cores]
};

__global__ __launch_bounds__(256) void k_s2p(S2pArgs a) {
  __shared__ int jbt[64];
  __shared__ int kpar[64];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const int64_t K = blockIdx.x;
  const int v = blockIdx.y;
  const uint64_t kb = a.kb[K];
  if (tid < a.n) {
    jbt[tid] = ((kb >> tid) & 1) ? -1 : str_index(a.sb, a.sb.rtab, kb | (1ull << tid));
    kpar[tid] = __popcll(kb & low_mask(tid)) & 1;
  }
  __syncthreads();
  const double* X = a.X + (int64_t)v * a.ld;
  const uint64_t free_k = ~kb & low_mask(a.n);
  double acc = 0.0;
  const int64_t total = a.nap * a.R;
  for (int64_t f = tid; f < total; f += 256) {
    int64_t api = f, off = 0;
    if (a.R > 1) {
      api = f / a.R;
      int64_t rho = f - api * a.R;
      if (a.nother == 2) {
        const int64_t i0 = rho / a.onstr[1];
        off = i0 * a.ostr[0] + (rho - i0 * a.onstr[1]) * a.ostr[1];
      } else {
        off = rho * a.ostr[0];
      }
    }
    const uint64_t A = a.ap[api];
    double val = 0.0;
    for (uint64_t q = A & free_k; q; q &= q - 1) {
      const int i = ctz64(q);
      const int jb = jbt[i];
      if (jb < 0) continue;
      const int ja = str_index(a.sa, a.sa.rtab, A ^ (1ull << i));
      if (ja < 0) continue;
      const double x = X[(int64_t)jb * a.stride_b + (int64_t)ja * a.stride_a + off];
      val += ((__popcll(A & low_mask(i)) + kpar[i]) & 1) ? -x : x;
    }
    acc += val * val;
  }
  acc = s2_wave_sum(acc);
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) a.part[(int64_t)v * gridDim.x + K] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void k_norm2(const double* __restrict__ X, int64_t ld, int64_t n,
                                               double* __restrict__ part) {
  __shared__ double red[4];
  const double* x = X + (int64_t)blockIdx.y * ld;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) acc += x[i] * x[i];
  acc = s2_wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// all k-subsets of n orbitals, ascending value (empty when k < 0 or k > n)
static std::vector<uint64_t> all_subsets(int n, int k) {
  std::vector<uint64_t> out;
  if (k < 0 || k > n) return out;
  const uint64_t top = n >= 64 ? ~0ull : (1ull << n);
  for (uint64_t K = k >= 64 ? ~0ull : (1ull << k) - 1; K < top;) {
    out.push_back(K);
    if (K == 0) break;
    const uint64_t c = K & (~K + 1), r = K + c;
    if (r == 0) break;
    K = (((r ^ K) >> 2) / c) | r;
  }
  return out;
}

static double binomd(int n, int k) {
  if (k < 0 || k > n) return 0.0;
  double v = 1.0;
  for (int i = 0; i < k; i++) v = v * (n - i) / (i + 1);
  return v;
}

// The spin-raising form applies unless its tables would exceed 2^26 entries (or the direct form
// is asked for).
static bool s2_raising_ok(const pqg_ci* ci, int p) {
  const Slot& sa = ci->slot[2 * p];
  const Slot& sb = ci->slot[2 * p + 1];
  const int n = sa.n_orb;
  return !std::getenv("PQG_S2_DIRECT") && binomd(n, sb.nocc - 1) <= double(1 << 26) &&
         binomd(n, sa.nocc + 1) <= double(1 << 26);
}

// |S+ C|^2 + Sz (Sz + 1) |C|^2 of species p for k internal-order vectors X (leading dimension
// ldx).  Internal (colex) order keeps the strings A+ - i of neighbouring A+ close together, so the
// gathers of a workgroup stay in few cache lines (reference-order indices made it 4x slower).
static void s2_raising(pqg_ci* ci, int p, const double* X, int64_t ldx, int k, double* out, hipStream_t st) {
  const Slot& sa = ci->slot[2 * p];
  const Slot& sb = ci->slot[2 * p + 1];
  const int n = sa.n_orb, na = sa.nocc, nb = sb.nocc;
  const int64_t N = ci->n_int;
  const int nblk_n = (int)std::min<int64_t>((N + 255) / 256, 1024);
  DevBuf<double> pn(ci->ctx, (size_t)nblk_n * k);
  k_norm2<<<dim3((unsigned)nblk_n, (unsigned)k), 256, 0, st>>>(X, ldx, N, pn.get());
  PQG_HIP(hipGetLastError());
  const std::vector<uint64_t> kbv = all_subsets(n, nb - 1), apv = all_subsets(n, na + 1);
  std::vector<double> hp;
  const int64_t nkb = (int64_t)kbv.size();
  if (nkb > 0 && !apv.empty()) {
    DevBuf<double> xt(ci->ctx, (size_t)N * k), part(ci->ctx, (size_t)nkb * k);
    DevBuf<uint64_t> dkb(ci->ctx, kbv.size()), dap(ci->ctx, apv.size());
    PQG_HIP(hipMemcpyAsync(dkb.get(), kbv.data(), kbv.size() * 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(dap.get(), apv.data(), apv.size() * 8, hipMemcpyHostToDevice, st));
    const int tb = 2 * p + 1;
    slot_outer(ci, tb, X, ldx, xt.get(), k, st);
    // strides with the beta set outermost: sets before it shrink by its string count
    S2pArgs a{};
    a.sa = sa.dev();
    a.sb = sb.dev();
    a.X = xt.get();
    a.ld = N;
    a.stride_b = N / sb.nstr;
    a.stride_a = ci->stride[2 * p] / sb.nstr;  // alpha comes before beta in slot order
    a.R = 1;
    for (int t = 0; t < ci->nslot; t++) {
      if (t == 2 * p || t == tb) continue;
      a.ostr[a.nother] = t < tb ? ci->stride[t] / sb.nstr : ci->stride[t];
      a.onstr[a.nother] = ci->slot[t].nstr;
      a.R *= ci->slot[t].nstr;
      a.nother++;
    }
    a.kb = dkb.get();
    a.ap = dap.get();
    a.nap = (int64_t)apv.size();
    a.n = n;
    a.part = part.get();
    PQG_REQUIRE(nkb < (1ll << 31) && a.nother <= 2 && n <= 64, PQG_ERR_UNSUPPORTED, "S^2 tables out of range");
    PQG_REQUIRE(a.X && a.part && a.kb && a.ap && a.sa.bits && a.sb.bits, PQG_ERR_HIP, "S^2 buffers not allocated");
    {
      ProfScope prof(ci->ctx, st, "s2", k);
      k_s2p<<<dim3((unsigned)nkb, (unsigned)k), 256, 0, st>>>(a);
      PQG_HIP(hipGetLastError());
    }
    hp.resize((size_t)nkb * k);
    PQG_HIP(hipMemcpyAsync(hp.data(), part.get(), hp.size() * 8, hipMemcpyDeviceToHost, st));
  }
  std::vector<double> hn((size_t)nblk_n * k);
  PQG_HIP(hipMemcpyAsync(hn.data(), pn.get(), hn.size() * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  const double sz = 0.5 * (na - nb);
  for (int v = 0; v < k; v++) {
    double s = 0.0, c2 = 0.0;
    for (int64_t b = 0; b < (int64_t)hp.size() / k; b++) s += hp[(size_t)v * nkb + b];
    for (int b = 0; b < nblk_n; b++) c2 += hn[(size_t)v * nblk_n + b];
    out[v + (size_t)p * k] = s + sz * (sz + 1.0) * c2;
  }
}

}  // namespace pqg

using namespace pqg;

extern "C" {

pqg_status pqg_ci_s2(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S2, int on_device) {
  return guard([&] {
    PQG_REQUIRE(ci && C && S2 && k >= 1 && k <= 65535, PQG_ERR_ARG, "invalid argument");
    PQG_REQUIRE(ldc >= ci->n_ref, PQG_ERR_ARG, "leading dimension smaller than N");
    hipStream_t st = ci->stream;
    const int64_t N = ci->n_int, Nr = ci->n_ref;
    DevBuf<double> cin(ci->ctx, (size_t)N * k), cref;
    const double* Cd = C;
    if (!on_device) {
      cref.reset(ci->ctx, (size_t)Nr * k);
      PQG_HIP(hipMemcpy2DAsync(cref.get(), Nr * 8, C, ldc * 8, Nr * 8, k, hipMemcpyHostToDevice, st));
      Cd = cref.get();
      ldc = Nr;
    }
    std::vector<double> out((size_t)k * ci->nsp);
    std::vector<int> direct;
    for (int p = 0; p < ci->nsp; p++)
      if (!s2_raising_ok(ci, p)) direct.push_back(p);
    to_internal(ci, Cd, ldc, cin.get(), N, k, st);
    for (int p = 0; p < ci->nsp; p++)
      if (s2_raising_ok(ci, p)) s2_raising(ci, p, cin.get(), N, k, out.data(), st);
    const int nblk = (int)std::min<int64_t>((N + 255) / 256, 4096);
    DevBuf<double> part;
    if (!direct.empty()) part.reset(ci->ctx, (size_t)nblk * k * ci->nsp);
    for (int p : direct) {
      S2Args a{};
      a.sa = ci->slot[2 * p].dev();
      a.sb = ci->slot[2 * p + 1].dev();
      a.X = cin.get();
      a.ld = N;
      a.n = N;
      a.stride_a = ci->stride[2 * p];
      a.stride_b = ci->stride[2 * p + 1];
      a.part = part.get() + (size_t)p * nblk * k;
      PQG_REQUIRE(a.X && a.part && a.sa.bits && a.sb.bits && a.sa.rtab && a.sb.rtab, PQG_ERR_HIP,
                  "S^2 buffers not allocated");
      ProfScope prof(ci->ctx, st, "s2", k);
      if (N < (1ll << 32))
        k_s2<uint32_t><<<dim3((unsigned)nblk, (unsigned)k), 256, 0, st>>>(a);
      else
        k_s2<uint64_t><<<dim3((unsigned)nblk, (unsigned)k), 256, 0, st>>>(a);
      PQG_HIP(hipGetLastError());
    }
    std::vector<double> h((size_t)nblk * k * ci->nsp);
    if (!direct.empty()) {
      PQG_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * 8, hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
    }
    for (int p : direct)
      for (int v = 0; v < k; v++) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += h[((size_t)p * k + v) * nblk + b];
        out[v + (size_t)p * k] = s;
      }
    std::copy(out.begin(), out.end(), S2);
  });
}

}  // extern "C"
