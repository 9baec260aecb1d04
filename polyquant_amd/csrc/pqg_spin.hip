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
// k_s2: one thread per determinant (internal order; entries outside a truncated space are zero
// and are skipped), string indices of the flipped strings from the colex rank tables, C_J
// gathered, block sums written per block and added on the host in a fixed order.
#include <algorithm>
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
    to_internal(ci, Cd, ldc, cin.get(), N, k, st);
    const int nblk = (int)std::min<int64_t>((N + 255) / 256, 4096);
    DevBuf<double> part(ci->ctx, (size_t)nblk * k * ci->nsp);
    for (int p = 0; p < ci->nsp; p++) {
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
    PQG_HIP(hipMemcpyAsync(h.data(), part.get(), h.size() * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    std::vector<double> out((size_t)k * ci->nsp);
    for (int p = 0; p < ci->nsp; p++)
      for (int v = 0; v < k; v++) {
        double s = 0.0;
        for (int b = 0; b < nblk; b++) s += h[((size_t)p * k + v) * nblk + b];
        out[v + (size_t)p * k] = s;
      }
    std::copy(out.begin(), out.end(), S2);
  });
}

}  // extern "C"
