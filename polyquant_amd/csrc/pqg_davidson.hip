// Davidson on the GPU: block of nev Ritz pairs, diagonal preconditioner, Rayleigh-Ritz in FP64.
//
// Restates the behaviour of the eigensolver the reference drives at src/ci/epci.cpp:589-621
// (Spectra DavidsonSymEigsSolver, SortRule::SmallestAlge): start from unit vectors at the
// nvec_init smallest diagonal entries, extend by t_k = r_k / (theta_k - diag) (with a perturbed
// start only for the roots whose residual is not yet below tol), restart on the current Ritz
// vectors when the basis would exceed nvec_max, converge when every one of the nev residual
// norms is below tol.
//
// Everything stays in HBM: tall-skinny products (V^T W, V Y, Gram-Schmidt) are FP64 kernels
// with a deterministic two-level reduction, and the projected eigenproblem is solved by a
// one-workgroup parallel-order Jacobi kernel in FP64.  Only the nev residual norms travel to
// the host per iteration (convergence decision).
//
// Mixed precision (PQG_MIXED): sigma of new search directions runs in FP32.  When the
// FP32-based residual has dropped below max(tol, min(1e-3, 0.05 * last true residual)), or
// stalls at the level where the FP32 error can hold it, sigma of the Ritz vectors is
// recomputed in FP64 ("true" residual).  Converged if that is below tol; otherwise the basis
// restarts on the Ritz vectors with their exact FP64 sigma (defect correction) and extends with
// corrections built from the true residuals.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "pqg_davidson.hpp"

namespace pqg {

constexpr int kTB = 256;

// out[a][b] = sum over the chunks of the partials in a fixed order (one block per output)
__global__ __launch_bounds__(256) void k_tn_reduce(const double* __restrict__ part, int nchunk, int na, int nb,
                                                   double* __restrict__ out, int ldo, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;  // a gated product that is not needed (uniform over the grid)
  const int o = blockIdx.x, G = na * nb;
  double v = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += 256) v += part[(size_t)c * G + o];
  __shared__ double red[4];
  for (int s = 32; s > 0; s >>= 1) v += __shfl_xor(v, s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[(o / nb) + (o % nb) * ldo] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Partials of A^T B for tall-skinny A (nrows x na) and B (nrows x nb), column-major, every column
// 16-byte aligned (even leading dimensions, aligned bases).  Block b sums the rows
// [b rpb, (b + 1) rpb); each of its waves owns a task of 8 columns of A x 4 columns of B
// (blockIdx.y selects further groups of four tasks), so a lane keeps 32 accumulators and takes
// two consecutive rows per step with 16-byte loads: many waves per CU, each with twelve 1 KB
// loads in flight, which keeps HBM busy (a B column read by the several waves of a block comes
// from L2 after the first).  The wave's sums meet by shuffles; part[b][a][c] is summed over b in
// a fixed order afterwards: deterministic.
constexpr int kTnQA = 8;
__global__ __launch_bounds__(256) void k_tn_part(const double* __restrict__ A, int64_t lda, int na,
                                                 const double* __restrict__ B, int64_t ldb, int nb, int64_t nrows,
                                                 int64_t rpb, double* __restrict__ part,
                                                 const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x & 63;
  const int task = (threadIdx.x >> 6) + blockIdx.y * (blockDim.x >> 6);
  const int ta = (na + kTnQA - 1) / kTnQA;
  if (task >= ta * ((nb + 3) / 4)) return;  // idle waves of the last task group
  const int a0 = (task % ta) * kTnQA, b0 = (task / ta) * 4;
  const int nal = min(kTnQA, na - a0), nbl = min(4, nb - b0);
  const int64_t r0 = (int64_t)blockIdx.x * rpb, r1 = min(nrows, r0 + rpb);
  double acc[kTnQA][4];
#pragma unroll
  for (int i = 0; i < kTnQA; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  for (int64_t r = r0 + 2 * lane; r < r1; r += 128) {
    const bool full = r + 1 < r1;
    double2 y[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      y[j] = make_double2(0.0, 0.0);
      if (j < nbl) {
        const double* q = B + r + (int64_t)(b0 + j) * ldb;
        y[j] = full ? *reinterpret_cast<const double2*>(q) : make_double2(q[0], 0.0);
      }
    }
#pragma unroll
    for (int i = 0; i < kTnQA; i++) {
      if (i < nal) {
        const double* q = A + r + (int64_t)(a0 + i) * lda;
        const double2 x = full ? *reinterpret_cast<const double2*>(q) : make_double2(q[0], 0.0);
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += x.x * y[j].x + x.y * y[j].y;
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kTnQA; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double v = acc[i][j];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      if (lane == 0 && i < nal && j < nbl) part[((int64_t)blockIdx.x * na + a0 + i) * nb + b0 + j] = v;
    }
}

// out[r][j] = beta * out[r][j] + alpha * sum_c A[r][c] Bs[c][j]   (Bs small, ka x nb, in LDS).
// A thread takes two consecutive rows (16-byte loads and stores; columns 16-byte aligned) and
// up to NB output columns at once; wider products run in column chunks of NB.  blockIdx.y
// selects (A0, out0) or (A1, out1), so two products with the same Bs share a launch.  A row is
// read completely before its outputs are written, so out may alias A when nb <= NB.
template <int NB>
__global__ __launch_bounds__(kTB) void k_nn(const double* A0, const double* A1, int64_t lda, int ka,
                                            const double* __restrict__ Bs, int ldb, int nb, double* out0,
                                            double* out1, int64_t ldo, int64_t nrows, double alpha, double beta,
                                            const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  __shared__ double bs[64 * NB];
  const int64_t r = ((int64_t)blockIdx.x * kTB + threadIdx.x) * 2;
  const double* A = blockIdx.y ? A1 : A0;
  double* out = blockIdx.y ? out1 : out0;
  const bool full = r + 1 < nrows;
  for (int j0 = 0; j0 < nb; j0 += NB) {
    const int nbl = min(NB, nb - j0);
    __syncthreads();
    for (int i = threadIdx.x; i < ka * NB; i += kTB) {
      const int c = i / NB, j = i % NB;
      bs[i] = j < nbl ? Bs[c + (j0 + j) * ldb] : 0.0;
    }
    __syncthreads();
    if (r >= nrows) continue;
    double2 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) acc[j] = make_double2(0.0, 0.0);
#pragma unroll 8
    for (int c = 0; c < ka; c++) {
      const double* q = A + r + c * lda;
      const double2 a = full ? *reinterpret_cast<const double2*>(q) : make_double2(q[0], 0.0);
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const double b = bs[c * NB + j];
        acc[j].x += a.x * b;
        acc[j].y += a.y * b;
      }
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (j >= nbl) break;
      double* o = out + r + (int64_t)(j0 + j) * ldo;
      if (full) {
        double2 v = make_double2(alpha * acc[j].x, alpha * acc[j].y);
        if (beta != 0.0) {
          const double2 q = *reinterpret_cast<const double2*>(o);
          v.x += beta * q.x;
          v.y += beta * q.y;
        }
        *reinterpret_cast<double2*>(o) = v;
      } else {
        o[0] = (beta == 0.0 ? 0.0 : beta * o[0]) + alpha * acc[j].x;
      }
    }
  }
}

// R = AX - X theta, then T = R / (theta - diag) (safeguarded); invalid entries (diag >= 1e29)
// give zero corrections.
__global__ void k_residual(const double* __restrict__ X, const double* __restrict__ AX, const double* __restrict__ theta,
                           const double* __restrict__ Hd, double* __restrict__ R, double* __restrict__ T, int64_t n,
                           int64_t ld, int nev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = Hd[i];
  for (int k = 0; k < nev; k++) {
    const double r = AX[i + k * ld] - theta[k] * X[i + k * ld];
    R[i + k * ld] = r;
    double den = theta[k] - d;
    if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
    T[i + k * ld] = (d >= 1e29) ? 0.0 : r / den;
  }
}

// Ritz vectors, residuals and corrections in one pass over the basis (nev <= NEV):
//   X = V Y, AX = W Y (Y: the first nev eigenvectors of the projected matrix, in LDS),
//   r = AX - X theta, t = r / (theta - diag) (safeguarded, zero on invalid entries),
// writing X and T and summing |r_k|^2, |t_k|^2 per block (part[block][2 NEV], fixed order); AX
// and R are never stored.  A thread takes two consecutive rows per step (16-byte accesses).
template <int NEV>
__global__ __launch_bounds__(kTB) void k_ritz_res(const double* __restrict__ V, const double* __restrict__ W,
                                                  int64_t ld, int m, const double* __restrict__ Y,
                                                  const double* __restrict__ theta, const double* __restrict__ Hd,
                                                  int nev, double* __restrict__ X, double* __restrict__ T,
                                                  int64_t nrows, double* __restrict__ part) {
  __shared__ double ys[64 * NEV];
  __shared__ double th[NEV];
  __shared__ double red[kTB / 64][2 * NEV];
  const int tid = threadIdx.x;
  for (int i = tid; i < m * NEV; i += kTB) {
    const int c = i / NEV, k = i % NEV;
    ys[i] = k < nev ? Y[c + k * 64] : 0.0;
  }
  if (tid < NEV) th[tid] = tid < nev ? theta[tid] : 0.0;
  __syncthreads();
  double ss[2 * NEV];
#pragma unroll
  for (int j = 0; j < 2 * NEV; j++) ss[j] = 0.0;
  for (int64_t r = ((int64_t)blockIdx.x * kTB + tid) * 2; r < nrows; r += (int64_t)gridDim.x * kTB * 2) {
    const bool full = r + 1 < nrows;
    double2 x[NEV], ax[NEV];
#pragma unroll
    for (int k = 0; k < NEV; k++) {
      x[k] = make_double2(0.0, 0.0);
      ax[k] = make_double2(0.0, 0.0);
    }
#pragma unroll 4
    for (int c = 0; c < m; c++) {
      const double* pv = V + r + c * ld;
      const double* pw = W + r + c * ld;
      const double2 a = full ? *reinterpret_cast<const double2*>(pv) : make_double2(pv[0], 0.0);
      const double2 b = full ? *reinterpret_cast<const double2*>(pw) : make_double2(pw[0], 0.0);
#pragma unroll
      for (int k = 0; k < NEV; k++) {
        const double y = ys[c * NEV + k];
        x[k].x += a.x * y;
        x[k].y += a.y * y;
        ax[k].x += b.x * y;
        ax[k].y += b.y * y;
      }
    }
    const double d0 = Hd[r], d1 = full ? Hd[r + 1] : 0.0;
#pragma unroll
    for (int k = 0; k < NEV; k++) {
      if (k >= nev) break;
      const double r0 = ax[k].x - th[k] * x[k].x, r1 = full ? ax[k].y - th[k] * x[k].y : 0.0;
      double e0 = th[k] - d0, e1 = th[k] - d1;
      if (fabs(e0) < 1e-8) e0 = e0 < 0 ? -1e-8 : 1e-8;
      if (fabs(e1) < 1e-8) e1 = e1 < 0 ? -1e-8 : 1e-8;
      const double t0 = d0 >= 1e29 ? 0.0 : r0 / e0, t1 = (!full || d1 >= 1e29) ? 0.0 : r1 / e1;
      ss[k] += r0 * r0 + r1 * r1;
      ss[NEV + k] += t0 * t0 + t1 * t1;
      double* px = X + r + k * ld;
      double* pt = T + r + k * ld;
      if (full) {
        *reinterpret_cast<double2*>(px) = x[k];
        *reinterpret_cast<double2*>(pt) = make_double2(t0, t1);
      } else {
        px[0] = x[k].x;
        pt[0] = t0;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 2 * NEV; j++) {
    double v = ss[j];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((tid & 63) == 0) red[tid >> 6][j] = v;
  }
  __syncthreads();
  if (tid < 2 * NEV) {
    double v = 0.0;
    for (int w = 0; w < kTB / 64; w++) v += red[w][tid];
    part[(int64_t)blockIdx.x * 2 * NEV + tid] = v;
  }
}

// |r_k|^2 and |t_k|^2 summed over the blocks in a fixed order, into the diagonal of the [R T]
// Gram (rows k and nev + k, leading dimension 64)
__global__ __launch_bounds__(256) void k_ritz_res_reduce(const double* __restrict__ part, int nblk, int nev, int nevt,
                                                         double* __restrict__ gram) {
  const int j = blockIdx.x;  // 0 .. 2 nevt - 1
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) v += part[(int64_t)b * 2 * nevt + j];
  __shared__ double red[4];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int k = j % nevt;
    if (k < nev) {
      const int row = j < nevt ? k : nev + k;
      gram[row + row * 64] = (red[0] + red[1]) + (red[2] + red[3]);
    }
  }
}

__global__ void k_d2f_blk(const double* __restrict__ x, float* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (float)x[i];
}
__global__ void k_f2d_blk(const float* __restrict__ x, double* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (double)x[i];
}
__global__ void k_unit(double* __restrict__ v, int64_t n, int64_t ld, const int64_t* __restrict__ pos, int k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < k; j++) v[i + j * ld] = (i == pos[j]) ? 1.0 : 0.0;
}
// unit vectors at pos[j] plus eps * u(i, j), u uniform in [-1, 1) from a SplitMix64 hash of
// (j, i): a start that overlaps every eigenvector
__global__ void k_unit_perturbed(double* __restrict__ v, int64_t n, int64_t ld, const int64_t* __restrict__ pos, int k,
                                 double eps) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < k; j++) {
    uint64_t z = (uint64_t)i * 64ull + (uint64_t)j + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    v[i + j * ld] = ((i == pos[j]) ? 1.0 : 0.0) + eps * u;
  }
}
__global__ void k_scale_cols(double* __restrict__ v, int64_t n, int64_t ld, const double* __restrict__ s, int k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < k; j++) v[i + j * ld] *= s[j];
}

__device__ __forceinline__ double wave_sum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int tid = threadIdx.x;
  __syncthreads();
  if ((tid & 63) == 0) red[tid >> 6] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
  return t;
}
// Rotation (c, s) that annihilates the coupling apq of a pair with diagonal entries app, aqq
// (Rutishauser's form): t = sgn(th) / (|th| + sqrt(th^2 + 1)), th = (aqq - app) / (2 apq),
// c = 1 / sqrt(1 + t^2), s = t c.  The FP64 division and square root are long instruction
// sequences on the eigensolve's critical path, so the hardware reciprocal and reciprocal square
// root are refined by two Newton steps each instead; c is then accurate to rounding and the
// rotation orthogonal.
__device__ __forceinline__ double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ double rsq_nr(double y) {  // y >= 1
  double r = __builtin_amdgcn_rsq(y);
  r = r * fma(-0.5 * y * r, r, 1.5);
  r = r * fma(-0.5 * y * r, r, 1.5);
  return r;
}
__device__ __forceinline__ void jac_rot(double app, double aqq, double apq, double& c, double& s) {
  const double th = (aqq - app) * rcp_nr(2.0 * apq);
  const double ath = fabs(th);
  double t;
  if (ath > 1e100) {
    t = 0.5 * rcp_nr(th);
  } else {
    const double y = fma(th, th, 1.0);
    const double sq = y * rsq_nr(y);
    t = rcp_nr(ath + sq);
    if (th < 0) t = -t;
  }
  c = rsq_nr(fma(t, t, 1.0));
  s = t * c;
}
// where position j of the round robin moves between steps (slot k = positions 2k, 2k+1; the
// first of slot 0 stays, the other firsts move one slot down, the seconds one slot up)
__device__ __forceinline__ int rr_next(int mm, int j) {
  const int half = mm >> 1, k = j >> 1;
  if (mm == 2 || j == 0) return j;
  if (j == 1) return 3;
  if ((j & 1) == 0) return k == 1 ? 1 : j - 2;
  return k == half - 1 ? j - 1 : j + 2;
}

// 64-bit DPP lane shifts across the wave: dn(x)[l] = x[l + 1], up(x)[l] = x[l - 1] (the lanes at
// the ends keep their own value; callers never use those)
__device__ __forceinline__ double dpp_dn(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const int lo = (int)(u & 0xffffffffu), hi = (int)(u >> 32);
  const int l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x130, 0xf, 0xf, false);
  const int h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2);
}
__device__ __forceinline__ double dpp_up(double x) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const int lo = (int)(u & 0xffffffffu), hi = (int)(u >> 32);
  const int l2 = __builtin_amdgcn_update_dpp(lo, lo, 0x138, 0xf, 0xf, false);
  const int h2 = __builtin_amdgcn_update_dpp(hi, hi, 0x138, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)h2 << 32) | (unsigned)l2);
}

// Subspace eigensolve of the Rayleigh-Ritz step: parallel-order (round-robin) cyclic Jacobi for a
// symmetric m x m matrix, m <= 64, in one workgroup, warm-started: with P = blockdiag(Yp (warm x
// warm), I) it diagonalises P^T G P and accumulates the rotations onto P, so a projected matrix
// that grew by a few rows since the previous iteration is diagonal but for those rows.
//
// The matrix is kept by round-robin position, not by index: slot k is positions (2k, 2k + 1),
// the indices at a slot's two positions form one of the step's m/2 disjoint pairs, and between
// steps every position moves by one fixed permutation (rr_next), which walks each index past
// every other once per mm - 1 steps.  Thread (bi, bj) owns the 2x2 block of rows of slot bi and
// columns of slot bj: it reads the block as two 16-byte loads from the current buffer, rotates
// it from both sides and writes its entries straight to their next-step positions in the other
// buffer.  V's rows stay with their threads in registers (rows 2 bi, 2 bi + 1 at the columns of
// slot bj): a step rotates them and moves them one slot along the wave by DPP lane shifts, so V
// never goes through LDS during the sweeps.  A pair whose coupling is below the sweep's
// threshold keeps c = 1, s = 0.  Outputs ascending eigenvalues and the matching eigenvectors
// (columns of Y).
constexpr int kJacT = 1024;
constexpr int kJacLd = 66;  // row stride (doubles): 16-byte aligned rows
__global__ __launch_bounds__(kJacT) void k_jacobi(double* __restrict__ G, int m, int old,
                                                  const double* __restrict__ gt, const double* __restrict__ Yp, int warm,
                                                  double* __restrict__ theta, double* __restrict__ Y,
                                                  double* __restrict__ Ynext, double rel_tol2) {
  __shared__ __attribute__((aligned(16))) double Ab[2][64][kJacLd];
  __shared__ __attribute__((aligned(16))) double Vb[1][64][kJacLd];
  __shared__ double cs[32], sn[32];
  __shared__ int rot[32];
  __shared__ int pidx[2][64];
  __shared__ double red[kJacT / 64];
  __shared__ int order[64];
  const int tid = threadIdx.x;
  const int mm = m + (m & 1), half = mm / 2;
  for (int i = tid; i < 64 * 64; i += kJacT) {
    const int r = i / 64, c = i % 64;
    double v = (r == c) ? 1.0 : 0.0;
    if (r < warm && c < warm) v = Yp[r + c * 64];
    Vb[0][r][c] = v;
    double g = 0.0;
    if (r < m && c < m) {
      if (r < old && c < old)
        g = 0.5 * (G[r + c * 64] + G[c + r * 64]);
      else if (r < old)
        g = gt ? gt[(c - old) + r * 64] : G[r + c * 64];
      else if (c < old)
        g = gt ? gt[(r - old) + c * 64] : G[c + r * 64];
      else
        g = 0.5 * (G[r + c * 64] + G[c + r * 64]);
    }
    Ab[0][r][c] = g;
  }
  if (tid < 64) pidx[0][tid] = tid;
  if (tid < 32) {
    cs[tid] = 1.0;
    sn[tid] = 0.0;
    rot[tid] = 0;
  }
  __syncthreads();
  for (int i = tid; i < m * m; i += kJacT) {
    const int r = i / m, c = i % m;
    G[r + c * 64] = Ab[0][r][c];
  }
  if (warm > 0) {
    // A = P^T A P in 2x2 register tiles (Ab[1] holds A P); sums over k < m keep A zero outside m
    const int r0 = 2 * (tid >> 5), c0 = 2 * (tid & 31);
    double a00 = 0.0, a01 = 0.0, a10 = 0.0, a11 = 0.0;
    for (int k = 0; k < m; k++) {
      const double x0 = Ab[0][r0][k], x1 = Ab[0][r0 + 1][k], y0 = Vb[0][k][c0], y1 = Vb[0][k][c0 + 1];
      a00 = fma(x0, y0, a00);
      a01 = fma(x0, y1, a01);
      a10 = fma(x1, y0, a10);
      a11 = fma(x1, y1, a11);
    }
    Ab[1][r0][c0] = a00;
    Ab[1][r0][c0 + 1] = a01;
    Ab[1][r0 + 1][c0] = a10;
    Ab[1][r0 + 1][c0 + 1] = a11;
    __syncthreads();  // A P complete; every read of Ab[0] is done
    a00 = a01 = a10 = a11 = 0.0;
    for (int k = 0; k < m; k++) {
      const double x0 = Vb[0][k][r0], x1 = Vb[0][k][r0 + 1], y0 = Ab[1][k][c0], y1 = Ab[1][k][c0 + 1];
      a00 = fma(x0, y0, a00);
      a01 = fma(x0, y1, a01);
      a10 = fma(x1, y0, a10);
      a11 = fma(x1, y1, a11);
    }
    Ab[0][r0][c0] = a00;
    Ab[0][r0][c0 + 1] = a01;
    Ab[0][r0 + 1][c0] = a10;
    Ab[0][r0 + 1][c0 + 1] = a11;
  }
  __syncthreads();
  int ca = 0;
  const int bi = tid >> 5, bj = tid & 31;
  const bool blk = bi < half && bj < half;
  const bool dgb = bi == bj;
  // fixed per thread: where this block's rows and columns go after a step, as offsets (doubles)
  // into a buffer; buffer cb is at + kBuf * cb
  constexpr int kBuf = 64 * kJacLd;
  const int rd0 = (2 * bi) * kJacLd + 2 * bj, rd1 = rd0 + kJacLd;
  int wA00 = 0, wA01 = 0, wA10 = 0, wA11 = 0, wV00 = 0, wV01 = 0, wV10 = 0, wV11 = 0, Q0 = 0, Q1 = 0;
  if (blk) {
    const int P0 = rr_next(mm, 2 * bi), P1 = rr_next(mm, 2 * bi + 1);
    Q0 = rr_next(mm, 2 * bj);
    Q1 = rr_next(mm, 2 * bj + 1);
    wA00 = P0 * kJacLd + Q0;
    wA01 = P0 * kJacLd + Q1;
    wA10 = P1 * kJacLd + Q0;
    wA11 = P1 * kJacLd + Q1;
    wV00 = (2 * bi) * kJacLd + Q0;
    wV01 = (2 * bi) * kJacLd + Q1;
    wV10 = (2 * bi + 1) * kJacLd + Q0;
    wV11 = (2 * bi + 1) * kJacLd + Q1;
  }
  double* const Abase = &Ab[0][0][0];
  // this thread's V entries: rows 2 bi, 2 bi + 1 at the positions of slot bj, in registers; a
  // step rotates them and moves them one slot along the wave
  double vx0 = Vb[0][2 * bi][2 * bj], vy0 = Vb[0][2 * bi][2 * bj + 1];
  double vx1 = Vb[0][2 * bi + 1][2 * bj], vy1 = Vb[0][2 * bi + 1][2 * bj + 1];
  for (int sweep = 0; sweep < 40; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int i = tid; i < mm * mm; i += kJacT) {
      const int r = i / mm, c = i % mm;
      const double a = Ab[ca][r][c];
      if (r != c)
        off += a * a;
      else
        dia += a * a;
    }
    const double offs = block_sum(off, red);
    __syncthreads();
    const double dias = block_sum(dia, red);
    if (offs <= rel_tol2 * dias || offs < 1e-300) break;
    const double thr2 = fmax(rel_tol2 * dias / ((double)m * (double)m), 1e-300);
    for (int step = 0; step < mm - 1; step++) {
      __syncthreads();  // the previous step's buffers are complete
      if (tid < half) {
        const int p = 2 * tid, q = p + 1;
        const double apq = Ab[ca][p][q], app = Ab[ca][p][p], aqq = Ab[ca][q][q];
        double c = 1.0, s = 0.0;
        int r = 0;
        if (pidx[ca][p] < m && pidx[ca][q] < m && apq * apq > thr2) {
          jac_rot(app, aqq, apq, c, s);
          r = 1;
        }
        cs[tid] = c;
        sn[tid] = s;
        rot[tid] = r;
      }
      __syncthreads();
      const double ci = cs[bi & 31], si = sn[bi & 31], cj = cs[bj], sj = sn[bj];
      if (blk) {
        const double* A0 = Abase + ca * kBuf;
        double* A1 = Abase + (ca ^ 1) * kBuf;
        const double2 a0 = *reinterpret_cast<const double2*>(A0 + rd0);
        const double2 a1 = *reinterpret_cast<const double2*>(A0 + rd1);
        // new column p = c*col p - s*col q, new column q = s*col p + c*col q (same for rows);
        // pairs left alone have c = 1, s = 0, which reproduces their entries exactly
        const double r00 = ci * a0.x - si * a1.x, r01 = ci * a0.y - si * a1.y;
        const double r10 = si * a0.x + ci * a1.x, r11 = si * a0.y + ci * a1.y;
        double n01 = sj * r00 + cj * r01, n10 = cj * r10 - sj * r11;
        if (dgb && rot[bi]) n01 = n10 = 0.0;  // the rotated pair's coupling is annihilated exactly
        A1[wA00] = cj * r00 - sj * r01;
        A1[wA01] = n01;
        A1[wA10] = n10;
        A1[wA11] = sj * r10 + cj * r11;
        if (bi == 0) {
          pidx[ca ^ 1][Q0] = pidx[ca][2 * bj];
          pidx[ca ^ 1][Q1] = pidx[ca][2 * bj + 1];
        }
      }
      {
        // V: rotate the columns of slot bj, then move to the next step's positions: the first
        // of slot bj comes from slot bj + 1's first (slot 0 keeps its own, the last slot takes
        // its own second), the second from slot bj - 1's second (slot 0 takes slot 1's first)
        const double x0 = cj * vx0 - sj * vy0, y0 = sj * vx0 + cj * vy0;
        const double x1 = cj * vx1 - sj * vy1, y1 = sj * vx1 + cj * vy1;
        const double dx0 = dpp_dn(x0), dx1 = dpp_dn(x1), uy0 = dpp_up(y0), uy1 = dpp_up(y1);
        if (half == 1) {
          vx0 = x0; vy0 = y0; vx1 = x1; vy1 = y1;
        } else {
          vx0 = bj == 0 ? x0 : (bj == half - 1 ? y0 : dx0);
          vx1 = bj == 0 ? x1 : (bj == half - 1 ? y1 : dx1);
          vy0 = bj == 0 ? dx0 : uy0;
          vy1 = bj == 0 ? dx1 : uy1;
        }
      }
      ca ^= 1;
    }
    __syncthreads();
  }
  __syncthreads();
  if (blk) {
    Vb[0][2 * bi][2 * bj] = vx0;
    Vb[0][2 * bi][2 * bj + 1] = vy0;
    Vb[0][2 * bi + 1][2 * bj] = vx1;
    Vb[0][2 * bi + 1][2 * bj + 1] = vy1;
  }
  __syncthreads();
  // positions holding one of the m indices, ranked by eigenvalue (ties by position)
  if (tid < mm) {
    const bool ok = pidx[ca][tid] < m;
    const double d = Ab[ca][tid][tid];
    int rk = 0;
    for (int j = 0; j < mm; j++) {
      if (pidx[ca][j] >= m) continue;
      const double e = Ab[ca][j][j];
      rk += (e < d || (e == d && j < tid)) ? 1 : 0;
    }
    if (ok) order[rk] = tid;
  }
  __syncthreads();
  for (int i = tid; i < m; i += kJacT) theta[i] = Ab[ca][order[i]][order[i]];
  for (int i = tid; i < m * m; i += kJacT) {
    const int r = i % m, c = i / m;
    const double v = Vb[0][r][order[c]];
    Y[r + c * 64] = v;
    Ynext[r + c * 64] = v;
  }
}

// Coefficients of a thick restart that keeps the previous iteration's Ritz vectors (GD+k):
// columns 0..keep-nz-1 of Q are the current Ritz vectors Y[:, c]; column keep-nz+j is the
// previous Ritz vector j, Z_j = [Yk[:, j] (mp rows, the earlier basis is a prefix of the
// current one); 0], orthogonalised twice against the columns before it and normalised; when
// less than 1e-6 of it is left (the root has stopped moving) the next current Ritz vector
// Y[:, keep-nz+j] takes its place.  One wave, lane r holds row r of every column (m <= 64).
__global__ __launch_bounds__(64) void k_gdk_basis(const double* __restrict__ Y, const double* __restrict__ Yk,
                                                  int m, int mp, int keep, int nz, double* __restrict__ Q) {
  __shared__ double q[64][65];
  const int r = threadIdx.x;
  const int nc = keep - nz;
  for (int c = 0; c < nc; c++) q[c][r] = r < m ? Y[r + c * 64] : 0.0;
  for (int j = 0; j < nz; j++) {
    const int c = nc + j;
    for (int attempt = 0; attempt < 2; attempt++) {
      double v = attempt == 0 ? (r < mp ? Yk[r + j * 64] : 0.0) : (r < m ? Y[r + c * 64] : 0.0);
      const double v0 = sqrt(wave_sum64(v * v));
      for (int pass = 0; pass < 2; pass++)
        for (int b = 0; b < c; b++) {
          const double d = wave_sum64(q[b][r] * v);
          v -= d * q[b][r];
        }
      const double nrm = sqrt(wave_sum64(v * v));
      if (attempt == 1 || nrm > 1e-6 * v0) {
        q[c][r] = nrm > 0.0 ? v / nrm : 0.0;
        break;
      }
    }
  }
  for (int c = 0; c < keep; c++) Q[r + c * 64] = r < m ? q[c][r] : 0.0;
}

// Cholesky factor of a small Gram matrix B (n x n, leading dimension 64) with column dropping,
// on the device: column j is kept when its residual norm after the kept columns before it
// exceeds thr_j = max(rel * sqrt(g0[j][j]), 1e-30) (thr_j = 0 without g0), and, when rg is given,
// when the residual norm sqrt(rg[j][j]) of the root it corrects is at least skip (converged
// roots are not expanded).  Writes M (n x nk,
// leading dimension 64, zero elsewhere) with Q = T M orthonormal, and nk.  n comes from n_dev
// when given.  One wave: lane 0 factors, lane b inverts column b of the triangular factor.
__global__ __launch_bounds__(64) void k_cholqr_small(const double* __restrict__ B, const int* __restrict__ n_dev,
                                                     int n_host, const double* __restrict__ g0, double rel,
                                                     double* __restrict__ M, int* __restrict__ nk_out,
                                                     const double* __restrict__ rg = nullptr, double skip = 0.0) {
  __shared__ double R[64][65];
  __shared__ double Ri[64][65];
  __shared__ int kept[64];
  __shared__ int nk_s;
  const int lane = threadIdx.x;
  const int n = n_dev ? *n_dev : n_host;
  for (int i = lane; i < n * n; i += 64) {
    R[i / n][i % n] = 0.0;
    Ri[i / n][i % n] = 0.0;
  }
  __syncthreads();
  if (lane == 0) {
    int nk = 0;
    for (int j = 0; j < n; j++) {
      double d = B[j + j * 64];
      for (int a = 0; a < nk; a++) {
        const int i = kept[a];
        double v = B[i + j * 64];
        for (int b = 0; b < a; b++) v -= R[kept[b]][i] * R[kept[b]][j];
        R[i][j] = v / R[i][i];
        d -= R[i][j] * R[i][j];
      }
      const double t = g0 ? fmax(rel * sqrt(fmax(g0[j + j * 64], 0.0)), 1e-30) : 0.0;
      const bool converged = rg && rg[j + j * 64] < skip * skip;
      if (d > t * t && d > 0.0 && !converged) {
        R[j][j] = sqrt(d);
        kept[nk++] = j;
      }
    }
    nk_s = nk;
  }
  __syncthreads();
  const int nk = nk_s;
  if (lane < nk) {  // column b of the inverse of the kept factor Rk[a][c] = R[kept[a]][kept[c]]
    const int b = lane;
    Ri[b][b] = 1.0 / R[kept[b]][kept[b]];
    for (int a = b - 1; a >= 0; a--) {
      double v = 0.0;
      for (int c = a + 1; c <= b; c++) v += R[kept[a]][kept[c]] * Ri[c][b];
      Ri[a][b] = -v / R[kept[a]][kept[a]];
    }
  }
  __shared__ int slot[64];  // position of row r among the kept columns, or -1
  if (lane < 64) slot[lane] = -1;
  __syncthreads();
  if (lane < nk) slot[kept[lane]] = lane;
  __syncthreads();
  for (int i = lane; i < 64 * 64; i += 64) {
    const int r = i % 64, c = i / 64;  // M[r + 64 c]
    const int a = slot[r];
    M[i] = (c < nk && a >= 0) ? Ri[a][c] : 0.0;
  }
  if (lane == 0) *nk_out = nk;
}

// status block read by the host once per iteration: residual norms, Ritz values, appended count
__global__ void k_pack_status(const double* __restrict__ rgram, const double* __restrict__ theta,
                              const int* __restrict__ nk, int nev, double* __restrict__ out) {
  const int i = threadIdx.x;
  if (i < 64) {
    out[i] = i < nev ? sqrt(fmax(rgram[i + i * 64], 0.0)) : 0.0;
    out[64 + i] = theta[i];
  }
  if (i == 0) {
    out[128] = nk[0];
    out[129] = nk[1];
  }
}

static int g1(int64_t n, int b) { return (int)((n + b - 1) / b); }

// G (64 x 64): columns [old, m) are new.  Fill their mirror rows; gt (newc x old, ld 64),
// when given, supplies the new-vs-old couplings.
// Per-iteration log on stderr when PQG_DAV_LOG is set: the fields of the reference's
// DavidsonDerivedLogger (src/io/davidson_logging.hpp:17-81) -- iteration, converged roots,
// subspace size, eigenvalue and residual of every root.
static bool dav_log_enabled() {
  static const bool on = std::getenv("PQG_DAV_LOG") != nullptr;
  return on;
}
static void dav_log(const char* tag, int iter, int m, double tol, const std::vector<double>& th,
                    const std::vector<double>& rn) {
  if (!dav_log_enabled()) return;
  int nconv = 0;
  for (double r : rn) nconv += r < tol;
  std::fprintf(stderr, "[davidson %s] iter %d subspace %d converged %d/%zu |", tag, iter, m, nconv, rn.size());
  for (size_t k = 0; k < rn.size() && k < th.size(); k++) std::fprintf(stderr, " E%zu=%.12f r=%.3e", k, th[k], rn[k]);
  std::fprintf(stderr, "\n");
}

Davidson::Davidson(pqg_ci* ci, const pqg_dav_params& p) : ci_(ci), p_(p) {
  try {
    setup();
  } catch (...) {
    // members built so far are released by the unwinding: let queued work finish first
    (void)hipStreamSynchronize(ci_->stream);
    throw;
  }
}

void Davidson::setup() {
  pqg_ci* ci = ci_;
  const pqg_dav_params& p = p_;
  N_ = ci->n_int;
  ld_ = (N_ + 31) / 32 * 32;  // every column 256-byte aligned
  nev_ = std::max(1, p.nev);
  PQG_REQUIRE(nev_ <= 64 && nev_ <= ci->n_ref, PQG_ERR_ARG, "nev out of range");
  ninit_ = std::max(p.nvec_init, nev_);
  nmax_ = p.nvec_max > 0 ? p.nvec_max : std::max(2 * p.nvec_init, 10 * nev_);
  ninit_ = (int)std::min<int64_t>(ninit_, ci->n_ref);
  nmax_ = (int)std::min<int64_t>(std::max(nmax_, ninit_ + nev_), std::min<int64_t>(ci->n_ref, 64));
  PQG_REQUIRE(nmax_ >= ninit_, PQG_ERR_ARG, "subspace too small");
  mixed_ = p.precision == PQG_MIXED;
  // thick restarts keep the previous iteration's Ritz vectors unless PQG_DAV_GDK=0
  gdk_ = true;
  if (const char* e = std::getenv("PQG_DAV_GDK")) gdk_ = std::atoi(e) != 0;
  pqg_ctx* ctx = ci->ctx;
  V_.reset(ctx, (size_t)ld_ * nmax_);
  W_.reset(ctx, (size_t)ld_ * nmax_);
  G_.reset(ctx, 64 * 64);
  Y_.reset(ctx, 64 * 64);
  Yprev_.reset(ctx, 64 * 64);
  Ynext_.reset(ctx, 64 * 64);
  theta_.reset(ctx, 64);
  X_.reset(ctx, (size_t)ld_ * nev_);
  AX_.reset(ctx, (size_t)ld_ * nev_);
  RT_.reset(ctx, (size_t)ld_ * nev_ * 2);  // residuals, then corrections (one Gram covers both)
  tmp_.reset(ctx, (size_t)ld_ * ninit_);
  small_.reset(ctx, 64 * 64);
  Yk_.reset(ctx, 64 * 64);
  Qr_.reset(ctx, 64 * 64);
  gram0_.reset(ctx, 64 * 64);
  rgram_.reset(ctx, 64 * 64);
  mbuf_.reset(ctx, 64 * 64);
  nk_dev_.reset(ctx, 2);
  reorth_dev_.reset(ctx, 1);
  small2_.reset(ctx, 64 * 64);
  status_dev_.reset(ctx, 130);
  {
    std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
    PQG_HIP(hipHostMalloc((void**)&status_pin_.p, 130 * 8, hipHostMallocDefault));
  }
  status_h_ = status_pin_.p;
  // partials of the tall-skinny products: at most ceil(N / 512) row chunks
  nchunk_ = (int)((N_ + 511) / 512);
  part_.reset(ctx, (size_t)std::min<int64_t>(nchunk_, 1024) * 64 * 64);
  if (mixed_) {
    W64_.reset(ctx, (size_t)ld_ * nev_);
    c32_.reset(ctx, (size_t)N_ * nev_ > (size_t)N_ * ninit_ ? (size_t)N_ * nev_ : (size_t)N_ * ninit_);
    s32_.reset(ctx, c32_.n);
    ensure_mixed(ci);
  }
  t0_ = std::chrono::steady_clock::now();
  initial_guess(ci->stream);
}

uint64_t Davidson::device_bytes() const {
  return V_.bytes() + W_.bytes() + X_.bytes() + AX_.bytes() + RT_.bytes() + tmp_.bytes() + W64_.bytes() +
         c32_.bytes() + s32_.bytes() + part_.bytes();
}

// the tall-skinny kernels read two rows per 16-byte access: every column 16-byte aligned
static void ts_check(const double* p, int64_t ld) {
  PQG_REQUIRE(((uintptr_t)p & 15) == 0 && (ld & 1) == 0, PQG_ERR_ARG, "tall-skinny operand not 16-byte aligned");
}

void Davidson::ts_gemm_tn(hipStream_t st, const double* A, int64_t lda, int na, const double* B, int64_t ldb, int nb,
                          double* out, int ldo, const int* gate) {
  ts_check(A, lda);
  ts_check(B, ldb);
  PQG_REQUIRE(na >= 1 && nb >= 1 && na <= 64 && nb <= 64, PQG_ERR_ARG, "tall-skinny product shape");
  // rows per block: about 8192 blocks, at least 256 rows, and no more blocks than the partial
  // buffer holds; a multiple of the 128 rows a wave takes per step
  const int64_t G = (int64_t)na * nb;
  int64_t rpb = std::max<int64_t>(256, (N_ + 8191) / 8192);
  const int64_t maxblk = (int64_t)part_.n / G;
  rpb = std::max<int64_t>(rpb, (N_ + maxblk - 1) / maxblk);
  rpb = (rpb + 127) / 128 * 128;
  const int64_t nblk = (N_ + rpb - 1) / rpb;
  PQG_REQUIRE((size_t)(nblk * G) <= part_.n, PQG_ERR_ARG, "tall-skinny partial buffer too small");
  const int tasks = ((na + kTnQA - 1) / kTnQA) * ((nb + 3) / 4);
  const int per = std::min(tasks, 4);
  dim3 grid((unsigned)nblk, (unsigned)((tasks + per - 1) / per));
  k_tn_part<<<grid, 64 * per, 0, st>>>(A, lda, na, B, ldb, nb, N_, rpb, part_.get(), gate);
  PQG_HIP(hipGetLastError());
  k_tn_reduce<<<(unsigned)G, 256, 0, st>>>(part_.get(), (int)nblk, na, nb, out, ldo, gate);
  PQG_HIP(hipGetLastError());
}

static void nn_launch(dim3 g, hipStream_t st, const double* A0, const double* A1, int64_t lda, int ka,
                      const double* Bs, int ldb, int nb, double* out0, double* out1, int64_t ldo, int64_t nrows,
                      double alpha, double beta, const int* gate = nullptr) {
  if (nb <= 4)
    k_nn<4><<<g, kTB, 0, st>>>(A0, A1, lda, ka, Bs, ldb, nb, out0, out1, ldo, nrows, alpha, beta, gate);
  else if (nb <= 8)
    k_nn<8><<<g, kTB, 0, st>>>(A0, A1, lda, ka, Bs, ldb, nb, out0, out1, ldo, nrows, alpha, beta, gate);
  else
    k_nn<16><<<g, kTB, 0, st>>>(A0, A1, lda, ka, Bs, ldb, nb, out0, out1, ldo, nrows, alpha, beta, gate);
  PQG_HIP(hipGetLastError());
}

void Davidson::ts_gemm_nn(hipStream_t st, const double* A, int64_t lda, int ka, const double* Bs, int ldb, int nb,
                          double* out, int64_t ldo, double alpha, double beta, const int* gate) {
  ts_check(A, lda);
  ts_check(out, ldo);
  PQG_REQUIRE(ka >= 1 && ka <= 64 && nb >= 1, PQG_ERR_ARG, "tall-skinny product shape");
  PQG_REQUIRE(out != A || nb <= 16, PQG_ERR_ARG, "in-place product wider than one column chunk");
  nn_launch(dim3(g1((N_ + 1) / 2, kTB)), st, A, A, lda, ka, Bs, ldb, nb, out, out, ldo, N_, alpha, beta, gate);
}

void Davidson::ts_gemm_nn2(hipStream_t st, const double* A0, const double* A1, int64_t lda, int ka, const double* Bs,
                           int ldb, int nb, double* out0, double* out1, int64_t ldo) {
  ts_check(A0, lda);
  ts_check(A1, lda);
  ts_check(out0, ldo);
  ts_check(out1, ldo);
  PQG_REQUIRE(ka >= 1 && ka <= 64 && nb >= 1, PQG_ERR_ARG, "tall-skinny product shape");
  nn_launch(dim3(g1((N_ + 1) / 2, kTB), 2), st, A0, A1, lda, ka, Bs, ldb, nb, out0, out1, ldo, N_, 1.0, 0.0);
}

void Davidson::initial_guess(hipStream_t st) {
  // unit vectors at the ninit smallest diagonal entries (ties: lower index first)
  std::vector<double> hd(N_);
  PQG_HIP(hipMemcpyAsync(hd.data(), operator_diag(ci_), N_ * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> idx(N_);
  std::iota(idx.begin(), idx.end(), 0);
  std::partial_sort(idx.begin(), idx.begin() + ninit_, idx.end(),
                    [&](int64_t a, int64_t b) { return hd[a] < hd[b] || (hd[a] == hd[b] && a < b); });
  DevBuf<int64_t> pos(ci_->ctx, ninit_);
  PQG_HIP(hipMemcpyAsync(pos.get(), idx.data(), ninit_ * 8, hipMemcpyHostToDevice, st));
  if (p_.guess_perturb > 0.0) {
    // per-entry size guess_perturb * sqrt(3 / N): the admixture has norm ~guess_perturb
    k_unit_perturbed<<<g1(N_, 256), 256, 0, st>>>(V_.get(), N_, ld_, pos.get(), ninit_,
                                                  p_.guess_perturb * std::sqrt(3.0 / (double)N_));
    PQG_HIP(hipGetLastError());
    mask_invalid<double>(ci_, V_.get(), ld_, ninit_, st);
    for (int pass = 0; pass < 2; pass++) {  // CholQR2
      ts_gemm_tn(st, V_.get(), ld_, ninit_, V_.get(), ld_, ninit_, small_.get(), 64);
      k_cholqr_small<<<1, 64, 0, st>>>(small_.get(), nullptr, ninit_, nullptr, 0.0, mbuf_.get(), nk_dev_.get());
      PQG_HIP(hipGetLastError());
      ts_gemm_nn(st, V_.get(), ld_, ninit_, mbuf_.get(), 64, ninit_, tmp_.get(), ld_, 1.0, 0.0);
      PQG_HIP(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)ld_ * ninit_ * 8, hipMemcpyDeviceToDevice, st));
    }
  } else {
    k_unit<<<g1(N_, 256), 256, 0, st>>>(V_.get(), N_, ld_, pos.get(), ninit_);
    PQG_HIP(hipGetLastError());
  }
  PQG_HIP(hipStreamSynchronize(st));
  m_ = ninit_;
  nW_ = 0;
  work_ = SigmaWork{m_, mixed_ ? PQG_MIXED : PQG_FP64, V_.get(), W_.get(), ld_};
}

void Davidson::sigma_body(hipStream_t st) {
  const int k = work_.k;
  if (work_.precision == PQG_FP64) {
    sigma_internal(ci_, work_.C, work_.S, k, work_.ld, st);
  } else {
    const int64_t n = (int64_t)N_ * k;
    grow(c32_, ci_->ctx, n, st);
    grow(s32_, ci_->ctx, n, st);
    for (int j = 0; j < k; j++)
      k_d2f_blk<<<g1(N_, 256), 256, 0, st>>>(work_.C + j * work_.ld, c32_.get() + (int64_t)j * N_, N_);
    sigma_internal(ci_, c32_.get(), s32_.get(), k, N_, st);
    for (int j = 0; j < k; j++)
      k_f2d_blk<<<g1(N_, 256), 256, 0, st>>>(s32_.get() + (int64_t)j * N_, work_.S + j * work_.ld, N_);
    PQG_HIP(hipGetLastError());
  }
}

// Device buffers a sigma may grow (and so move): a captured sigma is valid while they stay put.
uint64_t Davidson::sigma_buffers() const {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) { h = (h ^ v) * 1099511628211ull; };
  mix((uintptr_t)ci_->scratch.get());
  mix(ci_->scratch.n);
  mix((uintptr_t)ci_->lg.U.get());
  mix(ci_->lg.U.n);
  mix((uintptr_t)ci_->lg.P.get());
  mix(ci_->lg.P.n);
  mix((uintptr_t)c32_.get());
  mix(c32_.n);
  mix((uintptr_t)s32_.get());
  mix(s32_.n);
  return h;
}

// Captures sigma_body into a graph (nothing is executed); on any failure the capture is ended,
// errors are cleared, graphs are switched off for the solve and nullptr comes back.  The capture
// holds device_op_mutex, so no allocation or free of this library runs meanwhile.
hipGraphExec_t Davidson::capture_sigma(hipStream_t st) {
  std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  bool ok = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) == hipSuccess;
  if (ok) {
    try {
      sigma_body(st);
    } catch (...) {
      ok = false;
    }
    ok = (hipStreamEndCapture(st, &g) == hipSuccess) && ok && g;
    if (ok) ok = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess;
    if (g) (void)hipGraphDestroy(g);
  }
  if (!ok) {
    (void)hipGetLastError();
    if (ex) (void)hipGraphExecDestroy(ex);
    ex = nullptr;
    graphs_.on = false;
  }
  return ex;
}

SigmaGraphs::~SigmaGraphs() {
  for (auto& e : entries)
    if (e.exec) (void)hipGraphExecDestroy(e.exec);
}

// With graphs on, a sigma block seen before (same vectors, count, precision and buffers) is
// captured into a hipGraph on its second run and replayed from then on: one launch instead of
// the few tens of kernels of a small space's sigma.  The first run of a block is a plain one,
// so every buffer the sigma grows is in place before anything is captured.
void Davidson::run_sigma(hipStream_t st) {
  if (work_.k == 0) return;
  const int k = work_.k;
  if (graphs_.on && !ci_->ctx->profiling.load()) {
    SigmaGraphs::Entry* e = nullptr;
    for (auto& x : graphs_.entries)
      if (x.C == work_.C && x.S == work_.S && x.k == k && x.precision == work_.precision && x.ld == work_.ld) e = &x;
    const uint64_t bufs = sigma_buffers();
    if (e && e->bufs == bufs && !e->exec) e->exec = capture_sigma(st);
    if (e && e->bufs == bufs && e->exec) {
      PQG_HIP(hipGraphLaunch(e->exec, st));
    } else if (e && e->bufs == bufs) {
      sigma_body(st);  // the capture failed: graphs are off for this solve from now on
    } else {
      sigma_body(st);
      // the block as run (after any growth) becomes capturable on its next run
      if (!e) {
        graphs_.entries.push_back(SigmaGraphs::Entry{work_.C, work_.S, k, work_.precision, work_.ld, 0, nullptr});
        e = &graphs_.entries.back();
      } else if (e->exec) {
        (void)hipGraphExecDestroy(e->exec);
        e->exec = nullptr;
      }
      e->bufs = sigma_buffers();
    }
  } else {
    sigma_body(st);
  }
  if (work_.precision == PQG_FP64) sigma_calls64_ += k;
  sigma_calls_ += k;
}

// PQG_DAV_REORTH=1: the second Gram-Schmidt sweep always runs (the gate below is then always open)
static int reorth_always() {
  static const int on = [] {
    const char* e = std::getenv("PQG_DAV_REORTH");
    return e && std::atoi(e) != 0 ? 1 : 0;
  }();
  return on;
}

// Second Gram-Schmidt sweep needed?  DGKS criterion: yes when a column kept less than 1/sqrt(2) of
// its norm in the first sweep (g0: Gram before, g1: after; columns of converged roots, which the
// Cholesky-QR drops anyway, do not count).  One thread; the flag gates the second sweep.
__global__ void k_reorth_gate(const double* __restrict__ g0, const double* __restrict__ g1, int n,
                              const double* __restrict__ rg, double skip, int always, int* __restrict__ gate) {
  if (threadIdx.x != 0) return;
  int need = always;
  for (int j = 0; j < n; j++) {
    if (rg && rg[j + j * 64] < skip * skip) continue;
    const double before = g0[j + j * 64], after = g1[j + j * 64];
    if (before > 0.0 && after < 0.5 * before) need = 1;
  }
  *gate = need;
}

// Orthogonalise T (nt columns) against V (m columns) by block Gram-Schmidt, a second sweep only
// where the first lost most of a column (k_reorth_gate), then orthonormalise the block by two
// Cholesky-QR passes (columns that vanish are dropped) into V[:, m:m+nt], all on the stream; the
// kept count (and the second pass's count, which must agree) lands in nk_dev_[0..1].  Nothing is
// read back here; finish_append() commits the columns.
void Davidson::append_async(hipStream_t st, double* T, int nt, const double* g0, const double* rg) {
  nt = std::min(nt, nmax_ - m_);
  pending_nt_ = nt;
  if (nt <= 0) return;
  if (!g0) {
    ts_gemm_tn(st, T, ld_, nt, T, ld_, nt, gram0_.get(), 64);
    g0 = gram0_.get();
  }
  ts_gemm_tn(st, V_.get(), ld_, m_, T, ld_, nt, small_.get(), 64);
  ts_gemm_nn(st, V_.get(), ld_, m_, small_.get(), 64, nt, T, ld_, -1.0, 1.0);
  ts_gemm_tn(st, T, ld_, nt, T, ld_, nt, small_.get(), 64);
  k_reorth_gate<<<1, 64, 0, st>>>(g0, small_.get(), nt, rg, p_.tol, reorth_always(), reorth_dev_.get());
  PQG_HIP(hipGetLastError());
  const int* gate = reorth_dev_.get();
  ts_gemm_tn(st, V_.get(), ld_, m_, T, ld_, nt, small2_.get(), 64, gate);
  ts_gemm_nn(st, V_.get(), ld_, m_, small2_.get(), 64, nt, T, ld_, -1.0, 1.0, gate);
  ts_gemm_tn(st, T, ld_, nt, T, ld_, nt, small_.get(), 64, gate);
  // Cholesky-QR pass 1: drop vanishing columns, orthonormalise the rest straight into V[:, m:]
  k_cholqr_small<<<1, 64, 0, st>>>(small_.get(), nullptr, nt, g0, 1e-10, mbuf_.get(), nk_dev_.get(), rg, p_.tol);
  PQG_HIP(hipGetLastError());
  double* dst = V_.get() + (int64_t)m_ * ld_;
  ts_gemm_nn(st, T, ld_, nt, mbuf_.get(), 64, nt, dst, ld_, 1.0, 0.0);
  // pass 2 (CholQR2): re-orthonormalise the new block
  ts_gemm_tn(st, dst, ld_, nt, dst, ld_, nt, small_.get(), 64);
  k_cholqr_small<<<1, 64, 0, st>>>(small_.get(), nk_dev_.get(), 0, nullptr, 0.0, mbuf_.get(), nk_dev_.get() + 1);
  PQG_HIP(hipGetLastError());
  if (nt <= 8) {  // in place: each row is read before it is written
    ts_gemm_nn(st, dst, ld_, nt, mbuf_.get(), 64, nt, dst, ld_, 1.0, 0.0);
  } else {
    ts_gemm_nn(st, dst, ld_, nt, mbuf_.get(), 64, nt, tmp_.get(), ld_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(dst, tmp_.get(), (size_t)ld_ * nt * 8, hipMemcpyDeviceToDevice, st));
  }
  mask_invalid<double>(ci_, dst, ld_, nt, st);
}

// Commits the block appended by append_async (status_h_ must hold the counts): returns how many
// columns were added.
int Davidson::finish_append() {
  if (pending_nt_ <= 0) return 0;
  pending_nt_ = 0;
  const int nk = (int)status_h_[128], nk2 = (int)status_h_[129];
  if (nk <= 0 || nk2 != nk) return 0;  // numerically rank deficient: do not extend
  m_ += nk;
  return nk;
}

// Reads residual norms, Ritz values and the appended count in one transfer (one sync).
void Davidson::read_status_async(hipStream_t st, const double* rgram) {
  k_pack_status<<<1, 64, 0, st>>>(rgram, theta_.get(), nk_dev_.get(), nev_, status_dev_.get());
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipMemcpyAsync(status_h_, status_dev_.get(), 130 * 8, hipMemcpyDeviceToHost, st));
}

int Davidson::orthonormalize_append(hipStream_t st, double* T, int nt) {
  append_async(st, T, nt, nullptr);
  k_pack_status<<<1, 64, 0, st>>>(small_.get(), theta_.get(), nk_dev_.get(), 0, status_dev_.get());
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipMemcpyAsync(status_h_, status_dev_.get(), 130 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  return finish_append();
}

// Thick restart on the first `keep` Ritz vectors of the current subspace (Y from the last
// Rayleigh-Ritz); the projected matrix is recomputed exactly.
// Thick restart to `keep` vectors: the lowest Ritz vectors, or (GD+k) the lowest keep - nev
// of them and the previous iteration's nev Ritz vectors
void Davidson::restart_to_ritz(hipStream_t st, int keep) {
  const double* Yr = Y_.get();
  if (gdk_ && yk_rows_ > 0 && keep > nev_ && keep <= m_) {
    k_gdk_basis<<<1, 64, 0, st>>>(Y_.get(), Yk_.get(), m_, yk_rows_, keep, nev_, Qr_.get());
    PQG_HIP(hipGetLastError());
    Yr = Qr_.get();
  }
  ts_gemm_nn(st, V_.get(), ld_, m_, Yr, 64, keep, tmp_.get(), ld_, 1.0, 0.0);
  PQG_HIP(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)ld_ * keep * 8, hipMemcpyDeviceToDevice, st));
  ts_gemm_nn(st, W_.get(), ld_, m_, Yr, 64, keep, tmp_.get(), ld_, 1.0, 0.0);
  PQG_HIP(hipMemcpyAsync(W_.get(), tmp_.get(), (size_t)ld_ * keep * 8, hipMemcpyDeviceToDevice, st));
  m_ = nW_ = keep;
  warm_ = 0;
  ts_gemm_tn(st, V_.get(), ld_, keep, W_.get(), ld_, keep, G_.get(), 64);
  restarts_++;
}

// X = V Y, T = (W Y - X theta) / (theta - diag) and the diagonal of the [R T] Gram in one pass
// (nev <= 8)
void Davidson::ritz_residual(hipStream_t st) {
  ts_check(V_.get(), ld_);
  ts_check(W_.get(), ld_);
  const int64_t pairs = (N_ + 1) / 2;
  const int nevt = nev_ <= 4 ? 4 : 8;
  const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(g1(pairs, kTB), (int64_t)part_.n / (2 * nevt)));
  const int nblk2 = std::min(nblk, 8192);
  if (nevt == 4)
    k_ritz_res<4><<<nblk2, kTB, 0, st>>>(V_.get(), W_.get(), ld_, m_, Y_.get(), theta_.get(), operator_diag(ci_),
                                         nev_, X_.get(), Tp(), N_, part_.get());
  else
    k_ritz_res<8><<<nblk2, kTB, 0, st>>>(V_.get(), W_.get(), ld_, m_, Y_.get(), theta_.get(), operator_diag(ci_),
                                         nev_, X_.get(), Tp(), N_, part_.get());
  PQG_HIP(hipGetLastError());
  k_ritz_res_reduce<<<2 * nevt, 256, 0, st>>>(part_.get(), nblk2, nev_, nevt, rgram_.get());
  PQG_HIP(hipGetLastError());
}

void Davidson::rayleigh_ritz(hipStream_t st) {
  // G[:, old:m] = V^T W[:, old:m]; the rows of the new columns against older columns follow by
  // symmetry.  Mixed precision: those couplings are taken as v_new^T (W v_old) instead, so that
  // basis vectors whose sigma is FP64 (the Ritz vectors after a defect-correction restart)
  // couple with FP64 accuracy.
  const int old = nW_;
  const int newc = m_ - nW_;
  const double* gt = nullptr;
  if (newc > 0) {
    ts_gemm_tn(st, V_.get(), ld_, m_, W_.get() + (int64_t)nW_ * ld_, ld_, newc, G_.get() + (int64_t)nW_ * 64, 64);
    if (mixed_ && old > 0) {
      ts_gemm_tn(st, V_.get() + (int64_t)old * ld_, ld_, newc, W_.get(), ld_, old, small_.get(), 64);
      gt = small_.get();
    }
    nW_ = m_;
  }
  // Yprev_ (warm start) and Y_ are distinct buffers: the kernel reads one and writes both
  // Off-diagonal tolerance of the subspace eigensolve: the eigenvectors only need to be accurate
  // well below the current residual level (1e-3 of the worst residual relative to the largest
  // Ritz value), and to full precision once the residuals approach the convergence threshold.
  // the previous iteration's Ritz vectors in this basis (its first warm_ columns are the
  // previous basis), for a restart that keeps them
  if (gdk_ && warm_ > 0) {
    PQG_HIP(hipMemcpyAsync(Yk_.get(), Y_.get(), (size_t)64 * nev_ * 8, hipMemcpyDeviceToDevice, st));
    yk_rows_ = warm_;
  } else {
    yk_rows_ = 0;
  }
  double rel = 1e-15;
  if (have_worst_ && theta_abs_max_ > 0.0)
    rel = std::max(1e-15, std::min(1e-6, 1e-3 * std::max(last_worst_, p_.tol) / theta_abs_max_));
  k_jacobi<<<1, kJacT, 0, st>>>(G_.get(), m_, old, gt, Yprev_.get(), std::min(warm_, m_), theta_.get(), Y_.get(),
                               Ynext_.get(), rel * rel);
  PQG_HIP(hipGetLastError());
  std::swap(Yprev_, Ynext_);
  warm_ = m_;
}

void Davidson::post_sigma(hipStream_t st) {
  post_sigma_begin(st);
  post_sigma_end(st);
}

// The device half of an iteration, up to the asynchronous copy of the status block: nothing here
// waits for the device, so many solves can queue their iterations back to back before one
// synchronisation (post_sigma_end completes each of them).
void Davidson::post_sigma_begin(hipStream_t st) {
  began_ = false;
  if (done_ || checking_) return;
  iter_++;
  rayleigh_ritz(st);
  // Ritz vectors and residuals; what the next steps read of the Gram of [R T] is its diagonal:
  // |r_k|^2 (residual norms) and |t_k|^2 (drop thresholds of the corrections)
  const double* g0 = nullptr;
  if (nev_ <= 8) {
    ritz_residual(st);
    g0 = rgram_.get() + nev_ + nev_ * 64;
  } else if (2 * nev_ <= 64) {
    ts_gemm_nn2(st, V_.get(), W_.get(), ld_, m_, Y_.get(), 64, nev_, X_.get(), AX_.get(), ld_);
    k_residual<<<g1(N_, 256), 256, 0, st>>>(X_.get(), AX_.get(), theta_.get(), operator_diag(ci_), Rp(), Tp(), N_,
                                            ld_, nev_);
    PQG_HIP(hipGetLastError());
    ts_gemm_tn(st, Rp(), ld_, 2 * nev_, Rp(), ld_, 2 * nev_, rgram_.get(), 64);
    g0 = rgram_.get() + nev_ + nev_ * 64;
  } else {
    ts_gemm_nn2(st, V_.get(), W_.get(), ld_, m_, Y_.get(), 64, nev_, X_.get(), AX_.get(), ld_);
    k_residual<<<g1(N_, 256), 256, 0, st>>>(X_.get(), AX_.get(), theta_.get(), operator_diag(ci_), Rp(), Tp(), N_,
                                            ld_, nev_);
    PQG_HIP(hipGetLastError());
    ts_gemm_tn(st, Rp(), ld_, nev_, Rp(), ld_, nev_, rgram_.get(), 64);
  }
  // The next block is prepared before the convergence test is read back, so an iteration costs
  // one host synchronisation.  A mixed-precision solve may still need the current subspace for
  // its FP64 check, so it does not restart ahead of the test.
  need_restart_ = m_ + nev_ > nmax_;
  speculate_ = !(mixed_ && need_restart_);
  if (speculate_) {
    if (need_restart_) restart_to_ritz(st, ninit_);
    // with the perturbed start every eigenvector has weight in the basis, so converged roots
    // need not be expanded; with unit vectors their corrections can still bring in missing roots
    append_async(st, Tp(), nev_, g0, (p_.guess_perturb > 0.0 && 2 * nev_ <= 64) ? rgram_.get() : nullptr);
  }
  read_status_async(st, rgram_.get());
  began_ = true;
}

// The host half: waits for the status block of post_sigma_begin (or runs an FP64 check of the
// mixed path), decides convergence and queues the next sigma block.
void Davidson::post_sigma_end(hipStream_t st) {
  if (done_) return;
  if (checking_) {
    // FP64 sigma of the Ritz vectors is in W64_: true residuals
    checking_ = false;
    since_check_ = 0;
    {  // final energies: Rayleigh quotients x^T H x of the (orthonormal) Ritz vectors in FP64
      ts_gemm_tn(st, X_.get(), ld_, nev_, W64_.get(), ld_, nev_, small_.get(), 64);
      std::vector<double> q(64 * 64);
      PQG_HIP(hipMemcpyAsync(q.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
      for (int k = 0; k < nev_; k++) theta_h_[k] = q[k + k * 64];
      PQG_HIP(hipMemcpyAsync(theta_.get(), theta_h_.data(), nev_ * 8, hipMemcpyHostToDevice, st));
    }
    k_residual<<<g1(N_, 256), 256, 0, st>>>(X_.get(), W64_.get(), theta_.get(), operator_diag(ci_), Rp(), Tp(),
                                            N_, ld_, nev_);
    PQG_HIP(hipGetLastError());
    ts_gemm_tn(st, Rp(), ld_, nev_, Rp(), ld_, nev_, small_.get(), 64);
    std::vector<double> g(64 * 64);
    PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    double worst = 0.0;
    for (int k = 0; k < nev_; k++) {
      rn_h_[k] = std::sqrt(std::max(g[k + k * 64], 0.0));
      worst = std::max(worst, rn_h_[k]);
    }
    last_true_ = worst;
    dav_log("fp64-check", iter_, m_, p_.tol, theta_h_, rn_h_);
    if (worst < p_.tol) {
      done_ = true;
      work_ = SigmaWork{};
      seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
      return;
    }
    if (iter_ >= p_.maxit) {
      done_ = true;
      status_ = PQG_ERR_NOCONV;
      work_ = SigmaWork{};
      seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
      return;
    }
    // defect correction: restart on the Ritz vectors, the first nev with their exact FP64
    // sigma, the next ones (up to nvec_init) with the sigma of the current basis
    const int keep = std::max(nev_, std::min(ninit_, m_));
    ts_gemm_nn(st, V_.get(), ld_, m_, Y_.get(), 64, keep, tmp_.get(), ld_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)ld_ * keep * 8, hipMemcpyDeviceToDevice, st));
    ts_gemm_nn(st, W_.get(), ld_, m_, Y_.get(), 64, keep, tmp_.get(), ld_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(W_.get(), tmp_.get(), (size_t)ld_ * keep * 8, hipMemcpyDeviceToDevice, st));
    PQG_HIP(hipMemcpyAsync(W_.get(), W64_.get(), (size_t)ld_ * nev_ * 8, hipMemcpyDeviceToDevice, st));
    m_ = nW_ = keep;
    warm_ = 0;
    // projected matrix: couplings to the exact columns are taken as v^T (H x) (FP64)
    ts_gemm_tn(st, V_.get(), ld_, keep, W_.get(), ld_, keep, small_.get(), 64);
    {
      std::vector<double> q(64 * 64), gk(64 * 64, 0.0);
      PQG_HIP(hipMemcpyAsync(q.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
      for (int a = 0; a < keep; a++)
        for (int b = 0; b < keep; b++) {
          double v;
          if (a < nev_ && b >= nev_)
            v = q[b + a * 64];  // v_b^T W64_a
          else if (b < nev_ && a >= nev_)
            v = q[a + b * 64];
          else
            v = 0.5 * (q[a + b * 64] + q[b + a * 64]);
          gk[a + b * 64] = v;
        }
      PQG_HIP(hipMemcpyAsync(G_.get(), gk.data(), 64 * 64 * 8, hipMemcpyHostToDevice, st));
    }
    restarts_++;
    const int before = m_;
    orthonormalize_append(st, Tp(), nev_);
    work_ = SigmaWork{m_ - before, PQG_MIXED, V_.get() + (int64_t)before * ld_, W_.get() + (int64_t)before * ld_, ld_};
    if (work_.k == 0) {  // nothing new to add: accept
      done_ = true;
      seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }
    return;
  }
  if (!began_) return;
  began_ = false;
  PQG_HIP(hipStreamSynchronize(st));
  const bool need_restart = need_restart_, speculate = speculate_;
  theta_h_.assign(status_h_ + 64, status_h_ + 64 + nev_);
  rn_h_.assign(status_h_, status_h_ + nev_);
  double worst = 0.0;
  for (int k = 0; k < nev_; k++) worst = std::max(worst, rn_h_[k]);
  last_worst_ = worst;
  have_worst_ = true;
  theta_abs_max_ = 0.0;
  for (int k = 0; k < nev_; k++) theta_abs_max_ = std::max(theta_abs_max_, std::fabs(theta_h_[k]));
  dav_log(mixed_ ? "fp32" : "fp64", iter_, m_, p_.tol, theta_h_, rn_h_);
  since_check_++;
  // Mixed precision: the FP32-based residual carries the error of the FP32 sigma of the
  // directions added since the last FP64 check, a few 1e-6 of |theta| at most, and that error
  // shrinks with their weight in the Ritz vectors.  So an FP64 check is spent when the residual
  // has dropped 20x below the last true one (first at 1e-3), or when it no longer improves
  // (10% over 6 iterations) below 1e-4 |theta| -- the level where the FP32 error can stall it --
  // and at the latest every 50 iterations.
  if (mixed_) {
    worst_hist_[since_check_ % 8] = worst;
    const double thr = std::max(p_.tol, std::min(1e-3, 0.05 * last_true_));
    const bool stalled = since_check_ > 6 && worst < 1e-4 * theta_abs_max_ &&
                         worst > 0.9 * worst_hist_[(since_check_ - 6) % 8];
    mixed_check_ = worst < thr || stalled || since_check_ >= 50;
  }
  const bool approx_conv = mixed_ ? mixed_check_ : (worst < p_.tol);
  if (approx_conv && mixed_) {
    // FP64 check of the Ritz vectors (the speculative block is dropped: m_ is unchanged)
    pending_nt_ = 0;
    checking_ = true;
    work_ = SigmaWork{nev_, PQG_FP64, X_.get(), W64_.get(), ld_};
    return;
  }
  if (approx_conv || iter_ >= p_.maxit) {
    done_ = true;
    status_ = approx_conv ? PQG_OK : PQG_ERR_NOCONV;
    work_ = SigmaWork{};
    seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    return;
  }
  int added;
  const int before = need_restart ? ninit_ : m_;
  if (speculate) {
    added = finish_append();
  } else {
    restart_to_ritz(st, ninit_);
    added = orthonormalize_append(st, Tp(), nev_);
  }
  if (added == 0) {  // subspace exhausted
    done_ = true;
    status_ = worst < p_.tol ? PQG_OK : PQG_ERR_NOCONV;
    work_ = SigmaWork{};
    seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    return;
  }
  work_ = SigmaWork{added, mixed_ ? PQG_MIXED : PQG_FP64, V_.get() + (int64_t)before * ld_,
                    W_.get() + (int64_t)before * ld_, ld_};
}

pqg_dav_report Davidson::report() const {
  pqg_dav_report r{};
  r.iterations = iter_;
  r.converged = (done_ && status_ == PQG_OK) ? 1 : 0;
  r.sigma_calls = sigma_calls_;
  r.sigma_calls_fp64 = sigma_calls64_;
  r.restarts = restarts_;
  for (int k = 0; k < nev_ && k < 64 && k < (int)rn_h_.size(); k++) r.residual[k] = rn_h_[k];
  r.seconds = seconds_;
  return r;
}

}  // namespace pqg

using namespace pqg;

struct pqg_dav {
  pqg_ci* ci;
  pqg::Davidson dav;
  pqg_dav(pqg_ci* c, const pqg_dav_params& p) : ci(c), dav(c, p) {}
};

namespace {

void dav_check_params(const pqg_ci* ci, const pqg_dav_params* params) {
  PQG_REQUIRE(ci && params, PQG_ERR_ARG, "NULL argument");
  PQG_REQUIRE(params->precision == PQG_FP64 || params->precision == PQG_MIXED, PQG_ERR_ARG, "unknown precision");
  PQG_REQUIRE(params->tol > 0 && params->maxit > 0, PQG_ERR_ARG, "tol and maxit must be positive");
}

// energies, Ritz vectors (reference order, host) and report of a finished solve; raises the
// reference's "did not converge" error after filling them
void dav_finish(pqg_ci* ci, Davidson& dav, double* E, double* X, int64_t ldx, pqg_dav_report* report) {
  hipStream_t st = ci->stream;
  PQG_HIP(hipStreamSynchronize(st));
  const auto& ev = dav.energies();
  if (E)
    for (int k = 0; k < dav.nev(); k++) E[k] = ev[k];
  if (X) {
    PQG_REQUIRE(ldx >= ci->n_ref, PQG_ERR_ARG, "ldx smaller than N");
    DevBuf<double> xr(ci->ctx, (size_t)ci->n_ref * dav.nev());
    to_reference(ci, dav.ritz_vectors(), dav.ld(), xr.get(), ci->n_ref, dav.nev(), st);
    PQG_HIP(hipMemcpy2DAsync(X, ldx * 8, xr.get(), ci->n_ref * 8, ci->n_ref * 8, dav.nev(), hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
  }
  if (report) *report = dav.report();
  if (dav.status() != PQG_OK) throw Error{dav.status(), "CI Calculation did not converge!"};
}

}  // namespace

extern "C" pqg_status pqg_ci_davidson(pqg_ci* ci, const pqg_dav_params* params, double* E, double* X, int64_t ldx,
                                      pqg_dav_report* report) {
  return guard([&] {
    dav_check_params(ci, params);
    PQG_REQUIRE(E, PQG_ERR_ARG, "NULL argument");
    hipStream_t st = ci->stream;
    Davidson dav(ci, *params);
    while (!dav.done()) {
      dav.run_sigma(st);
      dav.post_sigma(st);
    }
    dav_finish(ci, dav, E, X, ldx, report);
  });
}

extern "C" pqg_status pqg_dav_create(pqg_ci* ci, const pqg_dav_params* params, pqg_dav** out) {
  return guard([&] {
    dav_check_params(ci, params);
    PQG_REQUIRE(out, PQG_ERR_ARG, "NULL argument");
    *out = nullptr;
    *out = new pqg_dav(ci, *params);
  });
}

extern "C" pqg_status pqg_dav_step(pqg_dav* d, int* done) {
  return guard([&] {
    PQG_REQUIRE(d && done, PQG_ERR_ARG, "NULL argument");
    if (!d->dav.done()) {
      hipStream_t st = d->ci->stream;
      d->dav.run_sigma(st);
      d->dav.post_sigma(st);
    }
    *done = d->dav.done() ? 1 : 0;
  });
}

extern "C" pqg_status pqg_dav_result(pqg_dav* d, double* E, double* X, int64_t ldx, pqg_dav_report* report) {
  return guard([&] {
    PQG_REQUIRE(d, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(d->dav.done(), PQG_ERR_ARG, "the solve has not finished: call pqg_dav_step until done");
    dav_finish(d->ci, d->dav, E, X, ldx, report);
  });
}

extern "C" pqg_status pqg_dav_destroy(pqg_dav* d) {
  return guard([&] {
    if (!d) return;
    (void)hipStreamSynchronize(d->ci->stream);
    delete d;
  });
}
