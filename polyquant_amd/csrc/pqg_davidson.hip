// Davidson on the GPU: block of nev Ritz pairs, diagonal preconditioner, Rayleigh-Ritz in FP64.
//
// Restates the behaviour of the eigensolver the reference drives at src/ci/epci.cpp:589-621
// (Spectra DavidsonSymEigsSolver, SortRule::SmallestAlge): start from unit vectors at the
// nvec_init smallest diagonal entries, extend by t_k = r_k / (theta_k - diag), restart on the
// current Ritz vectors when the basis would exceed nvec_max, converge when every one of the
// nev residual norms is below tol.
//
// Everything stays in HBM: tall-skinny products (V^T W, V Y, Gram-Schmidt) are FP64 kernels
// with a deterministic two-level reduction, and the projected eigenproblem is solved by a
// one-workgroup parallel-order Jacobi kernel in FP64.  Only the nev residual norms travel to
// the host per iteration (convergence decision).
//
// Mixed precision (PQG_MIXED): sigma of new search directions runs in FP32.  When the
// FP32-based residual has dropped below max(tol, 0.05 * last true residual), sigma of the
// Ritz vectors is recomputed in FP64 ("true" residual).  Converged if that is below tol;
// otherwise the basis restarts on the Ritz vectors with their exact FP64 sigma (defect
// correction) and extends with corrections built from the true residuals.
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

// out[a][b] partials: sum over a row chunk of A[r][a] * B[r][b], a in [a0, a0+8), b in [b0, b0+4)
__global__ __launch_bounds__(kTB) void k_tn_partial(const double* __restrict__ A, int64_t lda, int na,
                                                    const double* __restrict__ B, int64_t ldb, int nb, int64_t nrows,
                                                    int64_t chunk, double* __restrict__ part) {
  const int a0 = blockIdx.y * 8, b0 = blockIdx.z * 4;
  const int64_t r0 = (int64_t)blockIdx.x * chunk;
  const int64_t r1 = min(nrows, r0 + chunk);
  double acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kTB) {
    double bv[4], av[8];
#pragma unroll
    for (int j = 0; j < 4; j++) bv[j] = (b0 + j < nb) ? B[r + (b0 + j) * ldb] : 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) av[i] = (a0 + i < na) ? A[r + (a0 + i) * lda] : 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
      for (int j = 0; j < 4; j++) acc[i][j] += av[i] * bv[j];
  }
  __shared__ double red[kTB / 64][32];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < 8; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) {
      double v = acc[i][j];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
      if (lane == 0) red[wv][i * 4 + j] = v;
    }
  __syncthreads();
  if (threadIdx.x < 32) {
    const int i = threadIdx.x / 4, j = threadIdx.x % 4;
    double v = 0.0;
    for (int w = 0; w < kTB / 64; w++) v += red[w][threadIdx.x];
    if (a0 + i < na && b0 + j < nb) part[((size_t)blockIdx.x * na + (a0 + i)) * nb + (b0 + j)] = v;
  }
}

__global__ void k_tn_reduce(const double* __restrict__ part, int nchunk, int na, int nb, double* __restrict__ out,
                            int ldo) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= na * nb) return;
  const int a = idx / nb, b = idx % nb;
  double v = 0.0;
  for (int c = 0; c < nchunk; c++) v += part[((size_t)c * na + a) * nb + b];
  out[a + b * ldo] = v;
}

// out[r][j] = beta * out[r][j] + alpha * sum_c A[r][c] Bs[c][j]   (Bs small, ka x nb, in LDS)
__global__ __launch_bounds__(kTB) void k_nn(const double* __restrict__ A, int64_t lda, int ka,
                                            const double* __restrict__ Bs, int ldb, int nb, double* __restrict__ out,
                                            int64_t ldo, int64_t nrows, double alpha, double beta) {
  __shared__ double bs[64 * 64];
  for (int i = threadIdx.x; i < ka * nb; i += kTB) {
    const int c = i % ka, j = i / ka;
    bs[c + j * 64] = Bs[c + j * ldb];
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * kTB + threadIdx.x;
  if (r >= nrows) return;
  for (int j0 = 0; j0 < nb; j0 += 8) {
    double acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = 0.0;
    for (int c = 0; c < ka; c++) {
      const double a = A[r + c * lda];
#pragma unroll
      for (int j = 0; j < 8; j++)
        if (j0 + j < nb) acc[j] += a * bs[c + (j0 + j) * 64];
    }
#pragma unroll
    for (int j = 0; j < 8; j++)
      if (j0 + j < nb) {
        double* o = out + r + (j0 + j) * ldo;
        *o = (beta == 0.0 ? 0.0 : beta * *o) + alpha * acc[j];
      }
  }
}

// R = AX - X theta, then T = R / (theta - diag) (safeguarded); invalid entries (diag >= 1e29)
// give zero corrections.
__global__ void k_residual(const double* __restrict__ X, const double* __restrict__ AX, const double* __restrict__ theta,
                           const double* __restrict__ Hd, double* __restrict__ R, double* __restrict__ T, int64_t n,
                           int nev) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = Hd[i];
  for (int k = 0; k < nev; k++) {
    const double r = AX[i + k * n] - theta[k] * X[i + k * n];
    R[i + k * n] = r;
    double den = theta[k] - d;
    if (fabs(den) < 1e-8) den = den < 0 ? -1e-8 : 1e-8;
    T[i + k * n] = (d >= 1e29) ? 0.0 : r / den;
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
__global__ void k_scale_cols(double* __restrict__ v, int64_t n, int64_t ld, const double* __restrict__ s, int k) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < k; j++) v[i + j * ld] *= s[j];
}

// Parallel-order (round-robin) cyclic Jacobi for a symmetric m x m matrix, m <= 64, in one
// workgroup, warm-started: with P = blockdiag(Yp (warm x warm), I), it diagonalises P^T G P
// and accumulates the rotations onto P, so a projected matrix that grew by a few rows since the
// previous iteration needs fewer sweeps.  In each step the indices form m/2 disjoint pairs
// (circle method); thread (bi, bj) owns the 2x2 block of A at the rows of pair bi and the
// columns of pair bj, and two rows of V at the columns of pair bj, so the rotation of a step
// is applied in place with no exchange between threads.  Outputs ascending eigenvalues and the
// matching eigenvectors (columns of Y).
__device__ __forceinline__ void rr_pair(int mm, int step, int k, int& p, int& q) {
  // circle method: index mm-1 stays, the others rotate
  const int n1 = mm - 1;
  if (k == 0) {
    p = n1;
    q = step;
  } else {
    p = (step + k) % n1;
    q = (step - k + n1) % n1;
  }
  if (p > q) {
    const int t = p;
    p = q;
    q = t;
  }
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

__global__ __launch_bounds__(1024) void k_jacobi(const double* __restrict__ G, int ldg, int m,
                                                 const double* __restrict__ Yp, int warm, double* __restrict__ theta,
                                                 double* __restrict__ Y, int ldy) {
  __shared__ double A[64][65];
  __shared__ double Vv[64][65];
  __shared__ double Tm[64][65];
  __shared__ double cs[32], sn[32];
  __shared__ int pp[32], qq[32];
  __shared__ double red[16];
  __shared__ int order[64];
  const int tid = threadIdx.x;
  const int mm = m + (m & 1), half = mm / 2;
  for (int i = tid; i < 64 * 64; i += 1024) {
    const int r = i / 64, c = i % 64;
    double v = (r == c) ? 1.0 : 0.0;
    if (r < warm && c < warm) v = Yp[r + c * 64];
    Vv[r][c] = v;
    A[r][c] = (r < m && c < m) ? 0.5 * (G[r + c * ldg] + G[c + r * ldg]) : 0.0;
  }
  __syncthreads();
  if (warm > 0) {  // A = P^T Gs P
    for (int i = tid; i < m * m; i += 1024) {
      const int r = i / m, c = i % m;
      double v = 0.0;
      for (int k = 0; k < m; k++) v += A[r][k] * Vv[k][c];
      Tm[r][c] = v;
    }
    __syncthreads();
    for (int i = tid; i < m * m; i += 1024) {
      const int r = i / m, c = i % m;
      double v = 0.0;
      for (int k = 0; k < m; k++) v += Vv[k][r] * Tm[k][c];
      A[r][c] = v;
    }
    __syncthreads();
  }
  const int bi = tid >> 5, bj = tid & 31;
  const bool active = bi < half && bj < half;
  const int vr0 = 2 * bi;  // rows of V owned by this thread
  for (int sweep = 0; sweep < 40; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int i = tid; i < m * m; i += 1024) {
      const int r = i / m, c = i % m;
      const double a = A[r][c];
      if (r != c)
        off += a * a;
      else
        dia += a * a;
    }
    const double offs = block_sum(off, red);
    __syncthreads();
    const double dias = block_sum(dia, red);
    if (offs <= 1e-30 * dias || offs < 1e-300) break;
    for (int step = 0; step < mm - 1; step++) {
      __syncthreads();  // previous step's updates are visible
      if (tid < half) {
        int p, q;
        rr_pair(mm, step, tid, p, q);
        double c = 1.0, s = 0.0;
        if (q < m) {
          const double apq = A[p][q];
          if (fabs(apq) > 1e-300) {
            const double th = (A[q][q] - A[p][p]) / (2.0 * apq);
            const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        cs[tid] = c;
        sn[tid] = s;
        pp[tid] = p;
        qq[tid] = q;
      }
      __syncthreads();
      if (active) {
        // new column p = c*col p - s*col q, new column q = s*col p + c*col q (same for rows)
        const int pi = pp[bi], qi = qq[bi], pj = pp[bj], qj = qq[bj];
        const double ci = cs[bi], si = sn[bi], cj = cs[bj], sj = sn[bj];
        const bool qiok = qi < m, qjok = qj < m;
        const double a00 = A[pi][pj];
        const double a01 = qjok ? A[pi][qj] : 0.0;
        const double a10 = qiok ? A[qi][pj] : 0.0;
        const double a11 = (qiok && qjok) ? A[qi][qj] : 0.0;
        // rows
        const double r00 = ci * a00 - si * a10, r01 = ci * a01 - si * a11;
        const double r10 = si * a00 + ci * a10, r11 = si * a01 + ci * a11;
        // columns
        const double n00 = cj * r00 - sj * r01, n01 = sj * r00 + cj * r01;
        const double n10 = cj * r10 - sj * r11, n11 = sj * r10 + cj * r11;
        const bool diag = bi == bj;  // the rotated pair's coupling is annihilated exactly
        A[pi][pj] = n00;
        if (qjok) A[pi][qj] = diag ? 0.0 : n01;
        if (qiok) A[qi][pj] = diag ? 0.0 : n10;
        if (qiok && qjok) A[qi][qj] = n11;
#pragma unroll
        for (int e = 0; e < 2; e++) {
          const int r = vr0 + e;
          if (r < m) {
            const double vp = Vv[r][pj], vq = qjok ? Vv[r][qj] : 0.0;
            Vv[r][pj] = cj * vp - sj * vq;
            if (qjok) Vv[r][qj] = sj * vp + cj * vq;
          }
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  if (tid < m) {  // rank of each eigenvalue (ties by index)
    const double d = A[tid][tid];
    int rk = 0;
    for (int j = 0; j < m; j++) {
      const double e = A[j][j];
      rk += (e < d || (e == d && j < tid)) ? 1 : 0;
    }
    order[rk] = tid;
  }
  __syncthreads();
  for (int i = tid; i < m; i += 1024) theta[i] = A[order[i]][order[i]];
  for (int i = tid; i < m * m; i += 1024) {
    const int r = i % m, c = i / m;
    Y[r + c * ldy] = Vv[r][order[c]];
  }
}

static int g1(int64_t n, int b) { return (int)((n + b - 1) / b); }

// G (64 x 64): columns [old, m) are new.  Fill their mirror rows; gt (newc x old, ld 64),
// when given, supplies the new-vs-old couplings.
__global__ void k_symmetrize(double* __restrict__ G, int old, int m, const double* __restrict__ gt) {
  for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
    const int r = i % 64, c = i / 64;
    if (c < old || c >= m || r >= m) continue;
    if (r < old) {
      const double v = gt ? gt[(c - old) + r * 64] : G[r + c * 64];
      G[r + c * 64] = v;
      G[c + r * 64] = v;
    } else if (r > c) {
      const double v = 0.5 * (G[r + c * 64] + G[c + r * 64]);
      G[r + c * 64] = v;
      G[c + r * 64] = v;
    }
  }
}

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
  N_ = ci->n_int;
  nev_ = std::max(1, p.nev);
  PQG_REQUIRE(nev_ <= 64 && nev_ <= ci->n_ref, PQG_ERR_ARG, "nev out of range");
  ninit_ = std::max(p.nvec_init, nev_);
  nmax_ = p.nvec_max > 0 ? p.nvec_max : std::max(2 * p.nvec_init, 10 * nev_);
  ninit_ = (int)std::min<int64_t>(ninit_, ci->n_ref);
  nmax_ = (int)std::min<int64_t>(std::max(nmax_, ninit_ + nev_), std::min<int64_t>(ci->n_ref, 64));
  PQG_REQUIRE(nmax_ >= ninit_, PQG_ERR_ARG, "subspace too small");
  mixed_ = p.precision == PQG_MIXED;
  pqg_ctx* ctx = ci->ctx;
  V_.reset(ctx, (size_t)N_ * nmax_);
  W_.reset(ctx, (size_t)N_ * nmax_);
  G_.reset(ctx, 64 * 64);
  Y_.reset(ctx, 64 * 64);
  Yprev_.reset(ctx, 64 * 64);
  theta_.reset(ctx, 64);
  X_.reset(ctx, (size_t)N_ * nev_);
  AX_.reset(ctx, (size_t)N_ * nev_);
  R_.reset(ctx, (size_t)N_ * nev_);
  T_.reset(ctx, (size_t)N_ * nev_);
  tmp_.reset(ctx, (size_t)N_ * ninit_);
  small_.reset(ctx, 64 * 64);
  nchunk_ = (int)std::max<int64_t>(1, std::min<int64_t>(512, N_ / 4096));
  part_.reset(ctx, (size_t)nchunk_ * 64 * 64);
  if (mixed_) {
    W64_.reset(ctx, (size_t)N_ * nev_);
    c32_.reset(ctx, (size_t)N_ * nev_ > (size_t)N_ * ninit_ ? (size_t)N_ * nev_ : (size_t)N_ * ninit_);
    s32_.reset(ctx, c32_.n);
    ensure_mixed(ci);
  }
  t0_ = std::chrono::steady_clock::now();
  initial_guess(ci->stream);
}

uint64_t Davidson::device_bytes() const {
  return V_.bytes() + W_.bytes() + X_.bytes() + AX_.bytes() + R_.bytes() + T_.bytes() + tmp_.bytes() + W64_.bytes() +
         c32_.bytes() + s32_.bytes() + part_.bytes();
}

void Davidson::ts_gemm_tn(hipStream_t st, const double* A, int64_t lda, int na, const double* B, int64_t ldb, int nb,
                          double* out, int ldo) {
  const int64_t chunk = (N_ + nchunk_ - 1) / nchunk_;
  dim3 g(nchunk_, (na + 7) / 8, (nb + 3) / 4);
  k_tn_partial<<<g, kTB, 0, st>>>(A, lda, na, B, ldb, nb, N_, chunk, part_.get());
  PQG_HIP(hipGetLastError());
  k_tn_reduce<<<g1(na * nb, 256), 256, 0, st>>>(part_.get(), nchunk_, na, nb, out, ldo);
  PQG_HIP(hipGetLastError());
}

void Davidson::ts_gemm_nn(hipStream_t st, const double* A, int64_t lda, int ka, const double* Bs, int ldb, int nb,
                          double* out, int64_t ldo, double alpha, double beta) {
  k_nn<<<g1(N_, kTB), kTB, 0, st>>>(A, lda, ka, Bs, ldb, nb, out, ldo, N_, alpha, beta);
  PQG_HIP(hipGetLastError());
}

void Davidson::initial_guess(hipStream_t st) {
  // unit vectors at the ninit smallest diagonal entries (ties: lower index first)
  std::vector<double> hd(N_);
  PQG_HIP(hipMemcpyAsync(hd.data(), ci_->hdiag.get(), N_ * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  std::vector<int64_t> idx(N_);
  std::iota(idx.begin(), idx.end(), 0);
  std::partial_sort(idx.begin(), idx.begin() + ninit_, idx.end(),
                    [&](int64_t a, int64_t b) { return hd[a] < hd[b] || (hd[a] == hd[b] && a < b); });
  DevBuf<int64_t> pos(ci_->ctx, ninit_);
  PQG_HIP(hipMemcpyAsync(pos.get(), idx.data(), ninit_ * 8, hipMemcpyHostToDevice, st));
  k_unit<<<g1(N_, 256), 256, 0, st>>>(V_.get(), N_, N_, pos.get(), ninit_);
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipStreamSynchronize(st));
  m_ = ninit_;
  nW_ = 0;
  work_ = SigmaWork{m_, mixed_ ? PQG_MIXED : PQG_FP64, V_.get(), W_.get(), N_};
}

void Davidson::run_sigma(hipStream_t st) {
  if (work_.k == 0) return;
  const int k = work_.k;
  if (work_.precision == PQG_FP64) {
    sigma_internal(ci_, work_.C, work_.S, k, work_.ld, st);
    sigma_calls64_ += k;
  } else {
    const int64_t n = (int64_t)N_ * k;
    if (c32_.n < (size_t)n) {
      c32_.reset(ci_->ctx, n);
      s32_.reset(ci_->ctx, n);
    }
    for (int j = 0; j < k; j++)
      k_d2f_blk<<<g1(N_, 256), 256, 0, st>>>(work_.C + j * work_.ld, c32_.get() + (int64_t)j * N_, N_);
    sigma_internal(ci_, c32_.get(), s32_.get(), k, N_, st);
    for (int j = 0; j < k; j++)
      k_f2d_blk<<<g1(N_, 256), 256, 0, st>>>(s32_.get() + (int64_t)j * N_, work_.S + j * work_.ld, N_);
    PQG_HIP(hipGetLastError());
  }
  sigma_calls_ += k;
}

// Cholesky factor of a small Gram matrix B (n x n, leading dimension 64) with column dropping:
// column j is kept when its residual norm after the kept columns before it exceeds thr[j].
// Returns M (n x nk, leading dimension 64) with Q = T M orthonormal, and the kept count.
static int small_cholqr(const std::vector<double>& B, int n, const std::vector<double>& thr, std::vector<double>& M) {
  std::vector<int> kept;
  std::vector<double> R(64 * 64, 0.0);  // R[i + j*64], i <= j, over original column indices
  for (int j = 0; j < n; j++) {
    double d = B[j + j * 64];
    for (size_t a = 0; a < kept.size(); a++) {
      const int i = kept[a];
      double v = B[i + j * 64];
      for (size_t b = 0; b < a; b++) v -= R[kept[b] + i * 64] * R[kept[b] + j * 64];
      R[i + j * 64] = v / R[i + i * 64];
      d -= R[i + j * 64] * R[i + j * 64];
    }
    if (d > thr[j] * thr[j] && d > 0.0) {
      R[j + j * 64] = std::sqrt(d);
      kept.push_back(j);
    }
  }
  const int nk = (int)kept.size();
  // inverse of the kept upper-triangular factor, placed on the kept rows of M
  std::vector<double> Rk(64 * 64, 0.0), Ri(64 * 64, 0.0);
  for (int a = 0; a < nk; a++)
    for (int b = a; b < nk; b++) Rk[a + b * 64] = R[kept[a] + kept[b] * 64];
  for (int b = 0; b < nk; b++) {
    Ri[b + b * 64] = 1.0 / Rk[b + b * 64];
    for (int a = b - 1; a >= 0; a--) {
      double v = 0.0;
      for (int c = a + 1; c <= b; c++) v += Rk[a + c * 64] * Ri[c + b * 64];
      Ri[a + b * 64] = -v / Rk[a + a * 64];
    }
  }
  M.assign(64 * 64, 0.0);
  for (int a = 0; a < nk; a++)
    for (int b = 0; b < nk; b++) M[kept[a] + b * 64] = Ri[a + b * 64];
  return nk;
}

// Orthogonalise T (nt columns) against V (m columns) with two block Gram-Schmidt sweeps, then
// orthonormalise the block by two Cholesky-QR passes (columns that vanish are dropped), and
// append it to V.  Returns how many columns were appended.
int Davidson::orthonormalize_append(hipStream_t st, double* T, int nt) {
  nt = std::min(nt, nmax_ - m_);
  if (nt <= 0) return 0;
  std::vector<double> g(64 * 64), M;
  ts_gemm_tn(st, T, N_, nt, T, N_, nt, small_.get(), 64);
  PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  std::vector<double> thr(nt);
  for (int j = 0; j < nt; j++) thr[j] = std::max(1e-10 * std::sqrt(std::max(g[j + j * 64], 0.0)), 1e-30);
  for (int pass = 0; pass < 2; pass++) {
    ts_gemm_tn(st, V_.get(), N_, m_, T, N_, nt, small_.get(), 64);
    ts_gemm_nn(st, V_.get(), N_, m_, small_.get(), 64, nt, T, N_, -1.0, 1.0);
  }
  // pass 1: drop vanishing columns, orthonormalise the rest straight into V[:, m:]
  ts_gemm_tn(st, T, N_, nt, T, N_, nt, small_.get(), 64);
  PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  const int nk = small_cholqr(g, nt, thr, M);
  if (nk == 0) return 0;
  double* dst = V_.get() + (int64_t)m_ * N_;
  PQG_HIP(hipMemcpyAsync(small_.get(), M.data(), 64 * 64 * 8, hipMemcpyHostToDevice, st));
  ts_gemm_nn(st, T, N_, nt, small_.get(), 64, nk, dst, N_, 1.0, 0.0);
  // pass 2 (CholQR2): re-orthonormalise the new block
  ts_gemm_tn(st, dst, N_, nk, dst, N_, nk, small_.get(), 64);
  PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  std::vector<double> thr2(nk, 0.0);
  const int nk2 = small_cholqr(g, nk, thr2, M);
  if (nk2 != nk) return 0;  // numerically rank deficient: do not extend
  PQG_HIP(hipMemcpyAsync(small_.get(), M.data(), 64 * 64 * 8, hipMemcpyHostToDevice, st));
  ts_gemm_nn(st, dst, N_, nk, small_.get(), 64, nk, tmp_.get(), N_, 1.0, 0.0);
  PQG_HIP(hipMemcpyAsync(dst, tmp_.get(), (size_t)N_ * nk * 8, hipMemcpyDeviceToDevice, st));
  mask_invalid<double>(ci_, dst, N_, nk, st);
  m_ += nk;
  return nk;
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
    ts_gemm_tn(st, V_.get(), N_, m_, W_.get() + (int64_t)nW_ * N_, N_, newc, G_.get() + (int64_t)nW_ * 64, 64);
    if (mixed_ && old > 0) {
      ts_gemm_tn(st, V_.get() + (int64_t)old * N_, N_, newc, W_.get(), N_, old, small_.get(), 64);
      gt = small_.get();
    }
    nW_ = m_;
  }
  k_symmetrize<<<1, 256, 0, st>>>(G_.get(), old, m_, gt);
  PQG_HIP(hipGetLastError());
  k_jacobi<<<1, 1024, 0, st>>>(G_.get(), 64, m_, Yprev_.get(), std::min(warm_, m_), theta_.get(), Y_.get(), 64);
  PQG_HIP(hipGetLastError());
  PQG_HIP(hipMemcpyAsync(Yprev_.get(), Y_.get(), 64 * 64 * 8, hipMemcpyDeviceToDevice, st));
  warm_ = m_;
}

void Davidson::restart(hipStream_t st, const double* Vk, const double* Wk, int keep) {
  // V[:, :keep] = Vk, W[:, :keep] = Wk; projected matrix becomes diag(theta)
  PQG_HIP(hipMemcpyAsync(V_.get(), Vk, (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
  PQG_HIP(hipMemcpyAsync(W_.get(), Wk, (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
  m_ = nW_ = 0;
  // recompute the keep x keep projection exactly
  ts_gemm_tn(st, V_.get(), N_, keep, W_.get(), N_, keep, G_.get(), 64);
  m_ = nW_ = keep;
  warm_ = 0;
  restarts_++;
}

void Davidson::post_sigma(hipStream_t st) {
  if (done_) return;
  if (checking_) {
    // FP64 sigma of the Ritz vectors is in W64_: true residuals
    checking_ = false;
    since_check_ = 0;
    {  // final energies: Rayleigh quotients x^T H x of the (orthonormal) Ritz vectors in FP64
      ts_gemm_tn(st, X_.get(), N_, nev_, W64_.get(), N_, nev_, small_.get(), 64);
      std::vector<double> q(64 * 64);
      PQG_HIP(hipMemcpyAsync(q.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
      for (int k = 0; k < nev_; k++) theta_h_[k] = q[k + k * 64];
      PQG_HIP(hipMemcpyAsync(theta_.get(), theta_h_.data(), nev_ * 8, hipMemcpyHostToDevice, st));
    }
    k_residual<<<g1(N_, 256), 256, 0, st>>>(X_.get(), W64_.get(), theta_.get(), ci_->hdiag.get(), R_.get(), T_.get(),
                                            N_, nev_);
    PQG_HIP(hipGetLastError());
    ts_gemm_tn(st, R_.get(), N_, nev_, R_.get(), N_, nev_, small_.get(), 64);
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
    ts_gemm_nn(st, V_.get(), N_, m_, Y_.get(), 64, keep, tmp_.get(), N_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
    ts_gemm_nn(st, W_.get(), N_, m_, Y_.get(), 64, keep, tmp_.get(), N_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(W_.get(), tmp_.get(), (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
    PQG_HIP(hipMemcpyAsync(W_.get(), W64_.get(), (size_t)N_ * nev_ * 8, hipMemcpyDeviceToDevice, st));
    m_ = nW_ = keep;
    warm_ = 0;
    // projected matrix: couplings to the exact columns are taken as v^T (H x) (FP64)
    ts_gemm_tn(st, V_.get(), N_, keep, W_.get(), N_, keep, small_.get(), 64);
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
    orthonormalize_append(st, T_.get(), nev_);
    work_ = SigmaWork{m_ - before, PQG_MIXED, V_.get() + (int64_t)before * N_, W_.get() + (int64_t)before * N_, N_};
    if (work_.k == 0) {  // nothing new to add: accept
      done_ = true;
      seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }
    return;
  }
  iter_++;
  rayleigh_ritz(st);
  // Ritz vectors and residuals
  ts_gemm_nn(st, V_.get(), N_, m_, Y_.get(), 64, nev_, X_.get(), N_, 1.0, 0.0);
  ts_gemm_nn(st, W_.get(), N_, m_, Y_.get(), 64, nev_, AX_.get(), N_, 1.0, 0.0);
  k_residual<<<g1(N_, 256), 256, 0, st>>>(X_.get(), AX_.get(), theta_.get(), ci_->hdiag.get(), R_.get(), T_.get(), N_,
                                          nev_);
  PQG_HIP(hipGetLastError());
  ts_gemm_tn(st, R_.get(), N_, nev_, R_.get(), N_, nev_, small_.get(), 64);
  std::vector<double> g(64 * 64);
  theta_h_.resize(64);
  rn_h_.resize(nev_);
  PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipMemcpyAsync(theta_h_.data(), theta_.get(), 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  theta_h_.resize(nev_);
  double worst = 0.0;
  for (int k = 0; k < nev_; k++) {
    rn_h_[k] = std::sqrt(std::max(g[k + k * 64], 0.0));
    worst = std::max(worst, rn_h_[k]);
  }
  dav_log(mixed_ ? "fp32" : "fp64", iter_, m_, p_.tol, theta_h_, rn_h_);
  since_check_++;
  const double thr = std::max(p_.tol, std::min(1e-3, 0.05 * last_true_));
  const bool approx_conv = mixed_ ? (worst < thr || since_check_ >= 12) : (worst < p_.tol);
  if (approx_conv && mixed_) {
    // FP64 check of the Ritz vectors
    checking_ = true;
    work_ = SigmaWork{nev_, PQG_FP64, X_.get(), W64_.get(), N_};
    return;
  }
  if (approx_conv || iter_ >= p_.maxit) {
    done_ = true;
    status_ = approx_conv ? PQG_OK : PQG_ERR_NOCONV;
    work_ = SigmaWork{};
    seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    return;
  }
  if (m_ + nev_ > nmax_) {
    const int keep = ninit_;
    ts_gemm_nn(st, V_.get(), N_, m_, Y_.get(), 64, keep, tmp_.get(), N_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(V_.get(), tmp_.get(), (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
    ts_gemm_nn(st, W_.get(), N_, m_, Y_.get(), 64, keep, tmp_.get(), N_, 1.0, 0.0);
    PQG_HIP(hipMemcpyAsync(W_.get(), tmp_.get(), (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
    m_ = nW_ = keep;
    warm_ = 0;
    // projection after restart = diag(theta[:keep]) up to rounding; recompute exactly
    ts_gemm_tn(st, V_.get(), N_, keep, W_.get(), N_, keep, G_.get(), 64);
    restarts_++;
  }
  const int before = m_;
  orthonormalize_append(st, T_.get(), nev_);
  const int added = m_ - before;
  if (added == 0) {  // subspace exhausted
    done_ = true;
    status_ = worst < p_.tol ? PQG_OK : PQG_ERR_NOCONV;
    work_ = SigmaWork{};
    seconds_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    return;
  }
  work_ = SigmaWork{added, mixed_ ? PQG_MIXED : PQG_FP64, V_.get() + (int64_t)before * N_,
                    W_.get() + (int64_t)before * N_, N_};
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

extern "C" pqg_status pqg_ci_davidson(pqg_ci* ci, const pqg_dav_params* params, double* E, double* X, int64_t ldx,
                                      pqg_dav_report* report) {
  return guard([&] {
    PQG_REQUIRE(ci && params && E, PQG_ERR_ARG, "NULL argument");
    PQG_REQUIRE(params->precision == PQG_FP64 || params->precision == PQG_MIXED, PQG_ERR_ARG, "unknown precision");
    PQG_REQUIRE(params->tol > 0 && params->maxit > 0, PQG_ERR_ARG, "tol and maxit must be positive");
    hipStream_t st = ci->stream;
    Davidson dav(ci, *params);
    while (!dav.done()) {
      dav.run_sigma(st);
      dav.post_sigma(st);
    }
    PQG_HIP(hipStreamSynchronize(st));
    const auto& ev = dav.energies();
    for (int k = 0; k < dav.nev(); k++) E[k] = ev[k];
    if (X) {
      PQG_REQUIRE(ldx >= ci->n_ref, PQG_ERR_ARG, "ldx smaller than N");
      DevBuf<double> xr(ci->ctx, (size_t)ci->n_ref * dav.nev());
      to_reference(ci, dav.ritz_vectors(), ci->n_int, xr.get(), ci->n_ref, dav.nev(), st);
      PQG_HIP(hipMemcpy2DAsync(X, ldx * 8, xr.get(), ci->n_ref * 8, ci->n_ref * 8, dav.nev(), hipMemcpyDeviceToHost, st));
      PQG_HIP(hipStreamSynchronize(st));
    }
    if (report) *report = dav.report();
    if (dav.status() != PQG_OK) throw Error{dav.status(), "CI Calculation did not converge!"};
  });
}
