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

// Parallel-order (round-robin) cyclic Jacobi for a symmetric m x m matrix, m <= 64, one
// workgroup.  Outputs ascending eigenvalues and the matching eigenvectors (columns of Y).
__global__ __launch_bounds__(256) void k_jacobi(const double* __restrict__ G, int ldg, int m, double* __restrict__ theta,
                                                double* __restrict__ Y, int ldy) {
  __shared__ double A[64][65];
  __shared__ double Vv[64][65];
  __shared__ int P[64];
  __shared__ double cs[32], sn[32];
  __shared__ double red[256];
  __shared__ int order[64];
  const int tid = threadIdx.x;
  const int mm = m + (m & 1);
  const int np = mm / 2;
  for (int i = tid; i < 64 * 64; i += 256) {
    const int r = i / 64, c = i % 64;
    A[r][c] = (r < m && c < m) ? 0.5 * (G[r + c * ldg] + G[c + r * ldg]) : 0.0;
    Vv[r][c] = (r == c) ? 1.0 : 0.0;
  }
  if (tid < 64) P[tid] = tid;
  __syncthreads();
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, dia = 0.0;
    for (int i = tid; i < m * m; i += 256) {
      const int r = i / m, c = i % m;
      if (r != c)
        off += A[r][c] * A[r][c];
      else
        dia += A[r][c] * A[r][c];
    }
    red[tid] = off;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    const double offs = red[0];
    __syncthreads();
    red[tid] = dia;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    const double dias = red[0];
    __syncthreads();
    if (offs <= 1e-30 * dias || offs < 1e-300) break;
    for (int step = 0; step < mm - 1; step++) {
      if (tid < np) {
        int p = P[tid], q = P[mm - 1 - tid];
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
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
      }
      __syncthreads();
      for (int i = tid; i < np * m; i += 256) {  // rows
        const int k = i / m, j = i % m;
        int p = P[k], q = P[mm - 1 - k];
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q >= m) continue;
        const double c = cs[k], s = sn[k];
        const double ap = A[p][j], aq = A[q][j];
        A[p][j] = c * ap - s * aq;
        A[q][j] = s * ap + c * aq;
      }
      __syncthreads();
      for (int i = tid; i < np * m; i += 256) {  // columns + eigenvectors
        const int k = i / m, j = i % m;
        int p = P[k], q = P[mm - 1 - k];
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        if (q >= m) continue;
        const double c = cs[k], s = sn[k];
        const double ap = A[j][p], aq = A[j][q];
        A[j][p] = c * ap - s * aq;
        A[j][q] = s * ap + c * aq;
        const double vp = Vv[j][p], vq = Vv[j][q];
        Vv[j][p] = c * vp - s * vq;
        Vv[j][q] = s * vp + c * vq;
      }
      __syncthreads();
      if (tid == 0) {  // rotate players 1..mm-1
        const int last = P[mm - 1];
        for (int i = mm - 1; i > 1; i--) P[i] = P[i - 1];
        P[1] = last;
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    for (int i = 0; i < m; i++) order[i] = i;
    for (int i = 0; i < m; i++)
      for (int j = i + 1; j < m; j++)
        if (A[order[j]][order[j]] < A[order[i]][order[i]]) {
          const int t = order[i];
          order[i] = order[j];
          order[j] = t;
        }
  }
  __syncthreads();
  for (int i = tid; i < m; i += 256) theta[i] = A[order[i]][order[i]];
  for (int i = tid; i < m * m; i += 256) {
    const int r = i % m, c = i / m;
    Y[r + c * ldy] = Vv[r][order[c]];
  }
}

static int g1(int64_t n, int b) { return (int)((n + b - 1) / b); }

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

// Orthogonalise T (nt columns) against V (m columns) twice, then among themselves (MGS);
// appends the surviving columns to V.  Returns how many were appended.
int Davidson::orthonormalize_append(hipStream_t st, double* T, int nt) {
  std::vector<double> nrm0(nt), nrm(nt);
  // norms before
  ts_gemm_tn(st, T, N_, nt, T, N_, nt, small_.get(), 64);
  std::vector<double> g(64 * 64);
  PQG_HIP(hipMemcpyAsync(g.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  for (int j = 0; j < nt; j++) nrm0[j] = std::sqrt(std::max(g[j + j * 64], 0.0));
  for (int pass = 0; pass < 2; pass++) {
    ts_gemm_tn(st, V_.get(), N_, m_, T, N_, nt, small_.get(), 64);
    ts_gemm_nn(st, V_.get(), N_, m_, small_.get(), 64, nt, T, N_, -1.0, 1.0);
  }
  int added = 0;
  for (int j = 0; j < nt && m_ < nmax_; j++) {
    double* t = T + (int64_t)j * N_;
    double* dst = V_.get() + (int64_t)m_ * N_;
    for (int pass = 0; pass < 2; pass++) {
      if (added > 0) {
        const double* Vnew = V_.get() + (int64_t)(m_ - added) * N_;
        ts_gemm_tn(st, Vnew, N_, added, t, N_, 1, small_.get(), 64);
        ts_gemm_nn(st, Vnew, N_, added, small_.get(), 64, 1, t, N_, -1.0, 1.0);
      }
    }
    ts_gemm_tn(st, t, N_, 1, t, N_, 1, small_.get(), 64);
    double nn = 0.0;
    PQG_HIP(hipMemcpyAsync(&nn, small_.get(), 8, hipMemcpyDeviceToHost, st));
    PQG_HIP(hipStreamSynchronize(st));
    nn = std::sqrt(std::max(nn, 0.0));
    if (!(nn > 1e-10 * std::max(nrm0[j], 1e-300)) || nn < 1e-30) continue;
    const double inv = 1.0 / nn;
    PQG_HIP(hipMemcpyAsync(small_.get(), &inv, 8, hipMemcpyHostToDevice, st));
    PQG_HIP(hipMemcpyAsync(dst, t, N_ * 8, hipMemcpyDeviceToDevice, st));
    k_scale_cols<<<g1(N_, 256), 256, 0, st>>>(dst, N_, N_, small_.get(), 1);
    PQG_HIP(hipGetLastError());
    PQG_HIP(hipStreamSynchronize(st));
    m_++;
    added++;
  }
  mask_invalid<double>(ci_, V_.get() + (int64_t)(m_ - added) * N_, N_, added, st);
  return added;
}

void Davidson::rayleigh_ritz(hipStream_t st) {
  // G[:, nW_old:m] = V^T W[:, nW_old:m]; the rows of the new columns against older columns
  // follow by symmetry.
  const int old = nW_;
  const int newc = m_ - nW_;
  if (newc > 0) {
    ts_gemm_tn(st, V_.get(), N_, m_, W_.get() + (int64_t)nW_ * N_, N_, newc, G_.get() + (int64_t)nW_ * 64, 64);
    nW_ = m_;
  }
  // Mixed precision: the couplings of new directions to older basis vectors are taken as
  // v_new^T (W v_old) rather than v_old^T (W v_new), so that basis vectors whose sigma is FP64
  // (the Ritz vectors after a defect-correction restart) couple with FP64 accuracy.
  std::vector<double> g(64 * 64), gt;
  if (mixed_ && old > 0 && newc > 0) {
    ts_gemm_tn(st, V_.get() + (int64_t)old * N_, N_, newc, W_.get(), N_, old, small_.get(), 64);
    gt.resize(64 * 64);
    PQG_HIP(hipMemcpyAsync(gt.data(), small_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  }
  PQG_HIP(hipMemcpyAsync(g.data(), G_.get(), 64 * 64 * 8, hipMemcpyDeviceToHost, st));
  PQG_HIP(hipStreamSynchronize(st));
  for (int c = old; c < m_; c++)
    for (int r = 0; r < m_; r++) {
      if (r < old) {
        const double v = gt.empty() ? g[r + c * 64] : gt[(c - old) + r * 64];
        g[c + r * 64] = g[r + c * 64] = v;
      } else if (r > c) {
        g[r + c * 64] = g[c + r * 64] = 0.5 * (g[c + r * 64] + g[r + c * 64]);
      }
    }
  PQG_HIP(hipMemcpyAsync(G_.get(), g.data(), 64 * 64 * 8, hipMemcpyHostToDevice, st));
  k_jacobi<<<1, 256, 0, st>>>(G_.get(), 64, m_, theta_.get(), Y_.get(), 64);
  PQG_HIP(hipGetLastError());
}

void Davidson::restart(hipStream_t st, const double* Vk, const double* Wk, int keep) {
  // V[:, :keep] = Vk, W[:, :keep] = Wk; projected matrix becomes diag(theta)
  PQG_HIP(hipMemcpyAsync(V_.get(), Vk, (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
  PQG_HIP(hipMemcpyAsync(W_.get(), Wk, (size_t)N_ * keep * 8, hipMemcpyDeviceToDevice, st));
  m_ = nW_ = 0;
  // recompute the keep x keep projection exactly
  ts_gemm_tn(st, V_.get(), N_, keep, W_.get(), N_, keep, G_.get(), 64);
  m_ = nW_ = keep;
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
