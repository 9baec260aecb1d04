// Davidson eigensolver resident on the GPU, as a step-wise state machine so that a scheduler
// can interleave many solves and fuse their sigma builds.
#pragma once

#include <chrono>

#include "pqg_ci.hpp"

namespace pqg {

// Page-locked host buffer (freed with the owner).
struct PinnedHost {
  double* p = nullptr;
  PinnedHost() = default;
  PinnedHost(const PinnedHost&) = delete;
  PinnedHost& operator=(const PinnedHost&) = delete;
  ~PinnedHost() {
    if (!p) return;
    std::lock_guard<std::recursive_mutex> dl(device_op_mutex());
    (void)hipHostFree(p);
  }
};

// A block of vectors whose sigma the solver needs next.
struct SigmaWork {
  int k = 0;               // number of vectors (0: nothing pending)
  int precision = PQG_FP64;
  const double* C = nullptr;  // internal order, leading dimension ld
  double* S = nullptr;
  int64_t ld = 0;
};

// Captured sigma launches of one solve, keyed by the block they apply to.
struct SigmaGraphs {
  struct Entry {
    const double* C;
    double* S;
    int k, precision;
    int64_t ld;
    uint64_t bufs;          // the sigma buffers the capture (or the first run) used
    hipGraphExec_t exec;    // nullptr until captured
  };
  bool on = false;
  std::vector<Entry> entries;
  SigmaGraphs() = default;
  SigmaGraphs(const SigmaGraphs&) = delete;
  SigmaGraphs& operator=(const SigmaGraphs&) = delete;
  ~SigmaGraphs();
};

class Davidson {
 public:
  Davidson(pqg_ci* ci, const pqg_dav_params& p);
  Davidson(const Davidson&) = delete;
  Davidson& operator=(const Davidson&) = delete;
  // the solve's queued work finishes before its buffers go back to the arena
  ~Davidson() { (void)hipStreamSynchronize(ci_->stream); }
  // Work the solver needs before post_sigma() can run (k == 0 when finished).
  SigmaWork pending() const { return work_; }
  // Runs the rest of one iteration after the pending sigma has been written.
  void post_sigma(hipStream_t st);
  // post_sigma in two halves: begin queues the device work of the iteration (no wait), end
  // synchronises the stream and takes the host decisions
  void post_sigma_begin(hipStream_t st);
  void post_sigma_end(hipStream_t st);
  // replay the sigma of a block seen before from a captured hipGraph (small spaces)
  void use_graphs(bool on) { graphs_.on = on; }
  // Run the pending sigma itself (FP32 conversion included for mixed precision).
  void run_sigma(hipStream_t st);
  bool done() const { return done_; }
  pqg_status status() const { return status_; }
  int nev() const { return nev_; }
  const std::vector<double>& energies() const { return theta_h_; }
  const std::vector<double>& residuals() const { return rn_h_; }
  const double* ritz_vectors() const { return X_.get(); }  // internal order, N x nev, leading dimension ld()
  int64_t ld() const { return ld_; }
  pqg_dav_report report() const;
  uint64_t device_bytes() const;

 private:
  void setup();
  void initial_guess(hipStream_t st);
  hipGraphExec_t capture_sigma(hipStream_t st);
  int orthonormalize_append(hipStream_t st, double* T, int nt);
  // rg: residual Gram of the roots the columns correct; columns of converged roots are dropped
  void append_async(hipStream_t st, double* T, int nt, const double* g0, const double* rg = nullptr);
  int finish_append();
  void read_status_async(hipStream_t st, const double* rgram);
  void sigma_body(hipStream_t st);
  uint64_t sigma_buffers() const;
  void restart_to_ritz(hipStream_t st, int keep);
  void rayleigh_ritz(hipStream_t st);
  void ritz_residual(hipStream_t st);
  // gate: a device flag; when given, the product runs only if the flag is nonzero
  void ts_gemm_tn(hipStream_t st, const double* A, int64_t lda, int na, const double* B, int64_t ldb, int nb,
                  double* out, int ldo, const int* gate = nullptr);
  void ts_gemm_nn(hipStream_t st, const double* A, int64_t lda, int ka, const double* Bs, int ldb, int nb,
                  double* out, int64_t ldo, double alpha, double beta, const int* gate = nullptr);
  void ts_gemm_nn2(hipStream_t st, const double* A0, const double* A1, int64_t lda, int ka, const double* Bs, int ldb,
                   int nb, double* out0, double* out1, int64_t ldo);
  double* Rp() const { return RT_.get(); }
  double* Tp() const { return RT_.get() + (size_t)ld_ * nev_; }

  pqg_ci* ci_;
  pqg_dav_params p_;
  int64_t N_;
  int64_t ld_;  // leading dimension of every N-vector block (N_ rounded up to 32)
  int nev_, ninit_, nmax_;
  bool mixed_;
  DevBuf<double> V_, W_, G_, Y_, Yprev_, Ynext_, theta_, X_, AX_, RT_, tmp_, part_, small_, W64_;
  DevBuf<double> gram0_, rgram_, mbuf_, status_dev_;
  DevBuf<double> Yk_, Qr_;  // previous Ritz coefficients and the restart coefficients (GD+k)
  bool gdk_ = true;         // thick restart keeps the previous Ritz vectors (GD+k)
  int yk_rows_ = 0;         // rows of Yk_ (0: none saved for this basis)
  DevBuf<int> nk_dev_;
  DevBuf<int> reorth_dev_;       // second Gram-Schmidt pass needed (1) or not (0)
  DevBuf<double> small2_;        // coefficients of the second Gram-Schmidt pass
  PinnedHost status_pin_;        // pinned host copy of status_dev_
  double* status_h_ = nullptr;
  int pending_nt_ = 0;
  DevBuf<float> c32_, s32_;
  int m_ = 0, nW_ = 0, iter_ = 0, restarts_ = 0, since_check_ = 0, warm_ = 0;
  int sigma_calls_ = 0, sigma_calls64_ = 0;
  bool done_ = false, checking_ = false, mixed_check_ = false;
  bool began_ = false, need_restart_ = false, speculate_ = false;  // state between the halves
  SigmaGraphs graphs_;
  double worst_hist_[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // FP32-based worst residuals since the last check
  pqg_status status_ = PQG_OK;
  double last_true_ = 1e300;
  double last_worst_ = 1e300, theta_abs_max_ = 0.0;  // from the last residual check
  bool have_worst_ = false;
  std::vector<double> theta_h_, rn_h_;
  SigmaWork work_;
  std::chrono::steady_clock::time_point t0_;
  double seconds_ = 0.0;
  int nchunk_ = 1;
};

}  // namespace pqg
