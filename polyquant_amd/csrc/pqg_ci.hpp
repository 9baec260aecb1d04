// CI handle: determinant set, integrals and per-string tables resident in HBM.
#pragma once

#include "pqg_internal.hpp"
#include "pqg_tensor.hpp"

namespace pqg {

constexpr int kSlots = 4;  // (electron alpha, electron beta, positron alpha, positron beta)

// Device view of one string set ("unique_dets[part][spin]" in internal order).
struct SlotDev {
  const uint64_t* bits;  // [nstr] ascending integer value
  const uint32_t* rtab;  // rank table [nchunk][256][nocc+1]
  const int32_t* r2i;    // colex rank -> index in the set, or nullptr when the set is complete
  int nstr, n_orb, nocc, nvirt, nchunk, npair, nsingle, ndouble;
};

// Host bookkeeping for one string set.
struct Slot {
  int species = 0, spin = 0, spinblk = 0;
  int n_orb = 1, nocc = 0, nvirt = 0, nstr = 1, nchunk = 1, npair = 1;
  bool complete = true;
  int64_t nsingle = 0, ndouble = 0;
  std::vector<uint64_t> ref;       // reference order (level-major)
  std::vector<uint64_t> sorted;    // internal order (ascending value)
  std::vector<int32_t> ref2int;    // reference index -> internal index
  std::vector<int32_t> int2ref;
  std::vector<int> deg;            // excitation degree vs the HF string, reference order
  DevBuf<uint64_t> bits;
  DevBuf<uint32_t> rtab;
  DevBuf<int32_t> r2i;
  DevBuf<int32_t> d_ref2int;
  // one-particle density matrix: groups of the strings that share nocc - 1 orbitals (built on
  // first use): bits of each shared core and the string index of each member (-1: not in set)
  int64_t ngroup = 0;
  int gm = 0;                      // members per group, n_orb - nocc + 1
  DevBuf<uint64_t> gbits;
  DevBuf<int32_t> gmem;
  DevBuf<double> A, D, E1;         // singles same-slot part, doubles values, diagonal part
  DevBuf<float> A32, D32;
  SlotDev dev() const {
    return SlotDev{bits.get(), rtab.get(), complete ? nullptr : r2i.get(), nstr, n_orb, nocc, nvirt,
                   nchunk, npair, (int)nsingle, (int)ndouble};
  }
};

template <typename T>
struct Tables {
  const T* A[kSlots] = {};
  const T* D[kSlots] = {};
  const T* X[kSlots][kSlots] = {};  // spectator sums [nstr_u][npair_t], charge factor included
  const T* V[kSlots][kSlots] = {};  // coupling blocks [npair_t][npair_u], charge factor included
};

// One-species full CI: link tables for the sigma terms computed as GEMMs on MFMA.
struct LinkTab {
  int L = 0, nv = 0, nb = 0;  // links per string, virtuals, off-diagonal orbital pairs
  int nin = 0;                // rows of the column-set map (electron sets)
  DevBuf<int2> links;         // [nstr][L]: {J | odd << 31, Q << 16 | rev}
  DevBuf<int32_t> tbm;        // [nin][nstr]: (column index << 1 | odd) or -1 (electron sets)
  DevBuf<int32_t> tbmT;       // BF16x3 kernel: off-diagonal pair rows of tbm as [nstr + 32][nb4] words
                              // (byte offset of the pair in the staged row | odd << 31)
};
struct LinkGemm {
  bool enabled = false;
  LinkTab tab[kSlots];
  DevBuf<char> U;             // GEMM output for one column chunk (FP32: two half blocks)
  DevBuf<uint32_t> P;         // BF16x3 kernel: the vector of a pass as bf16 (hi, lo) pairs
  DevBuf<float> V32s[kSlots]; // FP32 copies of the same-set coupling blocks
  hipStream_t side = nullptr; // the overlapped gathers (default priority), created on first use
  hipEvent_t ev[4] = {};      // GEMM done into block 0 / 1, gather done from block 0 / 1
  LinkGemm() = default;
  LinkGemm(const LinkGemm&) = delete;
  LinkGemm& operator=(const LinkGemm&) = delete;
  ~LinkGemm();
};

}  // namespace pqg

namespace pqg {
// Holds one context reference; declared first in pqg_ci so it is released after every
// device buffer of the handle has been returned to the context's arena.
struct CtxRef {
  pqg_ctx* c = nullptr;
  void set(pqg_ctx* x) {
    c = x;
    ctx_retain(x);
  }
  ~CtxRef() {
    if (c) ctx_release(c);
  }
};
}  // namespace pqg

struct pqg_ci {
  pqg::CtxRef ctx_ref;
  pqg_ctx* ctx = nullptr;
  hipStream_t stream = nullptr;  // all work of this handle is ordered on it
  int nsp = 1;
  int nslot = 2;               // 2 (one species) or 4
  pqg::Slot slot[pqg::kSlots];
  double charge[2] = {-1.0, 1.0};
  int nspin[2] = {1, 1};
  int64_t n_ref = 0;           // reference determinant count (N_dets)
  int64_t n_int = 0;           // internal product-space size
  bool fci = true;             // every product tuple is a determinant
  int64_t stride[pqg::kSlots] = {1, 1, 1, 1};  // internal address strides
  int64_t nref_str[pqg::kSlots] = {1, 1, 1, 1};
  pqg::DevBuf<int64_t> ref_addr;   // truncated spaces: reference address -> internal address
  pqg::DevBuf<uint8_t> valid;      // truncated spaces: internal address -> in space
  std::vector<int32_t> host_unfold;  // truncated spaces: reference address -> tuple (N x 4)
  pqg::DevBuf<double> hdiag;       // internal order; +1e30 outside the space
  // integrals (device, row-major blocks) and derived tables
  pqg::DevBuf<double> V64[pqg::kSlots][pqg::kSlots];
  pqg::DevBuf<float> V32[pqg::kSlots][pqg::kSlots];
  pqg::DevBuf<double> X64[pqg::kSlots][pqg::kSlots];
  pqg::DevBuf<float> X32[pqg::kSlots][pqg::kSlots];
  pqg::DevBuf<double> DW[pqg::kSlots][pqg::kSlots];
  bool mixed_ready = false;
  pqg::TensorPath tensor;          // single-particle strings: dense pair-GEMM sigma
  pqg::LinkGemm lg;                // one-species full CI: link GEMM sigma terms
  // scratch for sigma (transposed copies), grown on demand
  pqg::DevBuf<char> scratch;
  // boundary staging
  pqg::DevBuf<double> stage_in, stage_out;
  // stored explicit H (pqg_ci_set_explicit): both triangles, internal order, rows over every
  // internal address (empty outside the space); sigma in FP64 is then its product
  bool xh = false;
  pqg::DevBuf<int64_t> xh_ptr;
  pqg::DevBuf<int32_t> xh_col;
  pqg::DevBuf<double> xh_val;
  // spin penalties added to the stored H: per species P = S^2 - e (both triangles, internal
  // order); sigma += alpha P C (power 1) or alpha P (P C) (power 2)
  struct XPen {
    pqg::DevBuf<int64_t> ptr;
    pqg::DevBuf<int32_t> col;
    pqg::DevBuf<double> val;
    double alpha = 0.0;
    int power = 1;
  };
  std::vector<XPen> xpen;
  // diagonal of the stored operator (stored H plus the penalties): the Davidson's
  // preconditioner and start selection use it while the explicit H is stored
  pqg::DevBuf<double> xdiag;

  pqg_ci() = default;
  pqg_ci(const pqg_ci&) = delete;
  pqg_ci& operator=(const pqg_ci&) = delete;
  // Work queued on the handle's stream finishes before any of its buffers go back to the arena
  // (which may hand them to another stream, or return a chunk to the device), also when a
  // setup or a solve is abandoned on an error.
  ~pqg_ci() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (lg.side) (void)hipStreamSynchronize(lg.side);
  }
};

namespace pqg {

// Diagonal of the operator sigma applies: the stored explicit H with its penalties when
// one is stored, else diagonal_Hii (internal order, +1e30 outside the space).
inline const double* operator_diag(const pqg_ci* ci) { return ci->xh ? ci->xdiag.get() : ci->hdiag.get(); }
// Sigma on internal-order device vectors: S = H C for k vectors (leading dimension ld).
void sigma_internal(pqg_ci* ci, const double* C, double* S, int k, int64_t ld, hipStream_t st);
void sigma_internal(pqg_ci* ci, const float* C, float* S, int k, int64_t ld, hipStream_t st);
void ensure_mixed(pqg_ci* ci);
// Reference-order <-> internal-order vector permutation (device).
void to_internal(pqg_ci* ci, const double* ref, int64_t ldr, double* in, int64_t ldi, int k, hipStream_t st);
void to_reference(pqg_ci* ci, const double* in, int64_t ldi, double* ref, int64_t ldr, int k, hipStream_t st);
// zero entries outside a truncated space (no-op for full CI)
template <typename T>
void mask_invalid(pqg_ci* ci, T* v, int64_t ld, int k, hipStream_t st);
// moves string set t of k internal-order vectors to the outermost index (X[n_t][A][B])
void slot_outer(pqg_ci* ci, int t, const double* C, int64_t ldc, double* out, int k, hipStream_t st);
// bytes of device memory a CI of this description needs (admission control)
uint64_t estimate_ci_bytes(const pqg_ci_desc* d, const pqg_dav_params* p);
pqg_ci* create_ci(pqg_ctx* ctx, const pqg_ci_desc* desc, hipStream_t stream);
bool tensor_eligible(const pqg_ci* ci);
void tensor_setup(pqg_ci* ci, const std::vector<const double*>& h1);
void tensor_sigma(pqg_ci* ci, const double* C, double* S, int k, int64_t ld, hipStream_t st);
void tensor_sigma(pqg_ci* ci, const float* C, float* S, int k, int64_t ld, hipStream_t st);
bool link_gemm_eligible(const pqg_ci* ci);
void link_gemm_setup(pqg_ci* ci);
size_t lg_budget_bytes(int64_t N, size_t w);
// pass 0: everything but the diagonal and the e-beta singles/doubles (C with positron strings
// permuted in front of e-beta); pass 1: e-beta singles and doubles on the transposed vector.
// accumulate: add to S (one vector) or overwrite it.
void link_gemm_pass(pqg_ci* ci, int pass, const double* C, double* S, bool accumulate, hipStream_t st);
void link_gemm_pass(pqg_ci* ci, int pass, const float* C, float* S, bool accumulate, hipStream_t st);
// S += diag(H) C for k vectors
void link_gemm_diag(pqg_ci* ci, const double* C, int64_t ldc, double* S, int64_t lds, int k, hipStream_t st);
void link_gemm_diag(pqg_ci* ci, const float* C, int64_t ldc, float* S, int64_t lds, int k, hipStream_t st);

}  // namespace pqg
