/* pqgpu.h — C ABI of the MI355X configuration-interaction hot path.
 *
 * Drop-in boundary for Polyquant's CI solve (shivupa/polyquant v0.7.0).  Every entry point
 * takes plain pointers and sizes; no C++ or torch types cross it.  Matrices are column-major
 * like Eigen's defaults, determinant addresses and string orders are the reference's.
 *
 * Errors are status codes (never exit()); pqg_last_error() gives the message of the last
 * failing call on the calling thread.  A pqg_ci handle is used from one host thread at a
 * time; a pqg_ctx may be shared.
 */
#ifndef PQGPU_H
#define PQGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int pqg_status;
#define PQG_OK 0
#define PQG_ERR_ARG 1          /* invalid argument / unsupported shape               */
#define PQG_ERR_HIP 2          /* HIP runtime error                                  */
#define PQG_ERR_OOM 3          /* request does not fit the context's HBM budget      */
#define PQG_ERR_NOCONV 4       /* Davidson did not converge (epci.cpp:620 aborts)    */
#define PQG_ERR_UNSUPPORTED 5

#define PQG_FP64 0             /* all arithmetic in FP64                              */
#define PQG_MIXED 1            /* FP32 sigma bulk, FP64 correction / subspace / checks */

#define PQG_AO2MO_FP64 0       /* FP64 MFMA                                           */
#define PQG_AO2MO_BF16X3 1     /* BF16x3 split products on MFMA, FP32 accumulation    */

typedef struct pqg_ctx pqg_ctx;
typedef struct pqg_ci pqg_ci;

/* CI problem description.  Mirrors the data POLYQUANT_DETSET reads
 * (src/ci/determinant_set.hpp:84-119) and POLYQUANT_INTEGRAL's MO members
 * (src/integral/integral.hpp:189-205):
 *   h[p][s]          n x n one-body MO integrals (mo_one_body_ints[p][s]), column-major;
 *   eri[a][sa][b][sb] n_pair_a x n_pair_b two-body block (mo_two_body_ints[a][sa][b][sb]),
 *                    column-major, element (idx2(i,j), idx2(k,l)); required for b >= a.
 * Species index order is the reference's (electron = 0, positron = 1); spin block
 * index for beta is (1 % nspin_blocks) as in slater_condon_singleparticle.cpp:13.
 * ex_level is excitation_level (epci.hpp:107) = {alpha, beta, total}; a full CI passes
 * the largest levels.  Strings are one 64-bit word: n_orb <= 63.                       */
typedef struct {
  int nspecies;
  int n_orb[2];
  int n_alpha[2];
  int n_beta[2];
  double charge[2];
  int nspin_blocks[2];
  int ex_level[2][3];
  int max_collective;              /* max_collective_excitation_level (epci.hpp:108)    */
  const double* h[2][2];
  const double* eri[2][2][2][2];
} pqg_ci_desc;

/* Davidson parameters: the knobs POLYQUANT_EPCI::run passes to the eigensolver
 * (epci.cpp:589-608): num_states, num_subspace_vec, max(2*nsv, 10*nstates), iteration_max,
 * convergence_E.  precision selects PQG_FP64 or PQG_MIXED.                              */
typedef struct {
  int nev;                         /* 1..64                                              */
  int nvec_init;
  int nvec_max;                    /* <= 0: max(2*nvec_init, 10*nev); clamped to 64      */
  int maxit;
  double tol;                      /* residual-norm tolerance                            */
  int precision;
  /* 0: start from unit vectors at the smallest diagonal entries (the reference's start).
   * > 0 (no counterpart in the reference): add a deterministic pseudo-random admixture of
   * norm ~guess_perturb to each start vector, so the start overlaps every eigenvector (a
   * unit-vector start can converge to an excited root when the ground state has no weight
   * on the lowest-diagonal determinants).  It also stops expanding roots whose residual is
   * already below tol.                                                                     */
  double guess_perturb;
} pqg_dav_params;

typedef struct {
  int iterations;
  int converged;
  int sigma_calls;                 /* sigma launches (vectors), all precisions           */
  int sigma_calls_fp64;
  int restarts;
  double residual[64];             /* final FP64 residual norms per root                 */
  double seconds;                  /* wall time of the solve                             */
} pqg_dav_report;

/* --- context / memory manager --------------------------------------------------------- */
/* One HIP stream and a device-memory arena limited to hbm_budget bytes.                  */
pqg_status pqg_ctx_create(int device, uint64_t hbm_budget, pqg_ctx** out);
pqg_status pqg_ctx_destroy(pqg_ctx* ctx);
pqg_status pqg_ctx_memory(pqg_ctx* ctx, uint64_t* used, uint64_t* budget, uint64_t* peak);
pqg_status pqg_ctx_sync(pqg_ctx* ctx);
const char* pqg_last_error(void);
/* Kernel timing with HIP events recorded on the stream each kernel is launched on (the role of
 * POLYQUANT_TIMER, src/io/timer.hpp:15-154, for device work).  enable = 1 starts recording,
 * 0 stops.  pqg_ctx_profile_read sums the launches recorded under `name` since the last read
 * of that name: launch count, total milliseconds and total work units (vectors for sigma). */
pqg_status pqg_ctx_profile(pqg_ctx* ctx, int enable);
pqg_status pqg_ctx_profile_read(pqg_ctx* ctx, const char* name, int64_t* launches, double* total_ms,
                                int64_t* total_units);

/* --- determinant set (replaces POLYQUANT_DETSET<uint64_t>) ----------------------------- */
/* Builds strings in the reference order (determinant_creation.cpp:6-207, epci.cpp:334-388),
 * uploads integrals, and precomputes the diagonal (slater_condon.cpp:102-153) on the GPU. */
pqg_status pqg_ci_create(pqg_ctx* ctx, const pqg_ci_desc* desc, pqg_ci** out);
pqg_status pqg_ci_destroy(pqg_ci* ci);
/* rows()/cols() (determinant_set.hpp:127-134) */
int64_t pqg_ci_ndets(const pqg_ci* ci);
int64_t pqg_ci_nstrings(const pqg_ci* ci, int part, int spin);
/* Which sigma kernel family the handle runs (no counterpart in the reference, whose sigma has one
 * form): chosen at pqg_ci_create from the shape of the space and the limits of the fast paths
 * (the link GEMM needs full CI with a column set of at most 16 orbitals whose row fits LDS and at
 * most 64 single excitations per row tuple; the pair GEMM one particle per string), or the stored
 * explicit H after pqg_ci_set_explicit.  Returns a PQG_SIGMA_* value, -1 for a null handle.  */
#define PQG_SIGMA_ENUMERATION 0 /* general pull-form enumeration kernel                   */
#define PQG_SIGMA_LINK_GEMM 1   /* full CI, >= 2 particles per string set: link GEMM      */
#define PQG_SIGMA_PAIR_GEMM 2   /* one particle per string: pair GEMM                     */
#define PQG_SIGMA_STORED_H 3    /* stored explicit H (FP64)                               */
int pqg_ci_sigma_path(const pqg_ci* ci);
/* diagonal_Hii (determinant_set.hpp:259), reference address order */
pqg_status pqg_ci_diag(pqg_ci* ci, double* hdiag);
/* operator()(i, j) -> Slater_Condon (determinant_set.hpp:139, slater_condon.cpp:4-100) for
 * npairs (i, j) address pairs at once.                                                    */
pqg_status pqg_ci_elements(pqg_ci* ci, const int64_t* i, const int64_t* j, int64_t npairs, double* out);
/* create_ham + the mirrored dense copy of the exact_diag branch (explicit_ham.cpp:51-80,
 * epci.cpp:633-643): the full N x N Hamiltonian in reference address order, column-major with
 * leading dimension ldh, element (i, j) = Slater_Condon(min(i, j), max(i, j)) (diagonal
 * diagonal_Hii), so H is exactly symmetric.  on_device != 0: H is a device pointer.          */
pqg_status pqg_ci_dense_h(pqg_ci* ci, double* H, int64_t ldh, int on_device);
/* create_ham (explicit_ham.cpp:51-84 -> single_species_create_ham_singleshot,
 * explicit_ham_singleparticle.cpp:170-537, two_species_create_ham_singleshot,
 * explicit_ham_mixedparticle.cpp:186-550) as the reference's compressed row-major upper
 * triangle with int indices: row r holds the diagonal and every J > r with a nonzero
 * Slater_Condon element, columns ascending; screening > 0 drops every |v| <= screening
 * (ham.prune(screening, 1.0), explicit_ham.cpp:78-80).  Host arrays.  Call with cols = vals =
 * NULL for *nnz (and rowptr[N + 1] when rowptr != NULL), then with cap >= *nnz to fill.      */
pqg_status pqg_ci_csr(pqg_ci* ci, double screening, int32_t* rowptr, int32_t* cols, double* vals, int64_t cap,
                      int64_t* nnz);
/* build_matrix = true (determinant_set.hpp:196): enable != 0 stores the explicit H of
 * pqg_ci_csr (same entries, same screening) on the device, both triangles in internal order, and
 * from then on pqg_ci_sigma and pqg_ci_davidson with PQG_FP64 multiply by it, as the reference's
 * SparseSymMatProd does (epci.cpp:664-682); PQG_MIXED is refused while it is stored.
 * enable = 0 frees it and returns to the matrix-free sigma.  nnz (may be NULL): stored entries.  */
pqg_status pqg_ci_set_explicit(pqg_ci* ci, int enable, double screening, int64_t* nnz);
/* create_S_sq_penalty (s2.cpp:126-153) on the stored explicit H (pqg_ci_set_explicit first):
 * sigma then adds alpha (S^2 - expected) C of species part (power 1, first order) or
 * alpha (S^2 - expected)^2 C (power 2, second order, applied as two products).  Penalties
 * accumulate over calls; pqg_ci_set_explicit clears them.                                     */
pqg_status pqg_ci_add_spin_penalty(pqg_ci* ci, int part, double expected, int power, double alpha);
/* create_S_sq_minus_expected_S_sq_matrix_singleshot (s2.cpp:156-255) as a dense matrix, and
 * its square for the second-order penalty (s2.cpp:141-147): P = S^2 - expected * I of species
 * part (power 1) or P^2 (power 2, summed over the determinants of the space), N x N reference
 * order, column-major, leading dimension lds, exactly symmetric.  create_S_sq_penalty
 * (s2.cpp:126-153) adds spin_penalty * this to H.  on_device != 0: S is a device pointer.     */
pqg_status pqg_ci_dense_s2(pqg_ci* ci, int part, double expected, int power, double* S, int64_t lds,
                           int on_device);
/* operator*(C) (direct.cpp:5-19): S = H C for k vectors, N x k column-major.
 * on_device != 0: C and S are device pointers.  precision: PQG_FP64 or PQG_MIXED (FP32).   */
pqg_status pqg_ci_sigma(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S, int64_t lds,
                        int on_device, int precision);
/* Allocation-free variant of pqg_ci_sigma (no counterpart in the reference; for callers that
 * cannot allocate device memory: tight serving loops, harnesses that forbid it).  C and S are
 * device pointers (reference order, N x k column-major); the call makes no hipMalloc/hipFree:
 * every scratch buffer (the internal-order copies of C and S, transposes, the link-GEMM output
 * chunks) comes from one fixed device array of PQG_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends.  Sizes whose copies of C and S alone exceed it (2 x 8 N k bytes, plus 8 N k with
 * PQG_MIXED) are refused with PQG_ERR_ARG; PQG_MIXED also needs the FP32 tables that the
 * handle's first mixed-precision pqg_ci_sigma / pqg_ci_davidson call builds (PQG_ERR_UNSUPPORTED
 * before that), and a stored explicit H works as in pqg_ci_sigma.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (other threads, other handles) race.  The
 * call synchronises its handle's stream before it returns.                                   */
#define PQG_NOALLOC_SCRATCH_BYTES ((uint64_t)256 << 20)
pqg_status pqg_ci_sigma_noalloc(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S, int64_t lds,
                                int precision);
/* Full Davidson solve on the GPU (Spectra DavidsonSymEigsSolver, SmallestAlge,
 * epci.cpp:589-608).  E: nev eigenvalues (without constant shifts); X: N x nev Ritz vectors
 * (host, reference order, column-major, may be NULL).                                      */
pqg_status pqg_ci_davidson(pqg_ci* ci, const pqg_dav_params* params, double* E, double* X, int64_t ldx,
                           pqg_dav_report* report);

/* The same solve one iteration at a time (no counterpart in the reference, whose Spectra solver
 * runs its iterations inside compute(), epci.cpp:606-608): for callers that interleave solves
 * with other work or time single iterations.  pqg_dav_create builds the start vectors;
 * pqg_dav_step runs one iteration (the sigma of the pending block, Rayleigh-Ritz, residuals and
 * the next block; one host synchronisation) and sets *done once the solve has converged or
 * stopped; pqg_dav_result then gives what pqg_ci_davidson gives (E, X and report may be NULL,
 * PQG_ERR_NOCONV after filling them when the solve stopped unconverged).  The handle belongs to
 * its pqg_ci and must be destroyed before it; it is used from one host thread at a time.    */
typedef struct pqg_dav pqg_dav;
pqg_status pqg_dav_create(pqg_ci* ci, const pqg_dav_params* params, pqg_dav** out);
pqg_status pqg_dav_step(pqg_dav* dav, int* done);
pqg_status pqg_dav_result(pqg_dav* dav, double* E, double* X, int64_t ldx, pqg_dav_report* report);
pqg_status pqg_dav_destroy(pqg_dav* dav);

/* create_1rdm (src/ci/detset/rdm1.cpp:6-79) for k states at once: the one-particle reduced
 * density matrix of the (part, spin) string set for each column of C (N x k, reference order,
 * column-major).  D receives k consecutive n_orb x n_orb column-major blocks, element (k, l) =
 * <C| a+_l a_k |C> of that set (symmetric).  on_device != 0: C and D are device pointers.   */
pqg_status pqg_ci_rdm1(pqg_ci* ci, const double* C, int64_t ldc, int k, int part, int spin, double* D,
                       int on_device);

/* evaluate_s2 (src/ci/detset/s2.cpp:5-130): <C|S^2|C> of each species for each column of C
 * (N x k, reference order).  S2 receives k x nspecies values, column-major
 * (S2[state + species * k], the reference's S_squared(state, species)), always host memory.
 * on_device != 0: C is a device pointer.                                                   */
pqg_status pqg_ci_s2(pqg_ci* ci, const double* C, int64_t ldc, int k, double* S2, int on_device);

/* --- parity dumps ------------------------------------------------------------------------ */
/* unique_dets[part][spin] strings in the reference order (determinant_creation.cpp:139-146) */
pqg_status pqg_ci_dump_strings(pqg_ci* ci, int part, int spin, uint64_t* out);
/* reference string indices of each determinant address (det_idx_unfold, :378-388), N x 4 */
pqg_status pqg_ci_dump_unfold(pqg_ci* ci, int32_t* out);
/* Excitations of every string of (part, spin), enumerated on the GPU.  which = 1 (singles)
 * or 2 (doubles).  For string r (reference order) the entries [off[r], off[r+1]) hold the
 * excited string's reference index J, ascending (the order of unique_singles/doubles,
 * determinant_creation.cpp:209-351), and its phase (get_phase, determinant_operations.cpp:47-116)
 * as +1/-1.  Call with J == NULL to get off[] only (size nstrings+1).                      */
pqg_status pqg_ci_dump_excitations(pqg_ci* ci, int which, int part, int spin, int64_t* off, int32_t* J,
                                   int8_t* phase);

/* --- AO -> MO transform (replaces transform_mo_2_body_integrals, integral.cpp:207-477) -- */
/* eri8: AO integrals with 8-fold symmetry, packed as idx2(idx2(p,q), idx2(r,s)) (idx8,
 * integral.hpp:76-88), length npair*(npair+1)/2 with npair = nao*(nao+1)/2.
 * C: nao x nmo column-major, 1 <= nmo <= nao (the reference's num_mo = C.cols() - frozen_core -
 * deleted_virtual, integral.cpp:214-217: pass the active columns, C + frozen_core * nao).
 * out4: npair_mo x npair_mo column-major, element (idx2(i,j), idx2(k,l)).
 * on_device != 0: eri8, C and out4 are device pointers.
 * mode: PQG_AO2MO_FP64 or PQG_AO2MO_BF16X3.  seconds (may be NULL): GPU time.              */
pqg_status pqg_ao2mo(pqg_ctx* ctx, int nao, int nmo, const double* eri8, const double* C, double* out4,
                     int on_device, int mode, double* seconds);

/* transform_mo_2_body_integrals with two coefficient matrices (integral.cpp:207-477, mo_coeffs_a
 * and mo_coeffs_b of one basis, e.g. the alpha-beta block of an unrestricted species):
 * out4(idx2(i,j), idx2(k,l)) = (ij|kl) with i, j over the columns of Ca and k, l over those of
 * Cb, npair x npair column-major.  Arguments otherwise as pqg_ao2mo.                        */
pqg_status pqg_ao2mo_pair(pqg_ctx* ctx, int nao, int nmo, const double* eri8, const double* Ca, const double* Cb,
                          double* out4, int on_device, int mode, double* seconds);
/* transform_mo_2_body_integrals for two species with their own bases (integral.cpp:207-477 with
 * quantum_part_a_idx != quantum_part_b_idx: num_ao_a / num_ao_b, :214-217; the AO integrals
 * (pq|rs) with p, q in basis a and r, s in basis b that libint computes at :259-279), as
 * calculate_mo_2_body_integrals calls it for the electron-positron block (:479-516).
 * eri_ab: npair(nao_a) x npair(nao_b) column-major, element (idx2(p,q), idx2(r,s)) (no charge
 * factor: the CI applies the charge product).  Ca: nao_a x nmo_a, Cb: nao_b x nmo_b, column-major
 * (1 <= nmo <= nao).  out4: npair(nmo_a) x npair(nmo_b) column-major, element (idx2(i,j), idx2(k,l)),
 * the reference's mo_two_body_ints[a][sa][b][sb].  on_device, mode, seconds as pqg_ao2mo.        */
pqg_status pqg_ao2mo_mixed(pqg_ctx* ctx, int nao_a, int nmo_a, int nao_b, int nmo_b, const double* eri_ab,
                           const double* Ca, const double* Cb, double* out4, int on_device, int mode,
                           double* seconds);
/* compute_frozen_core_ints (integral.cpp:565-698) for one (species a, spin a) operator and one
 * (species b, spin b) core density, as called by calculate_frozen_core_ints
 * (integral.cpp:166-205): out (nao_a x nao_a, column-major, host) +=
 *   cj * J(Dc) - [same species and a_spin == b_spin] * K(Dx)
 * with J(D)_ij = sum_kl (ij|kl) D_kl, K(D)_ij = sum_kl (ik|jl) D_kl, Dc = 2 D (restricted b) or
 * D_alpha + D_beta, Dx = D (restricted b) or D_{b_spin}, cj = 1 (same species, restricted b),
 * 1/2 (same species, unrestricted b: one call per spin block) or charge_prod (two species).
 * dm_b: the core densities of species b, nao_b x nao_b blocks: exactly 1 is read when
 * restricted_b != 0, exactly 2 consecutive blocks (alpha, beta) when restricted_b == 0, so an
 * unrestricted species with one spin block passes a zero beta block.  eri: same species, the 8-fold packed AO integrals (as pqg_ao2mo); two species,
 * the npair_a x npair_b column-major block (ij|kl), idx2 packing.                            */
pqg_status pqg_fc_ints(pqg_ctx* ctx, int nao_a, int nao_b, const double* eri, int same_particle, int a_spin,
                       int b_spin, int restricted_b, double charge_prod, const double* dm_b, double* out);

/* --- serving (continuous batching of CI solves) ------------------------------------------ */
typedef struct pqg_server pqg_server;
typedef struct {
  int64_t request_id;
  int status;                      /* PQG_OK, PQG_ERR_OOM, PQG_ERR_NOCONV, ...           */
  int nev;
  double energies[64];
  double t_submit, t_admit, t_done; /* seconds on the server clock                       */
  int iterations;
  char message[128];               /* error text of a failed request ("" on success)      */
} pqg_result;

/* max_slots: solves resident on the device at once; budget from the context.            */
pqg_status pqg_serve_create(pqg_ctx* ctx, int max_slots, pqg_server** out);
pqg_status pqg_serve_destroy(pqg_server* srv);
/* Queue one request (copies desc and params).  Returns its id.                           */
pqg_status pqg_serve_submit(pqg_server* srv, const pqg_ci_desc* desc, const pqg_dav_params* params,
                            int64_t* request_id);
/* Collect finished requests.  Admission and the solves themselves run inside the server:
 * each resident slot has a worker thread and its own HIP stream, so the solves of different
 * slots overlap on the device; a freed slot admits the next queued request whose memory fits.
 * Returns the number of results written (<= cap).                                        */
pqg_status pqg_serve_step(pqg_server* srv, pqg_result* results, int cap, int* nresults);
pqg_status pqg_serve_stats(pqg_server* srv, int* queued, int* resident, uint64_t* bytes_resident);
double pqg_serve_clock(pqg_server* srv);

#ifdef __cplusplus
}
#endif
#endif /* PQGPU_H */
