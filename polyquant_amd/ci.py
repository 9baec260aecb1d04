"""Host-side mirror of Polyquant's CI interfaces over the HIP library.

``DetSet`` stands in for ``POLYQUANT_DETSET<uint64_t>`` (src/ci/determinant_set.hpp:27-201):
``rows()``, ``cols()``, ``__call__(i, j)`` (Slater-Condon element), ``@``/``__mul__``
(``operator*``, the sigma build, src/ci/detset/direct.cpp:5-19) and ``diagonal_Hii``.
``EPCI`` restates the solve flow of ``POLYQUANT_EPCI::run`` (src/ci/epci.cpp:575-621):
Davidson with ``num_states``, ``num_subspace_vec``, ``max(2*nsv, 10*nstates)``,
``iteration_max``, ``convergence_E`` and the constant shift added to the energies; a failed
solve raises like the reference's ``APP_ABORT("CI Calculation did not converge!")``.

Addresses, string orders and matrices are the reference's (column-major N x k).
"""
from __future__ import annotations

import atexit
import ctypes as ct
import weakref
from typing import Optional

import numpy as np

from . import _lib
from ._lib import CiDesc, DavParams, DavReport, check, load

_dp = ct.POINTER(ct.c_double)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


# Live handles, closed (CI handles first) before the interpreter tears down the HIP runtime.
_live_ci: "weakref.WeakSet" = weakref.WeakSet()
_live_ctx: "weakref.WeakSet" = weakref.WeakSet()


@atexit.register
def _close_all():
    for obj in list(_live_ci):
        obj.close()
    for obj in list(_live_ctx):
        obj.close()


class Context:
    """One GPU, one HIP stream and a device-memory arena bounded by ``hbm_budget`` bytes."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = ct.c_void_p()
        check(load().pqg_ctx_create(device, hbm_budget, ct.byref(self.h)))
        _live_ctx.add(self)

    def memory(self):
        used, budget, peak = ct.c_uint64(), ct.c_uint64(), ct.c_uint64()
        check(load().pqg_ctx_memory(self.h, ct.byref(used), ct.byref(budget), ct.byref(peak)))
        return used.value, budget.value, peak.value

    def profile(self, enable: bool = True):
        check(load().pqg_ctx_profile(self.h, 1 if enable else 0))

    def profile_read(self, name: str):
        """(launches, total_ms, total_units) recorded for kernel `name` since the last read."""
        n, ms, u = ct.c_int64(), ct.c_double(), ct.c_int64()
        check(load().pqg_ctx_profile_read(self.h, name.encode(), ct.byref(n), ct.byref(ms), ct.byref(u)))
        return n.value, ms.value, u.value

    def close(self):
        if self.h:
            load().pqg_ctx_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context()
    return _default_ctx


def make_desc(prob, keep: list) -> CiDesc:
    """pqg_ci_desc from a workloads.CIProblem (arrays kept alive in ``keep``)."""
    d = CiDesc()
    d.nspecies = prob.nspecies
    d.max_collective = int(min(prob.max_collective, 2**31 - 1))
    for p, sp in enumerate(prob.species):
        d.n_orb[p] = sp.n_orb
        d.n_alpha[p] = sp.n_alpha
        d.n_beta[p] = sp.n_beta
        d.charge[p] = sp.charge
        d.nspin_blocks[p] = sp.nspin_blocks
        lv = sp.levels()
        for q in range(3):
            d.ex_level[p][q] = int(lv[q])
        for s in range(sp.nspin_blocks):
            h = np.asfortranarray(prob.h[p][s], dtype=np.float64)
            keep.append(h)
            d.h[p][s] = h.ctypes.data_as(_dp)
    for (a, sa, b, sb), v in prob.eri.items():
        arr = np.asfortranarray(v, dtype=np.float64)
        keep.append(arr)
        d.eri[a][sa][b][sb] = arr.ctypes.data_as(_dp)
    return d


class DetSet:
    """GPU determinant set: the drop-in for POLYQUANT_DETSET<uint64_t>."""

    def __init__(self, prob, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.prob = prob
        self._keep = []
        self.desc = make_desc(prob, self._keep)
        self.h = ct.c_void_p()
        check(load().pqg_ci_create(self.ctx.h, ct.byref(self.desc), ct.byref(self.h)))
        _live_ci.add(self)
        self._keep = []  # the library copied everything it needs
        self.N_dets = load().pqg_ci_ndets(self.h)
        self._diag = None

    def close(self):
        if self.h:
            load().pqg_ci_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- POLYQUANT_DETSET interface ---------------------------------------------------
    def rows(self) -> int:
        return self.N_dets

    def cols(self) -> int:
        return self.N_dets

    @property
    def diagonal_Hii(self) -> np.ndarray:
        if self._diag is None:
            out = np.zeros(self.N_dets)
            check(load().pqg_ci_diag(self.h, _ptr(out)))
            self._diag = out
        return self._diag

    def elements(self, i, j) -> np.ndarray:
        i = np.ascontiguousarray(np.atleast_1d(i), dtype=np.int64)
        j = np.ascontiguousarray(np.atleast_1d(j), dtype=np.int64)
        out = np.zeros(len(i))
        check(load().pqg_ci_elements(self.h, _ptr(i), _ptr(j), len(i), _ptr(out)))
        return out

    def create_ham_dense(self) -> np.ndarray:
        """create_ham plus the mirrored dense copy that exact_diag diagonalises
        (src/ci/detset/explicit_ham.cpp:51-80, src/ci/epci.cpp:633-643): N x N, reference
        address order, built on the GPU."""
        n = self.N_dets
        H = np.zeros((n, n), order="F")
        check(load().pqg_ci_dense_h(self.h, _ptr(H), n, 0))
        return H

    def create_ham_csr(self, screening: float = 0.0):
        """create_ham (src/ci/detset/explicit_ham.cpp:51-84): the upper triangle of H as
        compressed rows (rowptr int32 [N + 1], cols int32, vals), columns ascending, built on the
        GPU; screening > 0 drops |v| <= screening as ham.prune(screening, 1.0)."""
        n = self.N_dets
        rowptr = np.zeros(n + 1, dtype=np.int32)
        nnz = ct.c_int64()
        check(load().pqg_ci_csr(self.h, float(screening), _ptr(rowptr), None, None, 0, ct.byref(nnz)))
        cols = np.zeros(nnz.value, dtype=np.int32)
        vals = np.zeros(nnz.value)
        check(load().pqg_ci_csr(self.h, float(screening), None, _ptr(cols), _ptr(vals), nnz.value, ct.byref(nnz)))
        return rowptr, cols, vals

    def set_explicit(self, enable: bool = True, screening: float = 0.0) -> int:
        """build_matrix = true: store the explicit H (create_ham's entries and screening) on the
        GPU, so sigma and the Davidson solve (FP64) multiply by it; False returns to the
        matrix-free sigma.  Returns the number of stored entries (both triangles)."""
        nnz = ct.c_int64()
        check(load().pqg_ci_set_explicit(self.h, int(enable), float(screening), ct.byref(nnz)))
        return nnz.value

    def add_spin_penalty(self, part: int, expected: float, power: int, alpha: float):
        """Adds alpha (S^2 - expected)^power of species part to the stored explicit H
        (set_explicit first); sigma and the Davidson solve then include it."""
        check(load().pqg_ci_add_spin_penalty(self.h, part, float(expected), power, float(alpha)))

    def create_s2_penalty_dense(self, part: int, expected: float = 0.0, power: int = 1) -> np.ndarray:
        """create_S_sq_minus_expected_S_sq_matrix_singleshot (src/ci/detset/s2.cpp:156-255) as a
        dense N x N matrix P = S^2 - expected I of species part (power 1), or P @ P (power 2, the
        second-order penalty of s2.cpp:141-147), built on the GPU."""
        n = self.N_dets
        S = np.zeros((n, n), order="F")
        check(load().pqg_ci_dense_s2(self.h, part, float(expected), power, _ptr(S), n, 0))
        return S

    def __call__(self, i: int, j: int) -> float:
        """operator()(i, j) -> Slater_Condon(i, j)."""
        return float(self.elements([i], [j])[0])

    def sigma(self, C: np.ndarray, precision: int = _lib.PQG_FP64) -> np.ndarray:
        """operator*(C): sigma = H C for the columns of C (host arrays)."""
        C = np.asfortranarray(np.atleast_2d(np.asarray(C, dtype=np.float64).T).T)
        S = np.zeros_like(C, order="F")
        check(load().pqg_ci_sigma(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(S), S.shape[0], 0, precision))
        return S

    def sigma_device(self, C_ptr: int, ldc: int, k: int, S_ptr: int, lds: int, precision: int = _lib.PQG_FP64):
        """sigma on device buffers (e.g. torch tensors' data_ptr())."""
        check(load().pqg_ci_sigma(self.h, ct.c_void_p(C_ptr), ldc, k, ct.c_void_p(S_ptr), lds, 1, precision))

    def create_1rdm(self, C: np.ndarray, part: int = 0, spin: int = 0) -> np.ndarray:
        """create_1rdm (src/ci/detset/rdm1.cpp:6-79) of the (part, spin) string set for every
        column of C (host, reference order): an (n_orb, n_orb) array for a 1-D C, else
        (k, n_orb, n_orb).  Computed on the GPU."""
        one = np.ndim(C) == 1
        C = np.asfortranarray(np.atleast_2d(np.asarray(C, dtype=np.float64).T).T)
        # an out-of-range part is passed through so that the library reports it
        n = self.prob.species[part].n_orb if 0 <= part < self.prob.nspecies else 1
        D = np.zeros((C.shape[1], n, n))
        check(load().pqg_ci_rdm1(self.h, _ptr(C), C.shape[0], C.shape[1], part, spin, _ptr(D), 0))
        D = D.transpose(0, 2, 1)  # column-major blocks -> D[k][row][col]
        return D[0] if one else D

    def evaluate_s2(self, C: np.ndarray) -> np.ndarray:
        """evaluate_s2 (src/ci/detset/s2.cpp:5-130): <C|S^2|C> per state (rows) and species
        (columns) for the columns of C (host, reference order).  Computed on the GPU."""
        C = np.asfortranarray(np.atleast_2d(np.asarray(C, dtype=np.float64).T).T)
        out = np.zeros((self.prob.nspecies, C.shape[1]))
        check(load().pqg_ci_s2(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(out), 0))
        return out.T.copy()

    def evaluate_s2_device(self, C_ptr: int, ldc: int, k: int) -> np.ndarray:
        """evaluate_s2 of k device-resident vectors; the (k, nspecies) result on the host."""
        out = np.zeros((self.prob.nspecies, k))
        check(load().pqg_ci_s2(self.h, ct.c_void_p(C_ptr), ldc, k, _ptr(out), 1))
        return out.T.copy()

    def create_1rdm_device(self, C_ptr: int, ldc: int, k: int, part: int, spin: int, D_ptr: int):
        """create_1rdm on device buffers: D receives k column-major n_orb x n_orb blocks."""
        check(load().pqg_ci_rdm1(self.h, ct.c_void_p(C_ptr), ldc, k, part, spin, ct.c_void_p(D_ptr), 1))

    def __matmul__(self, C):
        return self.sigma(C)

    __mul__ = __matmul__

    @property
    def sigma_path(self) -> str:
        """The sigma kernel family of this determinant set: "link_gemm" (full CI on MFMA),
        "pair_gemm" (one particle per string), "enumeration" (every other space, including full
        CI beyond the link GEMM's limits) or "stored_h" (after set_explicit)."""
        names = {_lib.PQG_SIGMA_ENUMERATION: "enumeration", _lib.PQG_SIGMA_LINK_GEMM: "link_gemm",
                 _lib.PQG_SIGMA_PAIR_GEMM: "pair_gemm", _lib.PQG_SIGMA_STORED_H: "stored_h"}
        return names[load().pqg_ci_sigma_path(self.h)]

    # --- parity dumps -------------------------------------------------------------------
    def nstrings(self, part: int, spin: int) -> int:
        return load().pqg_ci_nstrings(self.h, part, spin)

    def strings(self, part: int, spin: int) -> np.ndarray:
        out = np.zeros(self.nstrings(part, spin), dtype=np.uint64)
        check(load().pqg_ci_dump_strings(self.h, part, spin, _ptr(out)))
        return out

    def unfold_all(self) -> np.ndarray:
        out = np.zeros((self.N_dets, 4), dtype=np.int32)
        check(load().pqg_ci_dump_unfold(self.h, _ptr(out)))
        return out

    def excitations(self, which: int, part: int, spin: int):
        """(offsets, J, phase) of every string's singles (1) or doubles (2), reference order."""
        n = self.nstrings(part, spin)
        off = np.zeros(n + 1, dtype=np.int64)
        check(load().pqg_ci_dump_excitations(self.h, which, part, spin, _ptr(off), None, None))
        J = np.zeros(max(int(off[-1]), 1), dtype=np.int32)
        ph = np.zeros(max(int(off[-1]), 1), dtype=np.int8)
        check(load().pqg_ci_dump_excitations(self.h, which, part, spin, _ptr(off), _ptr(J), _ptr(ph)))
        return off, J[: off[-1]], ph[: off[-1]]

    # --- solver -------------------------------------------------------------------------
    def davidson(self, nev: int, nvec_init: int = 5, nvec_max: int = 0, maxit: int = 500, tol: float = 1e-6,
                 precision: int = _lib.PQG_FP64, want_vectors: bool = True, perturb: float = 0.0):
        p = DavParams(nev, nvec_init, nvec_max, maxit, tol, precision, perturb)
        E = np.zeros(nev)
        X = np.zeros((self.N_dets, nev), order="F") if want_vectors else None
        rep = DavReport()
        st = load().pqg_ci_davidson(self.h, ct.byref(p), _ptr(E), _ptr(X) if X is not None else None,
                                    self.N_dets, ct.byref(rep))
        check(st)
        return E, X, rep

    def davidson_solver(self, nev: int, nvec_init: int = 5, nvec_max: int = 0, maxit: int = 500, tol: float = 1e-6,
                        precision: int = _lib.PQG_FP64, want_vectors: bool = True,
                        perturb: float = 0.0) -> DavidsonSolver:
        """The same solve as davidson(), run an iteration at a time through DavidsonSolver.step()."""
        return DavidsonSolver(self, DavParams(nev, nvec_init, nvec_max, maxit, tol, precision, perturb), want_vectors)


def spin_penalty_defaults(prob, spin_penalty_type, spin_penalty=None, expected_S2=None):
    """spin_penalty and expected_S2 as POLYQUANT_EPCI::setup fills them (src/ci/epci.cpp:15-26):
    with a penalty type, an empty spin_penalty becomes default_spin_penalty = 0.1 per species
    (epci.hpp:130) and an empty expected_S2 the SCF <S^2> of each species (input_epscf->S_squared,
    here ``prob.scf_S2``).  A problem without an SCF value needs expected_S2 from the caller for the
    second-order penalty; the first order uses 0 (s2.cpp:136-140).  Returns both lists."""
    nsp = len(prob.species)
    if spin_penalty is not None:
        pen = [float(x) for x in spin_penalty]
    else:
        pen = [0.1] * nsp if spin_penalty_type is not None else [0.0] * nsp
    if expected_S2 is not None:
        exp = [float(x) for x in expected_S2]
    elif spin_penalty_type is not None and getattr(prob, "scf_S2", None) is not None:
        exp = [float(x) for x in prob.scf_S2]
    elif spin_penalty_type == "second_order":
        raise ValueError("second_order spin penalty needs expected_S2 (the problem carries no SCF <S^2>)")
    else:
        exp = [0.0] * nsp
    if len(pen) < nsp or len(exp) < nsp:
        raise ValueError("spin_penalty and expected_S2 need one value per species")
    return pen, exp


class DavidsonSolver:
    """One Davidson solve driven an iteration at a time (pqg_dav_*): step() runs one iteration
    and returns True once the solve has finished; result() then returns (E, X, report) as
    DetSet.davidson does."""

    def __init__(self, ds: "DetSet", params: DavParams, want_vectors: bool = True):
        self.ds = ds
        self.nev = params.nev
        self.want_vectors = want_vectors
        self.h = ct.c_void_p()
        check(load().pqg_dav_create(ds.h, ct.byref(params), ct.byref(self.h)))
        self.done = False

    def step(self) -> bool:
        d = ct.c_int()
        check(load().pqg_dav_step(self.h, ct.byref(d)))
        self.done = bool(d.value)
        return self.done

    def result(self):
        E = np.zeros(self.nev)
        X = np.zeros((self.ds.N_dets, self.nev), order="F") if self.want_vectors else None
        rep = DavReport()
        check(load().pqg_dav_result(self.h, _ptr(E), _ptr(X) if X is not None else None, self.ds.N_dets,
                                    ct.byref(rep)))
        return E, X, rep

    def close(self):
        if self.h:
            load().pqg_dav_destroy(self.h)
            self.h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EPCI:
    """Solve flow of POLYQUANT_EPCI::run (src/ci/epci.cpp:575-662): the matrix-free Davidson path
    or, with exact_diag, the dense Hamiltonian diagonalised in full."""

    def __init__(self, prob, ctx: Optional[Context] = None, num_states: int = 1, num_subspace_vec: int = 5,
                 iteration_max: int = 500, convergence_E: float = 1e-6, precision: int = _lib.PQG_FP64,
                 exact_diag: bool = False, spin_penalty_type: Optional[str] = None,
                 spin_penalty: Optional[list] = None, expected_S2: Optional[list] = None,
                 screening_threshold: float = 0.0, build_matrix: bool = False):
        # ci_keywords spin_penalty_type / spin_penalty / expected_S2 (calculation.cpp:280-312)
        if spin_penalty_type not in (None, "first_order", "second_order"):
            raise ValueError("spin_penalty_type can only be first_order or second_order")
        if spin_penalty_type is not None and not exact_diag and precision != _lib.PQG_FP64:
            raise ValueError("the stored explicit Hamiltonian is FP64 only")
        # ci_keywords screening_threshold (determinant_set.hpp:197, calculation.cpp:370-372): the
        # explicit H drops |H_ij| <= threshold (explicit_ham.cpp:78-80); the matrix-free sigma
        # has no stored H to screen
        # explicit H for the Davidson path (build_matrix, or a screening threshold, which only an
        # explicit H can carry); FP64 only
        if (build_matrix or float(screening_threshold) != 0.0) and not exact_diag and precision != _lib.PQG_FP64:
            raise ValueError("the stored explicit Hamiltonian is FP64 only")
        self.spin_penalty, self.expected_S2 = spin_penalty_defaults(prob, spin_penalty_type, spin_penalty,
                                                                    expected_S2)
        self.prob = prob
        self.detset = DetSet(prob, ctx)
        self.num_states = num_states
        self.num_subspace_vec = num_subspace_vec
        self.iteration_max = iteration_max
        self.convergence_E = convergence_E
        self.precision = precision
        self.exact_diag = exact_diag
        self.first_order_spin_penalty = spin_penalty_type == "first_order"
        self.second_order_spin_penalty = spin_penalty_type == "second_order"
        self.screening_threshold = float(screening_threshold)
        # the penalty is added to the explicit H (epci.cpp:623-631), so it implies build_matrix
        self.build_matrix = bool(build_matrix) or self.screening_threshold != 0.0 or spin_penalty_type is not None
        if self.build_matrix and not exact_diag:
            self.detset.set_explicit(True, self.screening_threshold)
            if self.first_order_spin_penalty or self.second_order_spin_penalty:
                self.add_spin_penalty(None)
        self.constant_shift = prob.e_shift  # E_nuc + frozen-core energy (epci.cpp:584-594)
        self.energies = None
        self.C_ci = None
        self.report = None

    def run(self):
        if self.exact_diag:
            # epci.cpp:622-662: dense H (+ the spin penalty), all eigenpairs on the host, the
            # lowest num_states kept
            H = self.detset.create_ham_dense()
            if self.screening_threshold != 0.0:
                H[np.abs(H) <= self.screening_threshold] = 0.0  # ham.prune(threshold, 1.0)
            if self.first_order_spin_penalty or self.second_order_spin_penalty:
                self.add_spin_penalty(H)
            w, v = np.linalg.eigh(H)
            self.all_energies = w + self.constant_shift
            self.energies = self.all_energies[:self.num_states].copy()
            self.C_ci = np.asfortranarray(v[:, :self.num_states])
            self.report = None
            return self.energies
        nvec_max = max(2 * self.num_subspace_vec, 10 * self.num_states)
        E, X, rep = self.detset.davidson(self.num_states, self.num_subspace_vec, nvec_max, self.iteration_max,
                                         self.convergence_E, self.precision)
        self.energies = E + self.constant_shift
        self.C_ci = X
        self.report = rep
        return self.energies

    def add_spin_penalty(self, H: Optional[np.ndarray]):
        """create_S_sq_penalty (src/ci/detset/s2.cpp:126-153): H += a (S^2) (first order,
        expected 0) or H += a (S^2 - expected)^2 (second order) per species, a = spin_penalty of
        that species, into the dense H (exact_diag) or, for H None, onto the stored explicit H
        of the Davidson path.  As in the reference, the species counter advances only past a nonzero
        penalty, so a zero penalty also switches off every later species."""
        power = 1 if self.first_order_spin_penalty else 2
        idx = 0
        for _ in self.prob.species:
            alpha = self.spin_penalty[idx]
            if alpha == 0.0:
                continue
            expected = 0.0 if power == 1 else self.expected_S2[idx]
            if H is None:  # onto the stored explicit H of the Davidson path
                self.detset.add_spin_penalty(idx, expected, power, alpha)
            else:
                H += alpha * self.detset.create_s2_penalty_dense(idx, expected, power)
            idx += 1
        return H

    def calculate_S_squared(self):
        """POLYQUANT_EPCI::calculate_S_squared (src/ci/epci.cpp:317-332): <S^2> of every state
        (rows) and species (columns)."""
        if self.C_ci is None:
            raise RuntimeError("run() first")
        self.S_squared = self.detset.evaluate_s2(self.C_ci)
        return self.S_squared

    def dump_post_mf_for_qmcpack(self, path: Optional[str] = None, frozen_core: Optional[list] = None) -> dict:
        """POLYQUANT_CALCULATION::dump_post_mf_for_qmcpack (src/calculation/calculation.cpp:622-657):
        the /MultiDet datasets of every state (strings padded with frozen_core orbitals per species,
        see polyquant_amd.qmcpack); with path, also saved as an .npz archive."""
        from . import qmcpack

        if self.C_ci is None:
            raise RuntimeError("run() first")
        arrays = qmcpack.multidet_arrays(self.detset, self.C_ci, [sp.n_orb for sp in self.prob.species], frozen_core,
                                         self.num_states)
        if path is not None:
            qmcpack.save_multidet(path, arrays)
        return arrays

    def calculate_NOs(self, states=None, orbitals=None):
        """POLYQUANT_EPCI::calculate_NOs (src/ci/epci.cpp:200-315) without frozen core: the
        1-RDM of every (species, spin) for each listed state on the GPU, then diag_dm_helper
        (epci.cpp:131-146) on the host: occupations in descending order and the natural orbitals,
        in the MO basis or, when orbitals[part][spin] (MO coefficients) is given, in its basis.
        A restricted species (one spin block) takes its beta matrix from the alpha one
        (epci.cpp:233-235).  Returns {state: {(part, spin): (dm1, occ, orbs)}}."""
        if self.C_ci is None:
            raise RuntimeError("run() first")
        states = list(range(self.num_states)) if states is None else list(states)
        out = {st: {} for st in states}
        C = self.C_ci[:, states]
        for part, sp in enumerate(self.prob.species):
            for spin in range(2 if (sp.n_alpha + sp.n_beta) > 1 else 1):
                if spin == 1 and sp.nspin_blocks == 1:
                    for st in states:
                        out[st][(part, 1)] = out[st][(part, 0)]
                    continue
                D = self.detset.create_1rdm(C, part, spin)
                for q, st in enumerate(states):
                    occ, vec = np.linalg.eigh(D[q])
                    order = np.argsort(-occ, kind="stable")
                    occ, vec = occ[order], vec[:, order]
                    if orbitals is not None:
                        vec = np.asarray(orbitals[part][spin]) @ vec
                    out[st][(part, spin)] = (D[q], occ, vec)
        return out
