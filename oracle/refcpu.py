"""ctypes wrapper for the oracle (``librefcpu.so``, a CPU restatement of the reference CI).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librefcpu.so")
_lib = None

_dp = ct.POINTER(ct.c_double)


class RcDesc(ct.Structure):
    _fields_ = [
        ("nspecies", ct.c_int),
        ("n_orb", ct.c_int * 2),
        ("n_alpha", ct.c_int * 2),
        ("n_beta", ct.c_int * 2),
        ("charge", ct.c_double * 2),
        ("nspin_blocks", ct.c_int * 2),
        ("ex_level", (ct.c_int * 3) * 2),
        ("max_collective", ct.c_int),
        ("h", (_dp * 2) * 2),
        ("eri", (((_dp * 2) * 2) * 2) * 2),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.rc_create.restype = ct.c_void_p
        L.rc_create.argtypes = [ct.POINTER(RcDesc)]
        L.rc_destroy.argtypes = [ct.c_void_p]
        L.rc_last_error.restype = ct.c_char_p
        for f in ("rc_ndets", "rc_nstrings", "rc_address"):
            getattr(L, f).restype = ct.c_int64
        L.rc_ndets.argtypes = [ct.c_void_p]
        L.rc_nstrings.argtypes = [ct.c_void_p, ct.c_int, ct.c_int]
        L.rc_strings.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
        L.rc_unfold_all.argtypes = [ct.c_void_p, ct.c_void_p]
        L.rc_address.argtypes = [ct.c_void_p, ct.c_void_p]
        L.rc_excitations.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int]
        L.rc_diag.argtypes = [ct.c_void_p, ct.c_void_p]
        L.rc_slater_condon.restype = ct.c_double
        L.rc_slater_condon.argtypes = [ct.c_void_p, ct.c_int64, ct.c_int64]
        L.rc_sigma.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64]
        L.rc_sigma_slow.argtypes = L.rc_sigma.argtypes
        L.rc_rdm1.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
        L.rc_s2.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p]
        L.rc_rdm1_sample.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_int, ct.c_void_p, ct.c_int64]
        L.rc_sigma_sample.argtypes = L.rc_sigma.argtypes + [ct.c_int64]
        L.rc_sigma_rows.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int, ct.c_void_p, ct.c_int64,
                                    ct.c_void_p]
        L.rc_davidson.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_double,
                                  ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.rc_get_holes.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p]
        L.rc_get_parts.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int, ct.c_void_p]
        L.rc_get_phase.restype = ct.c_double
        L.rc_get_phase.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_int]
        L.rc_get_occ_virt.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                      ct.c_void_p]
        L.rc_num_excitation.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_int]
        L.rc_ao2mo.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        L.rc_ao2mo2.argtypes = [ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p,
                                ct.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


# --- free bit operations (multi-word determinants) -------------------------------------

def _words(d):
    return np.ascontiguousarray(np.asarray(d, dtype=np.uint64))


def get_holes(di, dj):
    di, dj = _words(di), _words(dj)
    out = np.zeros(64 * len(di), dtype=np.int32)
    n = lib().rc_get_holes(_ptr(di), _ptr(dj), len(di), _ptr(out))
    return [int(x) for x in out[:n]]


def get_parts(di, dj):
    di, dj = _words(di), _words(dj)
    out = np.zeros(64 * len(di), dtype=np.int32)
    n = lib().rc_get_parts(_ptr(di), _ptr(dj), len(di), _ptr(out))
    return [int(x) for x in out[:n]]


def get_phase(di, holes, parts):
    di = _words(di)
    h = np.asarray(holes, dtype=np.int32)
    p = np.asarray(parts, dtype=np.int32)
    return lib().rc_get_phase(_ptr(di), len(di), _ptr(h), _ptr(p), len(h))


def get_occ_virt(d, max_orb):
    d = _words(d)
    occ = np.zeros(64 * len(d), dtype=np.int32)
    virt = np.zeros(64 * len(d), dtype=np.int32)
    no, nv = ct.c_int(), ct.c_int()
    lib().rc_get_occ_virt(_ptr(d), len(d), max_orb, _ptr(occ), ct.byref(no), _ptr(virt), ct.byref(nv))
    return [int(x) for x in occ[: no.value]], [int(x) for x in virt[: nv.value]]


def num_excitation(di, dj):
    di, dj = _words(di), _words(dj)
    return lib().rc_num_excitation(_ptr(di), _ptr(dj), len(di))


# --- determinant set --------------------------------------------------------------------

class OracleCI:
    """The reference's CI determinant set and operators, restated on the CPU."""

    def __init__(self, prob):
        self.prob = prob
        d = RcDesc()
        d.nspecies = prob.nspecies
        d.max_collective = min(int(prob.max_collective), 2**31 - 1)
        self._keep = []
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
                self._keep.append(h)
                d.h[p][s] = h.ctypes.data_as(_dp)
        for (a, sa, b, sb), v in prob.eri.items():
            arr = np.asfortranarray(v, dtype=np.float64)
            self._keep.append(arr)
            d.eri[a][sa][b][sb] = arr.ctypes.data_as(_dp)
        self.desc = d
        self.h = lib().rc_create(ct.byref(d))
        if not self.h:
            raise RuntimeError(lib().rc_last_error().decode())
        self.N = lib().rc_ndets(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().rc_destroy(self.h)
            self.h = None

    def rows(self):
        return self.N

    def cols(self):
        return self.N

    def strings(self, p, s):
        n = lib().rc_nstrings(self.h, p, s)
        out = np.zeros(n, dtype=np.uint64)
        lib().rc_strings(self.h, p, s, _ptr(out))
        return out

    def unfold_all(self):
        out = np.zeros((self.N, 4), dtype=np.int32)
        lib().rc_unfold_all(self.h, _ptr(out))
        return out

    def address(self, t):
        t4 = np.zeros(4, dtype=np.int32)
        t4[: len(t)] = t
        return lib().rc_address(self.h, _ptr(t4))

    def excitations(self, which, p, s, idx):
        cap = 1 << 16
        out = np.zeros(cap, dtype=np.int32)
        n = lib().rc_excitations(self.h, which, p, s, idx, _ptr(out), cap)
        return out[:n].copy()

    def diag(self):
        out = np.zeros(self.N)
        lib().rc_diag(self.h, _ptr(out))
        return out

    def __call__(self, i, j):
        return lib().rc_slater_condon(self.h, int(i), int(j))

    def dense(self):
        H = np.zeros((self.N, self.N))
        for i in range(self.N):
            for j in range(self.N):
                H[i, j] = self(i, j)
        return H

    def sigma(self, C):
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        S = np.zeros_like(C, order="F")
        lib().rc_sigma(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(S), S.shape[0])
        return S

    def rdm1(self, C, part=0, spin=0):
        """create_1rdm (rdm1.cpp:6-79) of the (part, spin) string set: (n, n) for a 1-D C,
        else (k, n, n) with D[v][row][col]."""
        one = np.ndim(C) == 1
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        n = self.prob.species[part].n_orb
        D = np.zeros((C.shape[1], n, n))
        if lib().rc_rdm1(self.h, _ptr(C), C.shape[0], C.shape[1], part, spin, _ptr(D)) != 0:
            raise ValueError(lib().rc_last_error().decode())
        D = D.transpose(0, 2, 1)
        return D[0] if one else D

    def s2(self, C):
        """evaluate_s2 (s2.cpp:5-130): <C|S^2|C> per state (rows) and species (columns)."""
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        out = np.zeros((self.prob.nspecies, C.shape[1]))
        lib().rc_s2(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(out))
        return out.T.copy()

    def s2_penalty(self, part, expected=0.0, power=1):
        """create_S_sq_minus_expected_S_sq_matrix_singleshot (s2.cpp:156-255) as a dense matrix
        (pure-Python loops, small spaces): for every determinant the diagonal
        num_a + sz (sz - 1) - expected, and for every pair of an alpha single and a beta single
        of it whose target lies in the space, -phase_a phase_b when the alpha hole is the beta
        particle and the beta hole the alpha particle.  power 2: P @ P (s2.cpp:147)."""
        sa, sb = 2 * part, 2 * part + 1
        Sa, Sb = self.strings(part, 0), self.strings(part, 1)
        U = self.unfold_all()
        index = {tuple(int(x) for x in U[i]): i for i in range(self.N)}
        P = np.zeros((self.N, self.N))
        for i in range(self.N):
            ia, ib = int(U[i, sa]), int(U[i, sb])
            da, db = int(Sa[ia]), int(Sb[ib])
            x = da ^ db
            num_total = bin(x).count("1")
            num_a = bin(x & da).count("1")
            sz = num_a - 0.5 * num_total
            P[i, i] += num_a + sz * (sz - 1.0) - expected
            for ja in self.excitations(1, part, 0, ia):
                for jb in self.excitations(1, part, 1, ib):
                    t = [int(v) for v in U[i]]
                    t[sa], t[sb] = int(ja), int(jb)
                    j = index.get(tuple(t))
                    if j is None:
                        continue
                    ea, eb = int(Sa[ja]), int(Sb[jb])
                    ha, pa = get_holes([da], [ea]), get_parts([da], [ea])
                    hb, pb = get_holes([db], [eb]), get_parts([db], [eb])
                    if ha[0] == pb[0] and hb[0] == pa[0]:
                        phase = get_phase([da], ha, pa) * get_phase([db], hb, pb)
                        P[i, j] += -phase
        return P @ P if power == 2 else P

    def rdm1_sample(self, c, part, spin, stride):
        """create_1rdm over determinants i = 0 mod stride only (a timing sample)."""
        c = np.ascontiguousarray(c, dtype=np.float64)
        n = self.prob.species[part].n_orb
        D = np.zeros((n, n))
        if lib().rc_rdm1_sample(self.h, _ptr(c), c.shape[0], part, spin, _ptr(D), int(stride)) != 0:
            raise ValueError(lib().rc_last_error().decode())
        return D.T

    def sigma_sample(self, C, stride):
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        S = np.zeros_like(C, order="F")
        lib().rc_sigma_sample(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(S), S.shape[0], int(stride))
        return S

    def sigma_rows(self, C, rows):
        """Exact sigma on the listed rows (slow sigma over the connected determinants)."""
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        rows = np.ascontiguousarray(rows, dtype=np.int64)
        out = np.zeros((len(rows), C.shape[1]), order="F")
        lib().rc_sigma_rows(self.h, _ptr(C), C.shape[0], C.shape[1], rows.ctypes.data_as(ct.c_void_p), len(rows),
                            _ptr(out))
        return out

    def sigma_slow(self, C):
        C = np.asfortranarray(np.atleast_2d(C.T).T, dtype=np.float64)
        S = np.zeros_like(C, order="F")
        lib().rc_sigma_slow(self.h, _ptr(C), C.shape[0], C.shape[1], _ptr(S), S.shape[0])
        return S

    def davidson(self, nev, nvec_init=5, nvec_max=None, maxit=500, tol=1e-6):
        """Mirror of epci.cpp:589-608: nvec_max defaults to max(2*nvec_init, 10*nev)."""
        if nvec_max is None:
            nvec_max = max(2 * nvec_init, 10 * nev)
        ev = np.zeros(nev)
        res = np.zeros(nev)
        X = np.zeros((self.N, nev), order="F")
        it = lib().rc_davidson(self.h, nev, nvec_init, nvec_max, maxit, tol, _ptr(ev), _ptr(X), _ptr(res))
        return ev, X, res, it


def frozen_core_ints_block(out, eri, fc_dm_b, same_particle, a_spin, b_spin, restricted_b, charge_prod):
    """compute_frozen_core_ints (integral.cpp:611-698) restated over basis functions, each its
    own shell: unique pairs j <= i of species a and l <= k of species b (no screening), the
    pair degeneracy factors, the Coulomb update of (i, j) and (j, i) and, for the same species
    and spin, the eight exchange updates with weight 1/8.  eri: the full (na, na, nb, nb) AO
    tensor.  Adds into out and returns it (pure-Python loops, small bases)."""
    na, nb = eri.shape[0], eri.shape[2]
    exchange = same_particle and a_spin == b_spin
    for i in range(na):
        for j in range(i + 1):
            for k in range(nb):
                for l in range(k + 1):
                    v = eri[i, j, k, l]
                    deg = (1.0 if i == j else 2.0) * (1.0 if k == l else 2.0)
                    spinscale = 0.5 if (same_particle and not restricted_b) else 1.0
                    scaleall = 0.5 * spinscale if same_particle else 0.5 * charge_prod * spinscale
                    if restricted_b:
                        d_kl = 2.0 * fc_dm_b[0][k, l]
                    else:
                        d_kl = fc_dm_b[0][k, l] + fc_dm_b[1][k, l]
                    out[i, j] += scaleall * deg * d_kl * v
                    out[j, i] += scaleall * deg * d_kl * v
                    if exchange:
                        D = fc_dm_b[0] if restricted_b else fc_dm_b[b_spin]
                        w = 0.125 * deg * v
                        out[i, k] -= w * D[j, l]
                        out[k, i] -= w * D[j, l]
                        out[j, l] -= w * D[i, k]
                        out[l, j] -= w * D[i, k]
                        out[i, l] -= w * D[j, k]
                        out[l, i] -= w * D[j, k]
                        out[j, k] -= w * D[i, l]
                        out[k, j] -= w * D[i, l]
    return out


def ao2mo(ao_full, Ca, Cb=None):
    """Quarter-transform restatement; ao_full is the AO tensor (pq|rs) of shape
    (nao_a, nao_a, nao_b, nao_b): one basis (Cb defaults to Ca), or two bases for the
    electron-positron block, Ca (nao_a x nmo_a) transforming p, q and Cb (nao_b x nmo_b) r, s."""
    if Cb is None:
        Cb = Ca
    na, nb = ao_full.shape[0], ao_full.shape[2]
    assert ao_full.shape == (na, na, nb, nb) and Ca.shape[0] == na and Cb.shape[0] == nb
    nma, nmb = Ca.shape[1], Cb.shape[1]
    ao = np.ascontiguousarray(ao_full, dtype=np.float64)
    Ca = np.asfortranarray(Ca, dtype=np.float64)
    Cb = np.asfortranarray(Cb, dtype=np.float64)
    out = np.zeros((nma * (nma + 1) // 2, nmb * (nmb + 1) // 2), order="F")
    lib().rc_ao2mo2(na, nb, nma, nmb, _ptr(ao), _ptr(Ca), _ptr(Cb), _ptr(out))
    return out


# --- QMCPACK multideterminant dump (reference's word loops restated, 64-bit words) ------------
_M64 = (1 << 64) - 1


def get_det_withfcorbs(det, nfc, max_orb, frozen_core):
    """POLYQUANT_DETSET::get_det_withfcorbs (src/ci/detset/determinant_creation.cpp:391-454),
    step for step: det is a string as a list of 64-bit words, most significant first (orbital o
    in word W-1-(o>>6)); num_int from the largest max_orb + frozen_core over the species
    (:396-403); with no frozen core the string is returned, or prefixed by a zero word
    (:405-413); otherwise the overflow of the top word opens a new word (:418-426), each word
    is shifted left by nfc and takes the top nfc bits of the next word (:427-442), and the
    lowest word gets the nfc core orbitals set (:446-448)."""
    one = 1
    max_orb_with_fc = [m + f for m, f in zip(max_orb, frozen_core)]
    if nfc > 64:
        raise ValueError("Can't handle >=64 core orbitals")
    num_int = (max(max_orb_with_fc) >> 6) + 1
    if nfc == 0:
        if num_int == len(det):
            return list(det)
        return [0] + list(det)
    new_det = []
    mask = ((one << nfc) - one) & _M64
    if num_int != len(det):
        new_det.append((det[0] >> (64 - nfc)) & mask)
    for count, i in enumerate(det):
        j = (i << nfc) & _M64
        if count + 1 != len(det):
            j |= (det[count + 1] >> (64 - nfc)) & mask
        if count == len(det) - 1:
            j |= ~(~0 & (~(one << nfc) + one)) & _M64
        new_det.append(j)
    return new_det


def qmcpack_multidet(oc, C, max_orb, frozen_core, n_states):
    """The /MultiDet datasets of dump_post_mf_for_qmcpack (src/calculation/calculation.cpp:622-657)
    and dump_post_mf_to_hdf5_for_QMCPACK (src/io/hdf5_utilities.cpp:399-468): per (part, spin)
    the padded string of every determinant (address order, det_idx_unfold) with its words
    reversed (CI_{2 part + spin}, N_dets x N_int), one coefficient column per state (Coeff,
    Coeff_1, ...), NbDet, Nbits (words of part 0's strings), nexcitedstate and nstate (the
    largest max_orb + frozen_core)."""
    u = oc.unfold_all()
    nsp = len(max_orb)
    W = (max(max_orb) >> 6) + 1  # words of the unpadded strings (determinant_creation.cpp:12-13)
    out = {}
    nbits = None
    for p in range(nsp):
        for s in range(2):
            strs = oc.strings(p, s)
            rows = []
            for i in range(oc.N):
                x = int(strs[u[i, 2 * p + s]])
                det = [(x >> (64 * (W - 1 - w))) & _M64 for w in range(W)]
                rows.append(get_det_withfcorbs(det, frozen_core[p], max_orb, frozen_core))
            if nbits is None:
                nbits = len(rows[0])
            out[f"CI_{2 * p + s}"] = np.array([[r[nbits - j - 1] for j in range(nbits)] for r in rows],
                                             dtype=np.uint64)
    for i in range(n_states):
        out["Coeff" if i == 0 else f"Coeff_{i}"] = np.array([C[j, i] for j in range(oc.N)], dtype=np.float64)
    out["NbDet"] = oc.N
    out["Nbits"] = nbits
    out["nexcitedstate"] = n_states
    out["nstate"] = max(m + f for m, f in zip(max_orb, frozen_core))
    return out
