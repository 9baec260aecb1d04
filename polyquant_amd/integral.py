"""Host-side mirror of POLYQUANT_INTEGRAL's MO two-body transform over the HIP library.

``transform_mo_2_body_integrals`` stands in for the reference function of that name
(src/integral/integral.cpp:207-477) for one coefficient matrix: it returns the 4-fold
packed ``n_pair x n_pair`` matrix ``(ij|kl)`` addressed by ``idx2(i,j), idx2(k,l)``
(src/integral/integral.hpp:14-20).  The AO integrals come in 8-fold packed form
(``idx8``, integral.hpp:76-88) instead of being recomputed by libint.
"""
from __future__ import annotations

import ctypes as ct
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, load
from .ci import Context, default_context

FP64 = _lib.PQG_AO2MO_FP64
BF16X3 = _lib.PQG_AO2MO_BF16X3


def pack8(ao_full: np.ndarray) -> np.ndarray:
    """8-fold packed vector of a (n,n,n,n) AO tensor: eri8[idx2(idx2(p,q), idx2(r,s))]."""
    n = ao_full.shape[0]
    ii, jj = np.tril_indices(n)  # pairs in idx2 order
    mat = ao_full[ii[:, None], jj[:, None], ii[None, :], jj[None, :]]
    r, c = np.tril_indices(len(ii))
    return np.ascontiguousarray(mat[r, c])


def transform_mo_2_body_integrals(eri8: np.ndarray, C: np.ndarray, mode: int = FP64,
                                  ctx: Optional[Context] = None):
    """(ij|kl) on the GPU from 8-fold packed AO integrals and C (nao x nmo, nmo <= nao: the
    active columns when orbitals are frozen or deleted).  Returns (out4, gpu_seconds), out4 the
    npair(nmo) x npair(nmo) matrix."""
    ctx = ctx or default_context()
    nao, nmo = C.shape
    npa = nao * (nao + 1) // 2
    eri8 = np.ascontiguousarray(eri8, dtype=np.float64)
    assert eri8.size == npa * (npa + 1) // 2
    Cf = np.asfortranarray(C, dtype=np.float64)
    out = np.zeros((nmo * (nmo + 1) // 2,) * 2, order="F")
    sec = ct.c_double()
    check(load().pqg_ao2mo(ctx.h, nao, nmo, eri8.ctypes.data_as(ct.c_void_p), Cf.ctypes.data_as(ct.c_void_p),
                           out.ctypes.data_as(ct.c_void_p), 0, mode, ct.byref(sec)))
    return out, sec.value


def transform_mo_2_body_integrals_pair(eri8: np.ndarray, Ca: np.ndarray, Cb: np.ndarray, mode: int = FP64,
                                       ctx: Optional[Context] = None):
    """transform_mo_2_body_integrals with mo_coeffs_a != mo_coeffs_b (integral.cpp:207-477) in
    one basis: (ij|kl) with the pair ij over Ca (rows) and kl over Cb (columns)."""
    ctx = ctx or default_context()
    nao, nmo = Ca.shape
    npa = nao * (nao + 1) // 2
    eri8 = np.ascontiguousarray(eri8, dtype=np.float64)
    assert eri8.size == npa * (npa + 1) // 2 and Cb.shape == Ca.shape
    Af = np.asfortranarray(Ca, dtype=np.float64)
    Bf = np.asfortranarray(Cb, dtype=np.float64)
    out = np.zeros((nmo * (nmo + 1) // 2,) * 2, order="F")
    sec = ct.c_double()
    check(load().pqg_ao2mo_pair(ctx.h, nao, nmo, eri8.ctypes.data_as(ct.c_void_p), Af.ctypes.data_as(ct.c_void_p),
                                Bf.ctypes.data_as(ct.c_void_p), out.ctypes.data_as(ct.c_void_p), 0, mode,
                                ct.byref(sec)))
    return out, sec.value


def transform_mo_2_body_integrals_mixed(eri_ab: np.ndarray, Ca: np.ndarray, Cb: np.ndarray, mode: int = FP64,
                                        ctx: Optional[Context] = None):
    """transform_mo_2_body_integrals for two species with their own bases (integral.cpp:207-477,
    num_ao_a != num_ao_b): eri_ab is the npair(nao_a) x npair(nao_b) AO block (pq|rs), p, q in
    basis a and r, s in basis b; Ca (nao_a x nmo_a) transforms the row pair, Cb (nao_b x nmo_b)
    the column pair.  Returns (out4, gpu_seconds), out4 npair(nmo_a) x npair(nmo_b) --
    mo_two_body_ints[a][sa][b][sb] of calculate_mo_2_body_integrals (:479-516)."""
    ctx = ctx or default_context()
    nao_a, nmo_a = Ca.shape
    nao_b, nmo_b = Cb.shape
    eri_ab = np.asfortranarray(eri_ab, dtype=np.float64)
    assert eri_ab.shape == (nao_a * (nao_a + 1) // 2, nao_b * (nao_b + 1) // 2)
    Af = np.asfortranarray(Ca, dtype=np.float64)
    Bf = np.asfortranarray(Cb, dtype=np.float64)
    out = np.zeros((nmo_a * (nmo_a + 1) // 2, nmo_b * (nmo_b + 1) // 2), order="F")
    sec = ct.c_double()
    check(load().pqg_ao2mo_mixed(ctx.h, nao_a, nmo_a, nao_b, nmo_b, eri_ab.ctypes.data_as(ct.c_void_p),
                                 Af.ctypes.data_as(ct.c_void_p), Bf.ctypes.data_as(ct.c_void_p),
                                 out.ctypes.data_as(ct.c_void_p), 0, mode, ct.byref(sec)))
    return out, sec.value


def cholesky_eri8_device(L, block: int = 2048):
    """8-fold packed AO integrals of Cholesky-like vectors, built on the GPU (torch):
    eri8[idx2(P, Q)] = sum_k L[k, P] L[k, Q] for pair indices Q <= P.  L: (nvec, npair) float64,
    host array or device tensor.  Returns (eri8, L_device)."""
    import torch

    Lg = torch.as_tensor(L, dtype=torch.float64, device="cuda")
    npr = Lg.shape[1]
    eri8 = torch.empty(npr * (npr + 1) // 2, dtype=torch.float64, device=Lg.device)
    for p0 in range(0, npr, block):
        p1 = min(npr, p0 + block)
        M = Lg[:, p0:p1].T.contiguous() @ Lg[:, :p1]
        mask = torch.arange(p1, device=Lg.device)[None, :] <= torch.arange(p0, p1, device=Lg.device)[:, None]
        eri8[p0 * (p0 + 1) // 2: p1 * (p1 + 1) // 2] = M[mask]
        del M, mask
    return eri8, Lg


def transform_device(ctx: Context, n: int, eri8_ptr: int, C_ptr: int, out_ptr: int, mode: int = FP64) -> float:
    """Same on device buffers (C column-major n x n, out column-major npair x npair)."""
    sec = ct.c_double()
    check(load().pqg_ao2mo(ctx.h, n, n, ct.c_void_p(eri8_ptr), ct.c_void_p(C_ptr), ct.c_void_p(out_ptr), 1, mode,
                           ct.byref(sec)))
    return sec.value


# --- frozen core (src/ci/epci.cpp:47-129, src/integral/integral.cpp:166-205, 565-698) ---------

def fc_ints_block(out: np.ndarray, eri: np.ndarray, dm_b, same_particle: bool, a_spin: int, b_spin: int,
                  restricted_b: bool, charge_prod: float, ctx: Optional[Context] = None) -> np.ndarray:
    """compute_frozen_core_ints (integral.cpp:565-698) on the GPU: out += the frozen-core
    operator of one (species b, spin b) core density.  eri: 8-fold packed AO integrals (same
    species) or the npair_a x npair_b (ij|kl) block (two species); dm_b: fc_dm[b] (1 or 2
    matrices)."""
    ctx = ctx or default_context()
    na = out.shape[0]
    nb = dm_b[0].shape[0]
    if len(dm_b) not in (1, 2):
        raise ValueError("dm_b holds one (restricted) or two (unrestricted) core densities")
    if not restricted_b and len(dm_b) == 1:
        # an unrestricted species with one particle has only its alpha block (calculate_fc_energy
        # makes a second block only for num_parts > 1, epci.cpp:47-129): its beta core density is
        # zero, and pqg_fc_ints reads two blocks for an unrestricted species
        dm_b = [dm_b[0], np.zeros_like(np.asarray(dm_b[0], dtype=np.float64))]
    # block-major, each block column-major
    dm = np.concatenate([np.asarray(d, dtype=np.float64).ravel(order="F") for d in dm_b])
    eri = np.ascontiguousarray(eri, dtype=np.float64) if same_particle else np.asfortranarray(eri, dtype=np.float64)
    res = np.asfortranarray(out, dtype=np.float64).copy(order="F")
    check(load().pqg_fc_ints(ctx.h, na, nb, eri.ctypes.data_as(ct.c_void_p), int(same_particle), a_spin, b_spin,
                             int(restricted_b), float(charge_prod), dm.ctypes.data_as(ct.c_void_p),
                             res.ctypes.data_as(ct.c_void_p)))
    out[...] = res
    return out


def calculate_frozen_core_ints(species, fc_dm, frozen_core, eri_ao, kinetic, nuclear,
                               ctx: Optional[Context] = None):
    """POLYQUANT_INTEGRAL::calculate_frozen_core_ints (integral.cpp:166-205): for every species a
    and spin block of fc_dm[a], the sum over species b with frozen_core[b] != 0 and over the spin
    blocks of fc_dm[b] of compute_frozen_core_ints, plus kinetic[a] - charge_a nuclear[a].
    species: objects with .charge and .restricted; eri_ao[(a, b)]: the 8-fold packed block (a == b)
    or the npair_a x npair_b block (a < b; a > b uses its transpose)."""
    out = []
    for a, spa in enumerate(species):
        na = kinetic[a].shape[0]
        per_spin = []
        for sa in range(len(fc_dm[a])):
            F = np.zeros((na, na))
            for b, spb in enumerate(species):
                for sb in range(len(fc_dm[b])):
                    if frozen_core[b] != 0:
                        eri = eri_ao[(a, b)] if a <= b else np.asarray(eri_ao[(b, a)]).T
                        fc_ints_block(F, eri, fc_dm[b], a == b, sa, sb, spb.restricted, spa.charge * spb.charge, ctx)
            F += kinetic[a] + (-spa.charge * nuclear[a])
            per_spin.append(F)
        out.append(per_spin)
    return out


def frozen_core_density(C, nfc: int) -> np.ndarray:
    """form_occ_helper_aufbau (the nfc lowest orbitals, occupation 1) then form_DM_helper:
    D = sum_i occ_i C[:, i] C[:, i]^T (epci.cpp:80-94)."""
    C = np.asarray(C, dtype=np.float64)
    Cc = C[:, :nfc]
    return Cc @ Cc.T


def calculate_fc_energy(species, C, frozen_core, eri_ao, kinetic, nuclear, ctx: Optional[Context] = None):
    """POLYQUANT_EPCI::calculate_fc_energy (epci.cpp:47-129).  species: objects with .charge,
    .restricted and .num_parts; C[a][s]: MO coefficients (nao x nmo) per spin block.  Returns
    (fc_dm, frozen_core_ints, frozen_core_energy per species)."""
    fc_dm = []
    for a, sp in enumerate(species):
        nao = kinetic[a].shape[0]
        two = sp.num_parts > 1 and not sp.restricted
        blocks = [np.zeros((nao, nao)) for _ in range(2 if two else 1)]
        if frozen_core[a] != 0:
            blocks[0] = frozen_core_density(C[a][0], frozen_core[a])
            if two:
                blocks[1] = frozen_core_density(C[a][1], frozen_core[a])
        fc_dm.append(blocks)
    fci = calculate_frozen_core_ints(species, fc_dm, frozen_core, eri_ao, kinetic, nuclear, ctx)
    energy = []
    for a, sp in enumerate(species):
        e = 0.0
        if frozen_core[a] != 0:
            Hc = kinetic[a] + (-sp.charge * nuclear[a])
            if not sp.restricted:
                d1 = fc_dm[a][1] if len(fc_dm[a]) > 1 else fc_dm[a][0]
                f1 = fci[a][1] if len(fci[a]) > 1 else fci[a][0]
                e = 0.5 * float(np.sum((fc_dm[a][0] + d1) * Hc) + np.sum(fc_dm[a][0] * fci[a][0]) + np.sum(d1 * f1))
            else:
                e = float(np.sum(fc_dm[a][0] * (Hc + fci[a][0])))
        energy.append(e)
    return fc_dm, fci, energy


def calculate_mo_1_body_integrals(C, frozen_core, deleted_virtual, frozen_core_ints):
    """integral.cpp:565-610 with frozen-core operators: h[a][s] = C_sub^T F[a][s] C_sub,
    C_sub = the columns frozen_core[a] .. nmo - deleted_virtual[a] - 1 of C[a][s]."""
    out = []
    for a, Cs in enumerate(C):
        per = []
        for s, Cm in enumerate(Cs):
            Cm = np.asarray(Cm, dtype=np.float64)
            nmo = Cm.shape[1] - frozen_core[a] - deleted_virtual[a]
            sub = Cm[:, frozen_core[a]:frozen_core[a] + nmo]
            F = frozen_core_ints[a][min(s, len(frozen_core_ints[a]) - 1)]
            per.append(sub.T @ F @ sub)
        out.append(per)
    return out


def active_two_body(out4: np.ndarray, nmo: int, first: int, nact: int) -> np.ndarray:
    """The (ij|kl) block of the orbitals first .. first + nact - 1 from the full 4-fold matrix of
    nmo orbitals (the transform with the MO subset of integral.cpp:207-477)."""
    def pairs(n, off):
        ii, jj = [], []
        for i in range(n):
            for j in range(i + 1):
                ii.append(i + off)
                jj.append(j + off)
        return np.array(ii), np.array(jj)

    ia, ja = pairs(nact, first)
    idx = np.maximum(ia, ja) * (np.maximum(ia, ja) + 1) // 2 + np.minimum(ia, ja)
    return np.ascontiguousarray(np.asarray(out4)[np.ix_(idx, idx)])


def frozen_core_problem(species, C, kinetic, nuclear, eri_ao, e_nuc: float, frozen_core, deleted_virtual,
                        nroots: int = 1, max_collective: Optional[int] = None, ctx: Optional[Context] = None):
    """calculate_fc_energy, calculate_mo_1_body_integrals and calculate_mo_2_body_integrals with
    frozen core and deleted virtuals (epci.cpp:27-129, 575-594; integral.cpp:166-205, 207-516,
    565-610) for one species or two (electron, positron): returns the active-space CIProblem whose
    e_shift is e_nuc plus the frozen-core energies, and those energies.  species: objects with
    .charge, .restricted, .n_alpha, .n_beta (all particles, core included) and .ex_level;
    C[a][s]: the MO coefficients of species a (one block, or alpha and beta when unrestricted);
    eri_ao[(a, a)]: the 8-fold packed AO integrals of species a, eri_ao[(a, b)] (a < b): the
    npair_a x npair_b AO block (pq|rs) between the two bases.  Every two-body block is
    transformed over the active orbitals only (the columns frozen_core[a] ..
    nmo - deleted_virtual[a] - 1), as the reference's transform does (integral.cpp:214-217,
    :298-300)."""
    from types import SimpleNamespace

    from .workloads import CIProblem, Species

    ns = len(species)
    if not 1 <= ns <= 2:
        raise ValueError("frozen_core_problem takes one or two species")
    # the caller's objects are left as they are; calculate_fc_energy reads num_parts
    spc = [SimpleNamespace(charge=sp.charge, restricted=sp.restricted, num_parts=sp.n_alpha + sp.n_beta)
           for sp in species]
    fc_dm, fci, e_fc = calculate_fc_energy(spc, C, frozen_core, eri_ao, kinetic, nuclear, ctx)
    h = calculate_mo_1_body_integrals(C, frozen_core, deleted_virtual, fci)
    act, nact = [], []
    for a in range(ns):
        nmo = np.asarray(C[a][0]).shape[1]
        na = nmo - frozen_core[a] - deleted_virtual[a]
        if na < 1:
            raise ValueError(f"species {a}: no active orbitals")
        act.append([np.asfortranarray(np.asarray(Cs, dtype=np.float64)[:, frozen_core[a]:frozen_core[a] + na])
                    for Cs in C[a]])
        nact.append(na)
    # calculate_mo_2_body_integrals (integral.cpp:479-516): every (a, sa, b, sb) with b >= a
    eri = {}
    for a in range(ns):
        for s1 in range(len(C[a])):
            for b in range(a, ns):
                for s2 in range(len(C[b])):
                    if a == b and s2 < s1:
                        continue
                    Ca, Cb = act[a][s1], act[b][s2]
                    if a != b:
                        out4, _ = transform_mo_2_body_integrals_mixed(eri_ao[(a, b)], Ca, Cb, ctx=ctx)
                    elif s1 == s2:
                        out4, _ = transform_mo_2_body_integrals(eri_ao[(a, a)], Ca, ctx=ctx)
                    else:
                        out4, _ = transform_mo_2_body_integrals_pair(eri_ao[(a, a)], Ca, Cb, ctx=ctx)
                    eri[(a, s1, b, s2)] = out4
        if len(C[a]) == 2:
            eri[(a, 1, a, 0)] = np.ascontiguousarray(eri[(a, 0, a, 1)].T)
    acts = [Species(nact[a], sp.n_alpha - frozen_core[a], sp.n_beta - frozen_core[a], sp.charge, len(C[a]),
                    getattr(sp, "ex_level", None)) for a, sp in enumerate(species)]
    prob = CIProblem(species=acts, h=[h[a] for a in range(ns)], eri=eri, nroots=nroots,
                     e_shift=e_nuc + sum(e_fc),
                     name="frozen_core_" + "_".join(f"{frozen_core[a]}_{deleted_virtual[a]}" for a in range(ns)))
    if max_collective is not None:
        prob.max_collective = int(max_collective)
    return prob, e_fc
