"""Brute-force second quantisation on tiny CI spaces (test helper).

H is assembled term by term from creation/annihilation operators acting on occupation-number
strings, independently of any Slater-Condon rule:

    H = sum_{species} [ sum_{s,pq} h^s_pq E^s_pq
                        + 1/2 sum_{s,t,pqrs} (pq|rs)_{st} (E^s_pq E^t_rs - delta_st delta_qr E^s_ps) ]
        + sum_{s in e, t in p} q_e q_p sum_pqrs (pq|rs)_{e s, p t} E^{e s}_pq E^{p t}_rs

with E^s_pq = a+_{p s} a_{q s}.  Determinants are products of creation operators in
ascending orbital order, one string per (species, spin).  The sign of a+_p a_q on a string is
(-1)^(occupied orbitals strictly between p and q).
"""
from __future__ import annotations

import numpy as np

from polyquant_amd.workloads import idx2


def apply_E(string: int, p: int, q: int):
    """a+_p a_q |string>  ->  (sign, new_string) or None."""
    if not (string >> q) & 1:
        return None
    s = string & ~(1 << q)
    if p != q and (s >> p) & 1:
        return None
    lo, hi = min(p, q), max(p, q)
    between = s & (((1 << hi) - 1) & ~((1 << (lo + 1)) - 1)) if hi > lo else 0
    sign = -1.0 if bin(between).count("1") & 1 else 1.0
    return sign, s | (1 << p)


def _eri(prob, a, sa, b, sb, p, q, r, s):
    if b < a:
        a, sa, b, sb, p, q, r, s = b, sb, a, sa, r, s, p, q
    return prob.eri[(a, sa, b, sb)][idx2(p, q), idx2(r, s)]


def hamiltonian(prob, dets):
    """Dense H over `dets`, a list of tuples of strings (ea, eb[, pa, pb])."""
    index = {d: i for i, d in enumerate(dets)}
    nsp = prob.nspecies
    H = np.zeros((len(dets), len(dets)))
    slots = [(sp, s) for sp in range(nsp) for s in range(2)]

    def spin_block(sp, s):
        return s % prob.species[sp].nspin_blocks

    for j, det in enumerate(dets):
        out = {}

        def add(d, v):
            out[d] = out.get(d, 0.0) + v

        for si, (sp, s) in enumerate(slots):
            n = prob.species[sp].n_orb
            sb = spin_block(sp, s)
            for p in range(n):
                for q in range(n):
                    r1 = apply_E(det[si], p, q)
                    if r1 is None:
                        continue
                    sg, ns = r1
                    nd = det[:si] + (ns,) + det[si + 1:]
                    add(nd, sg * prob.h[sp][sb][p, q])
        # two-body: pairs of slots
        for si, (spa, sa) in enumerate(slots):
            for ti, (spb, sbb) in enumerate(slots):
                if spa == spb:
                    fac = 0.5
                elif spa < spb:
                    fac = prob.species[spa].charge * prob.species[spb].charge
                else:
                    continue
                na, nb = prob.species[spa].n_orb, prob.species[spb].n_orb
                ba, bb = spin_block(spa, sa), spin_block(spb, sbb)
                for r in range(nb):
                    for s_ in range(nb):
                        r1 = apply_E(det[ti], r, s_)
                        if r1 is None:
                            continue
                        sg1, ns1 = r1
                        d1 = det[:ti] + (ns1,) + det[ti + 1:]
                        for p in range(na):
                            for q in range(na):
                                r2 = apply_E(d1[si], p, q)
                                if r2 is None:
                                    continue
                                sg2, ns2 = r2
                                d2 = d1[:si] + (ns2,) + d1[si + 1:]
                                add(d2, fac * sg1 * sg2 * _eri(prob, spa, ba, spb, bb, p, q, r, s_))
                if spa == spb and sa == sbb:
                    # - 1/2 sum_pqs (pq|qs) E_ps
                    for p in range(na):
                        for s_ in range(na):
                            r1 = apply_E(det[si], p, s_)
                            if r1 is None:
                                continue
                            sg, ns = r1
                            v = sum(_eri(prob, spa, ba, spa, ba, p, q, q, s_) for q in range(na))
                            nd = det[:si] + (ns,) + det[si + 1:]
                            add(nd, -0.5 * sg * v)
        for d, v in out.items():
            i = index.get(d)
            if i is not None:
                H[i, j] += v
    return H


def oracle_dets(oc):
    """The oracle's determinants as tuples of strings, in address order."""
    prob = oc.prob
    strs = [oc.strings(p, s) for p in range(prob.nspecies) for s in range(2)]
    U = oc.unfold_all()
    return [tuple(int(strs[k][U[i, k]]) for k in range(len(strs))) for i in range(oc.N)]


def s_squared(prob, dets, species, c):
    """<c| S^2 |c> of one species, S^2 = S- S+ + Sz (Sz + 1), with S+ = sum_i a+_{i alpha} a_{i beta}
    acting on the species' combined spin-orbital string (alpha orbitals 0..n-1, beta n..2n-1)."""
    n = prob.species[species].n_orb
    sa, sb = 2 * species, 2 * species + 1
    mask = (1 << n) - 1
    index = {d: i for i, d in enumerate(dets)}
    total = 0.0
    for j, det in enumerate(dets):
        if c[j] == 0.0:
            continue
        comb = det[sa] | (det[sb] << n)
        nal, nbe = bin(det[sa]).count("1"), bin(det[sb]).count("1")
        sz = 0.5 * (nal - nbe)
        total += c[j] * c[j] * sz * (sz + 1.0)
        for i in range(n):
            r1 = apply_E(comb, i, n + i)  # S+
            if r1 is None:
                continue
            g1, s1 = r1
            for k in range(n):
                r2 = apply_E(s1, n + k, k)  # S-
                if r2 is None:
                    continue
                g2, s2 = r2
                nd = list(det)
                nd[sa], nd[sb] = s2 & mask, s2 >> n
                t = index.get(tuple(nd))
                if t is not None:
                    total += g1 * g2 * c[t] * c[j]
    return total


def s_squared_matrix(prob, dets, species):
    """The matrix of S^2 = S- S+ + Sz (Sz + 1) of one species between the determinants of the
    space (operator algebra as in s_squared; intermediate states unrestricted)."""
    n = prob.species[species].n_orb
    sa, sb = 2 * species, 2 * species + 1
    mask = (1 << n) - 1
    index = {d: i for i, d in enumerate(dets)}
    M = np.zeros((len(dets), len(dets)))
    for j, det in enumerate(dets):
        comb = det[sa] | (det[sb] << n)
        nal, nbe = bin(det[sa]).count("1"), bin(det[sb]).count("1")
        sz = 0.5 * (nal - nbe)
        M[j, j] += sz * (sz + 1.0)
        for i in range(n):
            r1 = apply_E(comb, i, n + i)
            if r1 is None:
                continue
            g1, s1 = r1
            for k in range(n):
                r2 = apply_E(s1, n + k, k)
                if r2 is None:
                    continue
                g2, s2 = r2
                nd = list(det)
                nd[sa], nd[sb] = s2 & mask, s2 >> n
                t = index.get(tuple(nd))
                if t is not None:
                    M[t, j] += g1 * g2
    return M
