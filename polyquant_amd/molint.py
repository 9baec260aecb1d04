"""Gaussian integrals over s and p shells and restricted Hartree-Fock, on the CPU.

The CI path takes MO integrals as input; the self-consistent field that produces them stays on
the CPU (as in the reference, whose SCF feeds POLYQUANT_EPCI through POLYQUANT_INTEGRAL).  This
module builds those inputs for the water workloads: one-electron (overlap, kinetic, nuclear
attraction) and two-electron repulsion integrals by the McMurchie-Davidson scheme (Hermite
expansion coefficients E_t^{ij} and Hermite Coulomb integrals R_{tuv}; Helgaker, Jorgensen and
Olsen, "Molecular Electronic-Structure Theory", ch. 9), then closed-shell RHF with DIIS.

Conventions kept from the reference: Angstrom -> Bohr with bohr_to_angstrom = 0.529177210903
(src/molecule/molecule.hpp:111); alpha = ceil(N/2), beta = floor(N/2) electrons
(src/molecule/molecule.cpp:304-321); basis data in the Gaussian94 layout of the library files
(S / SP shells with contraction coefficients for normalised primitives).
Only angular momentum <= 1 is supported (s, p: enough for STO-3G and 6-31G on H..Ne).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Sequence, Tuple

import numpy as np
from scipy.special import gammainc, gamma

BOHR_TO_ANGSTROM = 0.529177210903

# Basis library data (exponent, s coefficient[, p coefficient]) per shell; published STO-3G and
# 6-31G values (Basis Set Exchange, Gaussian94 format).
BASIS: Dict[str, Dict[str, List[Tuple[str, List[Tuple[float, ...]]]]]] = {
    "sto-3g": {
        "H": [("S", [(3.425250914, 0.1543289673), (0.6239137298, 0.5353281423), (0.1688554040, 0.4446345422)])],
        "O": [
            ("S", [(130.7093214, 0.1543289673), (23.80886605, 0.5353281423), (6.443608313, 0.4446345422)]),
            ("SP", [(5.033151319, -0.09996722919, 0.1559162750), (1.169596125, 0.3995128261, 0.6076837186),
                    (0.3803889600, 0.7001154689, 0.3919573931)]),
        ],
    },
    "6-31g": {
        "H": [
            ("S", [(18.73113696, 0.03349460434), (2.825394365, 0.2347269535), (0.6401216923, 0.8137573261)]),
            ("S", [(0.1612777588, 1.0)]),
        ],
        "O": [
            ("S", [(5484.671660, 0.001831074430), (825.2349460, 0.01395017220), (188.0469580, 0.06844507810),
                   (52.96450000, 0.2327143360), (16.89757040, 0.4701928980), (5.799635340, 0.3585208530)]),
            ("SP", [(15.53961625, -0.1107775495, 0.07087426823), (3.599933586, -0.1480262627, 0.3397528391),
                    (1.013761750, 1.130767015, 0.7271585773)]),
            ("SP", [(0.2700058226, 1.0, 1.0)]),
        ],
    },
}

# Custom basis sets of the reference's electron + positron test systems, same layout:
#   "psh_electron" / "psh_positron": tests/data/PsH_wpos/electron_basis.g94, positron_basis.g94
#   (the H atom of the PsH_wpos inputs);
#   "li_s": the s shells of tests/data/li-_custombasis_wpos/Li_basis.g94 (its p shells left out).
BASIS["psh_electron"] = {"H": [("S", [(17.37, 0.032828), (2.6173, 0.20121), (0.58994, 0.81724)]),
                               ("S", [(0.169029, 1.0)]), ("S", [(0.06, 1.0)]), ("S", [(0.02, 1.0)]),
                               ("S", [(0.009, 1.0)])]}
BASIS["psh_positron"] = {"H": [("S", [(0.02, 1.0)]), ("S", [(0.009, 1.0)])]}
BASIS["li_s"] = {"Li": [("S", [(642.4189, 0.00214), (96.79849, 0.01621), (22.09109, 0.07732), (6.20107, 0.24579),
                               (1.93512, 0.47019)]),
                        ("S", [(2.19146, 0.03509), (0.59613, 0.19123)]), ("S", [(0.07455, 1.0)]),
                        ("S", [(0.02867, 1.0)]), ("S", [(0.0092, 1.0)]), ("S", [(0.0031, 1.0)]),
                        ("S", [(0.0008, 1.0)])]}

ZNUC = {"H": 1, "Li": 3, "O": 8}


@dataclasses.dataclass
class BasisFunction:
    center: np.ndarray        # Bohr
    lmn: Tuple[int, int, int]  # Cartesian powers (l <= 1)
    exps: np.ndarray
    coefs: np.ndarray          # contraction coefficients including primitive normalisation


def _prim_norm(a: float, lmn) -> float:
    L = sum(lmn)
    df = 1.0  # (2l-1)!! products, l <= 1 -> 1
    return (2 * a / math.pi) ** 0.75 * (4 * a) ** (L / 2) / math.sqrt(df)


def build_basis(symbols: Sequence[str], coords_bohr: np.ndarray, name: str) -> List[BasisFunction]:
    """Basis functions in the order atom -> shell -> (s, px, py, pz)."""
    out: List[BasisFunction] = []
    for sym, R in zip(symbols, coords_bohr):
        for kind, prims in BASIS[name][sym]:
            exps = np.array([p[0] for p in prims])
            parts = [("S", 1)] if kind == "S" else [("S", 1), ("P", 2)]
            for shell, col in parts:
                cs = np.array([p[col] for p in prims])
                lmns = [(0, 0, 0)] if shell == "S" else [(1, 0, 0), (0, 1, 0), (0, 0, 1)]
                for lmn in lmns:
                    c = cs * np.array([_prim_norm(a, lmn) for a in exps])
                    bf = BasisFunction(np.array(R, dtype=float), lmn, exps, c)
                    # normalise the contracted function
                    s = _overlap_bf(bf, bf)
                    bf.coefs = c / math.sqrt(s)
                    out.append(bf)
    return out


# --- McMurchie-Davidson building blocks ------------------------------------------------------
def _E(i: int, j: int, t: int, Qx, a, b):
    """Hermite expansion coefficient E_t^{ij} (arrays over primitive pairs)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return np.zeros_like(Qx)
    if i == j == t == 0:
        return np.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p)) * _E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * _E(i - 1, j, t, Qx, a, b) + \
            (t + 1) * _E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * _E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * _E(i, j - 1, t, Qx, a, b) + \
        (t + 1) * _E(i, j - 1, t + 1, Qx, a, b)


def _boys(n: int, T):
    T = np.asarray(T, dtype=float)
    out = np.empty_like(T)
    small = T < 1e-10
    out[small] = 1.0 / (2 * n + 1) - T[small] / (2 * n + 3)
    Ts = T[~small]
    out[~small] = gamma(n + 0.5) * gammainc(n + 0.5, Ts) / (2 * Ts ** (n + 0.5))
    return out


def _R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    """Hermite Coulomb integral R^n_{tuv}."""
    if t < 0 or u < 0 or v < 0:
        return np.zeros_like(RPC)
    if t == u == v == 0:
        return (-2 * p) ** n * _boys(n, p * RPC * RPC)
    if t > 0:
        return (t - 1) * _R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC) + PCx * _R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    if u > 0:
        return (u - 1) * _R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC) + PCy * _R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
    return (v - 1) * _R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC) + PCz * _R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)


def _pair(A: BasisFunction, B: BasisFunction):
    a = A.exps[:, None]
    b = B.exps[None, :]
    p = a + b
    P = (a[..., None] * A.center + b[..., None] * B.center) / p[..., None]
    cc = A.coefs[:, None] * B.coefs[None, :]
    return a, b, p, P, cc


def _overlap_bf(A: BasisFunction, B: BasisFunction) -> float:
    a, b, p, P, cc = _pair(A, B)
    s = np.ones_like(p)
    for d in range(3):
        s = s * _E(A.lmn[d], B.lmn[d], 0, A.center[d] - B.center[d], a, b)
    return float(np.sum(cc * s * (math.pi / p) ** 1.5))


def _kinetic_bf(A: BasisFunction, B: BasisFunction) -> float:
    a, b, p, P, cc = _pair(A, B)
    X = A.center - B.center
    l1 = A.lmn
    l2 = B.lmn

    def S1(d, i, j):
        return _E(i, j, 0, X[d], a, b) * np.sqrt(math.pi / p)

    tot = 0.0
    for d in range(3):
        j = l2[d]
        # T_ij = -2 b^2 S_{i,j+2} + b (2j + 1) S_ij - 1/2 j (j - 1) S_{i,j-2}
        Tij = -2 * b * b * S1(d, l1[d], j + 2) + b * (2 * j + 1) * S1(d, l1[d], j)
        if j >= 2:
            Tij = Tij - 0.5 * j * (j - 1) * S1(d, l1[d], j - 2)
        term = Tij
        for e in range(3):
            if e != d:
                term = term * S1(e, l1[e], l2[e])
        tot += float(np.sum(cc * term))
    return tot


def _nuclear_bf(A: BasisFunction, B: BasisFunction, atoms: List[Tuple[float, np.ndarray]]) -> float:
    a, b, p, P, cc = _pair(A, B)
    X = A.center - B.center
    tot = 0.0
    for Z, C in atoms:
        PC = P - C
        RPC = np.linalg.norm(PC, axis=-1)
        acc = np.zeros_like(p)
        for t in range(A.lmn[0] + B.lmn[0] + 1):
            Et = _E(A.lmn[0], B.lmn[0], t, X[0], a, b)
            for u in range(A.lmn[1] + B.lmn[1] + 1):
                Eu = _E(A.lmn[1], B.lmn[1], u, X[1], a, b)
                for v in range(A.lmn[2] + B.lmn[2] + 1):
                    Ev = _E(A.lmn[2], B.lmn[2], v, X[2], a, b)
                    acc = acc + Et * Eu * Ev * _R(t, u, v, 0, p, PC[..., 0], PC[..., 1], PC[..., 2], RPC)
        tot += -Z * float(np.sum(cc * 2 * math.pi / p * acc))
    return tot


def _eri_bf(A, B, C, D) -> float:
    a, b, p, P, cab = _pair(A, B)
    c, d, q, Q, ccd = _pair(C, D)
    # broadcast (a, b) x (c, d)
    sh1 = p.shape + (1, 1)
    sh2 = (1, 1) + q.shape
    pp = p.reshape(sh1)
    qq = q.reshape(sh2)
    alpha = pp * qq / (pp + qq)
    PQ = P.reshape(sh1 + (3,)) - Q.reshape(sh2 + (3,))
    RPQ = np.linalg.norm(PQ, axis=-1)
    XAB = A.center - B.center
    XCD = C.center - D.center
    Eab = [[_E(A.lmn[k], B.lmn[k], t, XAB[k], a, b).reshape(sh1) for t in range(A.lmn[k] + B.lmn[k] + 1)]
           for k in range(3)]
    Ecd = [[_E(C.lmn[k], D.lmn[k], t, XCD[k], c, d).reshape(sh2) for t in range(C.lmn[k] + D.lmn[k] + 1)]
           for k in range(3)]
    acc = 0.0
    for t, Et in enumerate(Eab[0]):
        for u, Eu in enumerate(Eab[1]):
            for v, Ev in enumerate(Eab[2]):
                e1 = Et * Eu * Ev
                for tau, Etau in enumerate(Ecd[0]):
                    for nu, Enu in enumerate(Ecd[1]):
                        for phi, Ephi in enumerate(Ecd[2]):
                            sign = -1.0 if (tau + nu + phi) % 2 else 1.0
                            acc = acc + sign * e1 * Etau * Enu * Ephi * _R(
                                t + tau, u + nu, v + phi, 0, alpha, PQ[..., 0], PQ[..., 1], PQ[..., 2], RPQ)
    pref = 2 * math.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq))
    return float(np.sum(cab.reshape(sh1) * ccd.reshape(sh2) * pref * acc))


def ao_integrals(basis: List[BasisFunction], atoms: List[Tuple[float, np.ndarray]]):
    """(S, T, V, eri) in the AO basis; eri is the full (n, n, n, n) chemists' tensor."""
    n = len(basis)
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = _overlap_bf(basis[i], basis[j])
            T[i, j] = T[j, i] = _kinetic_bf(basis[i], basis[j])
            V[i, j] = V[j, i] = _nuclear_bf(basis[i], basis[j], atoms)
    eri = np.zeros((n, n, n, n))
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    for ij, (i, j) in enumerate(pairs):
        for kl, (k, l) in enumerate(pairs[: ij + 1]):
            v = _eri_bf(basis[i], basis[j], basis[k], basis[l])
            for (x, y) in ((i, j), (j, i)):
                for (z, w) in ((k, l), (l, k)):
                    eri[x, y, z, w] = v
                    eri[z, w, x, y] = v
    return S, T, V, eri


def one_body_integrals(basis: List[BasisFunction], atoms: List[Tuple[float, np.ndarray]]):
    """(S, T, V) in the AO basis; V is the attraction of a charge -1 to the nuclei (negative)."""
    n = len(basis)
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1):
            S[i, j] = S[j, i] = _overlap_bf(basis[i], basis[j])
            T[i, j] = T[j, i] = _kinetic_bf(basis[i], basis[j])
            V[i, j] = V[j, i] = _nuclear_bf(basis[i], basis[j], atoms)
    return S, T, V


def eri_two_bases(basis_a: List[BasisFunction], basis_b: List[BasisFunction]) -> np.ndarray:
    """(pq|rs) with p, q in basis a and r, s in basis b: the (na, na, nb, nb) Coulomb tensor between
    two particle species (no charge factor)."""
    na, nb = len(basis_a), len(basis_b)
    out = np.zeros((na, na, nb, nb))
    for i in range(na):
        for j in range(i + 1):
            for k in range(nb):
                for l in range(k + 1):
                    v = _eri_bf(basis_a[i], basis_a[j], basis_b[k], basis_b[l])
                    out[i, j, k, l] = out[j, i, k, l] = out[i, j, l, k] = out[j, i, l, k] = v
    return out


def electron_positron_atom(symbol: str, electron_basis: str, positron_basis: str):
    """AO integrals of one atom at the origin with an electron and a positron basis (the reference's
    PsH_wpos and Li_wpos test systems): per species S, T, V (V: attraction of a charge -1), the
    electron-electron and positron-positron tensors and the electron-positron tensor
    (ne, ne, np, np)."""
    R = np.zeros((1, 3))
    atoms = [(float(ZNUC[symbol]), R[0])]
    be = build_basis([symbol], R, electron_basis)
    bp = build_basis([symbol], R, positron_basis)
    Se, Te, Ve = one_body_integrals(be, atoms)
    Sp, Tp, Vp = one_body_integrals(bp, atoms)
    return {"S": [Se, Sp], "T": [Te, Tp], "V": [Ve, Vp], "eri_ee": eri_two_bases(be, be),
            "eri_pp": eri_two_bases(bp, bp), "eri_ep": eri_two_bases(be, bp), "Z": ZNUC[symbol]}


def nuclear_repulsion(atoms: List[Tuple[float, np.ndarray]]) -> float:
    e = 0.0
    for i in range(len(atoms)):
        for j in range(i):
            e += atoms[i][0] * atoms[j][0] / np.linalg.norm(atoms[i][1] - atoms[j][1])
    return e


def rhf(S, Hcore, eri, nocc: int, e_tol: float = 1e-11, d_tol: float = 1e-9, maxit: int = 200):
    """Closed-shell RHF with DIIS from the core guess; returns (E_elec, C, eps)."""
    X = _inv_sqrt(S)
    eps, Cp = np.linalg.eigh(X.T @ Hcore @ X)
    C = X @ Cp
    D = C[:, :nocc] @ C[:, :nocc].T
    E_old = 0.0
    focks, errs = [], []
    for _ in range(maxit):
        J = np.einsum("pqrs,rs->pq", eri, D)
        K = np.einsum("prqs,rs->pq", eri, D)
        F = Hcore + 2 * J - K
        E = float(np.sum(D * (Hcore + F)))
        err = F @ D @ S - S @ D @ F
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        if len(focks) >= 2:
            m = len(focks)
            Bm = -np.ones((m + 1, m + 1))
            Bm[m, m] = 0.0
            for i in range(m):
                for j in range(m):
                    Bm[i, j] = np.sum(errs[i] * errs[j])
            rhs = np.zeros(m + 1)
            rhs[m] = -1.0
            w = np.linalg.solve(Bm, rhs)[:m]
            F = sum(wi * Fi for wi, Fi in zip(w, focks))
        eps, Cp = np.linalg.eigh(X.T @ F @ X)
        C = X @ Cp
        Dn = C[:, :nocc] @ C[:, :nocc].T
        if abs(E - E_old) < e_tol and np.max(np.abs(Dn - D)) < d_tol:
            D = Dn
            break
        D, E_old = Dn, E
    J = np.einsum("pqrs,rs->pq", eri, D)
    K = np.einsum("prqs,rs->pq", eri, D)
    F = Hcore + 2 * J - K
    E = float(np.sum(D * (Hcore + F)))
    eps, Cp = np.linalg.eigh(X.T @ F @ X)
    C = X @ Cp
    return E, C, eps


def _inv_sqrt(S):
    w, U = np.linalg.eigh(S)
    return U @ np.diag(w ** -0.5) @ U.T


WATER_GEOMETRY_ANGSTROM = np.array([[0.0, 0.0, 0.0], [0.7569685, 0.0, -0.5858752], [-0.7569685, 0.0, -0.5858752]])


def water(basis_name: str):
    """RHF water at O-H 0.9572 A, HOH 104.52 deg (the reference's h2o.json geometry).

    Returns dict with E_nuc, E_rhf, C, h_mo (n x n), eri_mo (n^4, chemists' notation) and the AO
    kinetic T, electron-nuclear V and eri_ao (n^4) integrals.
    """
    symbols = ["O", "H", "H"]
    coords = WATER_GEOMETRY_ANGSTROM / BOHR_TO_ANGSTROM
    atoms = [(float(ZNUC[s]), c) for s, c in zip(symbols, coords)]
    basis = build_basis(symbols, coords, basis_name)
    S, T, V, eri = ao_integrals(basis, atoms)
    Hcore = T + V
    nel = sum(ZNUC[s] for s in symbols)
    E_elec, C, eps = rhf(S, Hcore, eri, nel // 2)
    Enuc = nuclear_repulsion(atoms)
    h_mo = C.T @ Hcore @ C
    eri_mo = np.einsum("pi,qj,rk,sl,pqrs->ijkl", C, C, C, C, eri, optimize=True)
    return {"E_nuc": Enuc, "E_rhf": E_elec + Enuc, "C": C, "eps": eps, "h_mo": h_mo, "eri_mo": eri_mo,
            "n_orb": len(basis), "n_alpha": (nel + 1) // 2, "n_beta": nel // 2,
            "T": T, "V": V, "eri_ao": eri}
