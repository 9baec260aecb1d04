"""Seeded synthetic CI problems (host side, numpy only).

Every workload the tests and the benchmark use is generated here, so the HIP path and the
oracle see byte-identical integrals.  Randomness is SplitMix64 (Steele, Lea & Flood 2014)
turned into doubles with the top 53 bits and into normals with Box-Muller; the seed of
workload ``c`` is ``0x5EED0000 + c``.

Integral conventions follow the reference's storage (src/integral/integral.hpp:14-20,
src/integral/integral.cpp:479-516): one-body ``h[species][spin]`` is an n x n matrix and the
two-body block ``eri[a][sa][b][sb]`` (b >= a) is the n_pair_a x n_pair_b matrix
``(ij|kl)`` addressed by ``idx2(i, j), idx2(k, l)``.

Synthetic integrals (``cholesky_problem``):
  * h = G + diag(sorted U[-2, 3]) with G symmetric, G_ij ~ N(0, 0.5^2) for i <= j;
  * L_k (k < 4n) symmetric with entries ~ N(0, 0.1^2) for i <= j; (pq|rs) = sum_k L_k,pq L_k,rs;
  * the electron-positron block is +sum_k L^e_k,pq L^p_k,rs; the reference multiplies every
    mixed-species element by q_e * q_p = -1 (src/ci/detset/direct_mixedparticle.cpp:276-278),
    which gives the negated coupling.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

SEED_BASE = 0x5EED0000


class SplitMix64:
    """Counter-based SplitMix64 stream (vectorised draws, wrap-around uint64 arithmetic)."""

    def __init__(self, seed: int):
        self.seed = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.counter = 0

    def raw(self, n: int) -> np.ndarray:
        idx = np.arange(self.counter + 1, self.counter + 1 + n, dtype=np.uint64)
        self.counter += n
        with np.errstate(over="ignore"):
            z = self.seed + idx * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        """U[0, 1) doubles from the top 53 bits."""
        return (self.raw(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int, sigma: float = 1.0) -> np.ndarray:
        m = (n + 1) // 2
        u1 = 1.0 - self.uniform(m)  # (0, 1]
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)])
        return sigma * z[:n]

    def integers(self, lo: int, hi: int, n: int = 1) -> np.ndarray:
        """Uniform integers in [lo, hi] (inclusive)."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64)


def npair(n: int) -> int:
    return n * (n + 1) // 2


def idx2(i: int, j: int) -> int:
    """symmetric_matrix_triangular_idx (src/integral/integral.hpp:14-20)."""
    return i * (i + 1) // 2 + j if i > j else j * (j + 1) // 2 + i


def pair_index_matrix(n: int) -> np.ndarray:
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    hi = np.maximum(ii, jj)
    lo = np.minimum(ii, jj)
    return hi * (hi + 1) // 2 + lo


def sym_gaussian(rng: SplitMix64, n: int, sigma: float) -> np.ndarray:
    iu = np.triu_indices(n)
    v = rng.normal(len(iu[0]), sigma)
    m = np.zeros((n, n))
    m[iu] = v
    return m + np.triu(m, 1).T


def packed_cholesky(rng: SplitMix64, n: int, nvec: int, sigma: float) -> np.ndarray:
    """nvec x n_pair matrix of symmetric Cholesky-like vectors, packed with idx2."""
    P = pair_index_matrix(n)
    iu = np.triu_indices(n)
    out = np.zeros((nvec, npair(n)))
    for k in range(nvec):
        m = sym_gaussian(rng, n, sigma)
        out[k, P[iu]] = m[iu]
    return out


@dataclasses.dataclass
class Species:
    """One quantum-particle type of a CI problem (electron or positron)."""

    n_orb: int
    n_alpha: int
    n_beta: int
    charge: float
    nspin_blocks: int = 1
    ex_level: tuple = None  # (alpha, beta, total); None -> full CI

    def levels(self):
        if self.ex_level is not None:
            return tuple(self.ex_level)
        la = min(self.n_alpha, self.n_orb - self.n_alpha)
        lb = min(self.n_beta, self.n_orb - self.n_beta)
        return (la, lb, la + lb)


@dataclasses.dataclass
class CIProblem:
    """Integrals and particle counts of one CI solve, in the reference's storage layout."""

    species: List[Species]
    h: list  # h[p][s] -> (n, n) float64
    eri: dict  # (a, sa, b, sb) -> (npair_a, npair_b) float64, b >= a
    nroots: int = 1
    max_collective: int = 2**31 - 1
    e_shift: float = 0.0  # constant added to the energies (E_nuc + frozen core)
    name: str = ""

    @property
    def nspecies(self) -> int:
        return len(self.species)

    def n_strings(self) -> List[int]:
        return [math.comb(sp.n_orb, k) for sp in self.species for k in (sp.n_alpha, sp.n_beta)]

    def n_dets_fci(self) -> int:
        return int(np.prod(self.n_strings()))


def cholesky_problem(seed: int, species: List[Species], nroots: int, nvec_per_orb: int = 4,
                     name: str = "") -> CIProblem:
    """Seeded synthetic integrals with the recipe in the module docstring."""
    rng = SplitMix64(seed)
    hs, Ls = [], []
    for sp in species:
        n = sp.n_orb
        g = sym_gaussian(rng, n, 0.5)
        eps = np.sort(-2.0 + 5.0 * rng.uniform(n))
        h = g + np.diag(eps)
        hs.append([h.copy() for _ in range(sp.nspin_blocks)])
        Ls.append(packed_cholesky(rng, n, nvec_per_orb * n, 0.1))
    eri = {}
    for a, spa in enumerate(species):
        for b in range(a, len(species)):
            spb = species[b]
            nv = min(Ls[a].shape[0], Ls[b].shape[0])
            v = Ls[a][:nv].T @ Ls[b][:nv]
            for sa in range(spa.nspin_blocks):
                for sb in range(spb.nspin_blocks):
                    eri[(a, sa, b, sb)] = np.ascontiguousarray(v)
    return CIProblem(species=species, h=hs, eri=eri, nroots=nroots, name=name)


# --- the named workloads -------------------------------------------------------------

def config2(nroots: int = 3) -> CIProblem:
    """Electron + positron FCI: 48 orbitals each, 1a1b electrons, 1 positron (110,592 dets)."""
    sp = [Species(48, 1, 1, -1.0, 1), Species(48, 1, 0, 1.0, 1)]
    return cholesky_problem(SEED_BASE + 2, sp, nroots, name="ep_fci_48e_48p")


def config3(nroots: int = 4) -> CIProblem:
    """Electronic FCI, 16 orbitals, 8a8b electrons (12870^2 = 165,636,900 dets)."""
    sp = [Species(16, 8, 8, -1.0, 1)]
    return cholesky_problem(SEED_BASE + 3, sp, nroots, name="fci_16o_8a8b")


def config4(n: int = 320, nvec: int = 1280, sigma: float = 0.05):
    """AO->MO transform input: (L, C).

    L: nvec x n_pair symmetric Cholesky-like vectors (idx2-packed), entries N(0, sigma^2), so
    that (mu nu|la si) = sum_k L[k, idx2(mu,nu)] L[k, idx2(la,si)].  C: Q factor of a Householder
    QR of an n x n Gaussian with diag(R) > 0 (column-major MO coefficients).
    """
    rng = SplitMix64(SEED_BASE + 4)
    L = packed_cholesky(rng, n, nvec, sigma)
    G = rng.normal(n * n).reshape(n, n)
    Q, R = np.linalg.qr(G)
    Q = Q * np.sign(np.diag(R))[None, :]
    return L, np.asfortranarray(Q)


def water_problem(basis: str = "6-31g", nroots: int = 1, ex_level=None) -> CIProblem:
    """Water (O-H 0.9572 A, HOH 104.52 deg), RHF orbitals on the CPU, all orbitals active.

    basis "6-31g" with ex_level None is the full CI of 13 orbitals, 5a5b (1,656,369 dets);
    "sto-3g" with ex_level (2, 2, 2) is the reference's H2O/STO-3G CISD case
    (tests/data/h2o_sto3glibrary_cisd/h2o.json).  e_shift = nuclear repulsion.
    """
    from . import molint

    w = molint.water(basis)
    n = w["n_orb"]
    P = pair_index_matrix(n)
    ii, jj = np.tril_indices(n)
    pairs = P[ii, jj]
    order = np.argsort(pairs)
    ii, jj = ii[order], jj[order]
    v = w["eri_mo"][ii[:, None], jj[:, None], ii[None, :], jj[None, :]]
    sp = [Species(n, w["n_alpha"], w["n_beta"], -1.0, 1, ex_level)]
    prob = CIProblem(species=sp, h=[[w["h_mo"].copy()]], eri={(0, 0, 0, 0): np.ascontiguousarray(v)},
                     nroots=nroots, e_shift=w["E_nuc"], name=f"h2o_{basis}")
    prob.e_rhf = w["E_rhf"]
    return prob


def small_problem(seed: int, n: int, na: int, nb: int, nroots: int = 1, nspin: int = 1,
                  positron: Optional[tuple] = None, ex_level=None, pos_ex_level=None,
                  max_collective: int = 2**31 - 1) -> CIProblem:
    """Seeded small problem for parity tests.  positron = (n_p, n_alpha_p, n_beta_p)."""
    sp = [Species(n, na, nb, -1.0, nspin, ex_level)]
    if positron is not None:
        npo, pa, pb = positron
        sp.append(Species(npo, pa, pb, 1.0, 1, pos_ex_level))
    prob = cholesky_problem(seed, sp, nroots)
    if nspin == 2:
        # distinct beta one-body and two-body blocks so spin-resolved paths are exercised
        rng = SplitMix64(seed ^ 0xABCDEF)
        hb = prob.h[0][0] + sym_gaussian(rng, n, 0.05)
        prob.h[0][1] = hb
        Lb = packed_cholesky(rng, n, 4 * n, 0.1) * 0.3 + 0.0
        La = packed_cholesky(SplitMix64(seed ^ 0x123457), n, 4 * n, 0.1)
        La = La * 0.3
        base = prob.eri[(0, 0, 0, 0)]
        prob.eri[(0, 0, 0, 0)] = base + La.T @ La
        prob.eri[(0, 1, 0, 1)] = base + Lb.T @ Lb
        prob.eri[(0, 0, 0, 1)] = base + La.T @ Lb
        prob.eri[(0, 1, 0, 0)] = np.ascontiguousarray((base + La.T @ Lb).T)
        if positron is not None:
            v = prob.eri[(0, 0, 1, 0)]
            prob.eri[(0, 1, 1, 0)] = v * 1.1
    prob.max_collective = max_collective
    return prob


def serving_requests(n_requests: int = 512, rate: float = 200.0, seed: int = SEED_BASE + 5, only=None):
    """Request mix of the serving workload: Poisson arrivals and uneven problem sizes.

    Returns a list of (arrival_time_s, CIProblem).  Per request: n ~ U{6..14},
    alpha, beta ~ U{2..n//2} (independent), positron in {0, 1} with n_p ~ U{6..12},
    roots ~ U{1..4}; integrals seeded per request with the Cholesky recipe.  ``only``: a
    collection of request numbers whose integrals are built (the others come back as None).
    """
    rng = SplitMix64(seed)
    t = 0.0
    out = []
    for r in range(n_requests):
        t += -math.log(1.0 - rng.uniform(1)[0]) / rate
        n = int(rng.integers(6, 14)[0])
        na = int(rng.integers(2, n // 2)[0])
        nb = int(rng.integers(2, n // 2)[0])
        has_pos = int(rng.integers(0, 1)[0])
        roots = int(rng.integers(1, 4)[0])
        sp = [Species(n, na, nb, -1.0, 1)]
        if has_pos:
            npo = int(rng.integers(6, 12)[0])
            sp.append(Species(npo, 1, 0, 1.0, 1))
        if only is not None and r not in only:
            out.append((t, None))
            continue
        prob = cholesky_problem(seed * 1000003 + r, sp, roots, name=f"req{r}")
        out.append((t, prob))
    return out
