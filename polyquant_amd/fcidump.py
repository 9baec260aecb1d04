"""FCIDUMP files: the writer of POLYQUANT_EPCI::fcidump and a reader that turns a file back into
a CI problem for the GPU path (a request format for ``DetSet`` and the serving loop).

Writer (``write_fcidump``) mirrors src/ci/epci.cpp:528-573 and src/io/fcidump_utilities.cpp:17-152
line for line in what it prints:
  * header only for a same-species file (fcidump_utilities.cpp:27-29, 40-61), with the
    reference's spelling ``M2S``, 2 x NORB orbital symmetry labels of 1 (epci.cpp:550-551),
    ``IUHF`` = 0/1 and ``PNT_GRP``;
  * the constant line ``E 0 0 0 0``, the orbital-energy lines ``e_i i 0 0 0`` (alpha, then beta
    when unrestricted), the lower triangle of the one-body matrix and the 4-fold two-body list
    (i >= j, a >= b over all pairs), each line ``{: >25.15f}`` then four ``{:>10d}`` indices
    (1-based), the same loops including the repetition of a one-body block per spin_b when
    unrestricted (fcidump_utilities.cpp:63-83).
A pair file (two species) holds only the two-body block (fcidump_utilities.cpp:64-66, 125-127).

Reader (``read_fcidump``) takes a same-species file in this layout (or any FCIDUMP whose two-body
list is 4- or 8-fold) and returns a ``CIProblem`` with ``e_shift`` = the constant line.
"""
from __future__ import annotations

import re
from typing import Optional, Sequence

import numpy as np

from .workloads import CIProblem, Species, idx2


def _line(v: float, i: int, j: int, k: int, l: int) -> str:
    return f"{v: >25.15f}{i:>10d}{j:>10d}{k:>10d}{l:>10d}\n"


def _setw(s, w: int) -> str:
    """std::setw(w) << s (right-aligned, never truncated)."""
    s = str(s)
    return s.rjust(w)


def header(num_mo: int, num_part_total: int, ms2: int, restricted: bool, orbsym: Sequence[int], isym: int,
           point_group: str) -> str:
    """dump_header (fcidump_utilities.cpp:40-61)."""
    out = (_setw(" &FCI  NORB= ", 10) + _setw(num_mo, 5) + _setw(", NELEC= ", 9) + _setw(num_part_total, 4)
           + _setw(", M2S= ", 5) + _setw(ms2, 4) + ", \n")
    out += _setw("ORBSYM= ", 10) + "".join(_setw(x, 3) + ", " for x in orbsym) + "\n"
    out += _setw("ISYM= ", 10) + f"{isym},\n"
    out += _setw("IUHF= ", 10) + f"{0 if restricted else 1},\n"
    out += _setw("PNT_GRP= ", 10) + _setw(point_group, 4) + ",\n"
    out += _setw(" &END", 3) + "\n"
    return out


def write_fcidump(path: str, prob: CIProblem, part_a: int = 0, part_b: Optional[int] = None,
                  e_constant: float = 0.0, mo_energies: Optional[Sequence[np.ndarray]] = None,
                  isym: int = 1, point_group: str = "C1") -> str:
    """POLYQUANT_FCIDUMP::dump for species part_a (and part_b for a pair file).  mo_energies:
    orbital energies per spin of part_a (zeros when the problem carries none).  Returns the text
    written."""
    part_b = part_a if part_b is None else part_b
    sa = prob.species[part_a]
    n = sa.n_orb
    num_part_total = sa.n_alpha + sa.n_beta
    ms2 = sa.n_alpha - sa.n_beta  # multiplicity - 1 of the high-spin state (epci.cpp:543)
    unique_beta = num_part_total > 1 and sa.nspin_blocks == 2
    restricted = not unique_beta
    spin_types = 1 if restricted else 2
    if mo_energies is None:
        mo_energies = [np.zeros(n), np.zeros(n)]
    out = []
    same = part_a == part_b
    if same:
        out.append(header(n, num_part_total, ms2, restricted, [1] * (2 * n), isym, point_group))
        out.append(_line(e_constant, 0, 0, 0, 0))
        for i in range(n):
            out.append(_line(float(mo_energies[0][i]), i + 1, 0, 0, 0))
        if spin_types == 2:
            for i in range(n):
                out.append(_line(float(mo_energies[1][i]), i + 1, 0, 0, 0))
        for spin_a in range(spin_types):
            for _spin_b in range(spin_types):
                h = prob.h[part_a][spin_a % sa.nspin_blocks]
                for i in range(n):
                    for a in range(i + 1):
                        out.append(_line(float(h[i, a]), i + 1, a + 1, 0, 0))
    sb = prob.species[part_b]
    for spin_a in range(spin_types):
        for spin_b in range(spin_types):
            V = prob.eri[(part_a, spin_a % sa.nspin_blocks, part_b, spin_b % sb.nspin_blocks)]
            na, nb = sa.n_orb, sb.n_orb
            for i in range(na):
                for j in range(i + 1):
                    row = V[idx2(i, j)]
                    for a in range(nb):
                        for b in range(a + 1):
                            out.append(_line(float(row[idx2(a, b)]), i + 1, j + 1, a + 1, b + 1))
    text = "".join(out)
    with open(path, "w") as f:
        f.write(text)
    return text


def parse_header(text: str) -> dict:
    """Key -> list of value tokens of the namelist header (&FCI ... &END or /).  Numeric lists
    (NORB, NELEC, MS2/M2S, ORBSYM, ISYM, IUHF) are what the reader uses."""
    body = text.split("&FCI", 1)[1] if "&FCI" in text else text
    out = {}
    for key, val in re.findall(r"([A-Za-z_][A-Za-z_0-9]*)\s*=\s*([^A-Za-z&/=]*)", body):
        out[key.upper()] = [v for v in re.split(r"[,\s]+", val.strip()) if v]
    return out


def read_fcidump(path: str, nroots: int = 1, charge: float = -1.0) -> CIProblem:
    """Same-species FCIDUMP -> CIProblem (one species; unrestricted when IUHF = 1)."""
    with open(path) as f:
        text = f.read()
    end = re.search(r"(&END|/)\s*\n", text)
    if not end:
        raise ValueError("FCIDUMP header has no &END")
    hdr = parse_header(text[:end.start()])
    n = int(hdr["NORB"][0])
    nelec = int(hdr["NELEC"][0])
    ms2 = int((hdr.get("MS2") or hdr.get("M2S") or ["0"])[0])
    uhf = int(hdr.get("IUHF", ["0"])[0]) != 0
    na, nb = (nelec + ms2) // 2, (nelec - ms2) // 2
    if na + nb != nelec or na < 0 or nb < 0:
        raise ValueError(f"inconsistent NELEC={nelec} MS2={ms2}")
    rows = np.loadtxt(text[end.end():].splitlines(), ndmin=2)
    vals, ijkl = rows[:, 0], rows[:, 1:5].astype(np.int64)
    npr = n * (n + 1) // 2
    spins = 2 if uhf else 1
    h = [np.zeros((n, n)) for _ in range(spins)]
    eri = {(0, a, 0, b): np.zeros((npr, npr)) for a in range(spins) for b in range(spins)}
    e_const = 0.0
    # Same-spin-pattern layout of the writer: constant, orbital energies, one-body blocks, then
    # two-body blocks in (spin_a, spin_b) order.  Lines are assigned by kind and, when
    # unrestricted, by their position within the kind.
    one_count, two_count = 0, 0
    one_block = npr  # lines per one-body block
    two_block = npr * npr
    for v, (i, j, k, l) in zip(vals, ijkl):
        if i == 0 and j == 0 and k == 0 and l == 0:
            e_const = float(v)
        elif j == 0 and k == 0 and l == 0:
            continue  # orbital energy
        elif k == 0 and l == 0:
            blk = (one_count // one_block) if uhf else 0
            s = min(blk // 2, 1) if uhf else 0  # the writer repeats each spin's block once per spin_b
            h[s][i - 1, j - 1] = v
            h[s][j - 1, i - 1] = v
            one_count += 1
        else:
            blk = (two_count // two_block) if uhf else 0
            sa, sb = (blk // 2, blk % 2) if uhf else (0, 0)
            p, q = idx2(i - 1, j - 1), idx2(k - 1, l - 1)
            eri[(0, sa, 0, sb)][p, q] = v
            if not uhf or sa == sb:
                eri[(0, sa, 0, sb)][q, p] = v
            two_count += 1
    species = [Species(n, na, nb, charge, 2 if uhf else 1)]
    return CIProblem(species=species, h=[h], eri=eri, nroots=nroots, e_shift=e_const,
                     name=f"fcidump:{path}")
