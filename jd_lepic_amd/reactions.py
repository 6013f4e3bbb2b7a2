"""Reaction tables in the reference's text format, and the analytic tables of
the synthetic workloads.

Host-side input code (no GPU work).  The reader follows the section grammar
of read_reactions.f90:112-431 (IONS / NEUTRALS particle sections, REACTION
blocks with label, ``Eth dE``, ``scale scale``, type, free text and a
``----``-delimited (E, sigma) table); surface-reaction sections are skipped.
Species numbering follows the reference: the electron is species 1, then the
particles in file order (read_reactions.f90:102-207); neutrals carry a
negative mass and density ni0*ngas (read_reactions.f90:200-205).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

QE = float(np.float32(1.60217646e-19))   # main.f90:120 (single-precision literal)
AMU = 1.66053886e-27                      # main.f90:124
ME = 9.10938188e-31                       # read_reactions.f90:105

# reaction type codes, read_reactions.f90:17-21
COLLISION, IONIZATION, EXCITATION, CHARGEEXCHANGE, DISSOCIATION = 1, 2, 3, 4, 5
_RTYPE_WORDS = {"COLLISION": 1, "IONIZATION": 2, "EXCITATION": 3, "CHARGEEXCHANGE": 4,
                "DISSOCIATION": 5}
_RTYPE_PREFIX = [("DATT", 2), ("DION", 2), ("DISS", 5), ("ELA", 1), ("EXC", 3), ("ION", 2),
                 ("VIBB", 3), ("VIBS", 3), ("CEX", 4), ("DRE", 5), ("RRE", 2)]


@dataclass
class Species:
    name: str
    mass: float          # kg (negative for neutrals)
    charge: float        # C
    T: float             # eV
    ni0: float           # n/n0 for charged, density (m^-3) for neutrals


@dataclass
class Reaction:
    label: str
    reactants: list      # 1-based species indices
    products: list
    rtype: int
    Eth: float
    dE: float
    E: np.ndarray        # eV
    S: np.ndarray        # m^2

    def col_info(self) -> list:
        """col_info(icol,1:10) of read_reactions.f90:11-22."""
        n_re, n_by = len(self.reactants), len(self.products)
        ci = [n_re, n_by] + list(self.reactants) + list(self.products) + [self.rtype]
        ci.append(sum(1 for p in self.products if p == 1))
        return (ci + [0] * 10)[:10]


@dataclass
class Chemistry:
    species: list = field(default_factory=list)    # index 0 is species 1 (the electron)
    reactions: list = field(default_factory=list)

    @property
    def ntype_all(self):
        return len(self.species)

    @property
    def ntype(self):
        return sum(1 for s in self.species if s.charge != 0.0)

    def index(self, name: str) -> int:
        for k, s in enumerate(self.species):
            if s.name == name:
                return k + 1
        raise KeyError(f"unknown particle {name} in reaction")

    def masses(self) -> dict:
        return {k + 1: s.mass for k, s in enumerate(self.species)}

    def tags(self) -> dict:
        """tag_neg, tag_neu, tag_b of read_reactions.f90:187-189: the last
        particle named [H-]/[D-], [H]/[D], [Hb]/[Db] (0 when absent)."""
        out = dict(tag_neg=0, tag_neu=0, tag_b=0)
        for k, sp in enumerate(self.species[1:], start=2):
            if sp.name in ("[H-]", "[D-]"):
                out["tag_neg"] = k
            if sp.name in ("[H]", "[D]"):
                out["tag_neu"] = k
            if sp.name in ("[Hb]", "[Db]"):
                out["tag_b"] = k
        return out


def _rtype(word: str) -> int:
    w = word.strip()
    if w.upper() in _RTYPE_WORDS and (w.isupper() or w.islower()):
        return _RTYPE_WORDS[w.upper()]
    for p, code in _RTYPE_PREFIX:
        if w[:len(p)].upper() == p and (w[:len(p)].isupper() or w[:len(p)].islower()):
            return code
    raise ValueError(f"unknown reaction type {word!r}")


def _floats(line: str):
    return [float(t.replace("d", "e").replace("D", "e")) for t in line.split()]


def parse(text: str, ngas: float) -> Chemistry:
    chem = Chemistry([Species("[e]", ME, -QE, 0.0, 1.0)], [])
    lines = text.splitlines()
    i = 0

    def skip_to_dashes(i):
        while i < len(lines) and not lines[i].strip().startswith("----"):
            i += 1
        return i + 1

    while i < len(lines):
        head = lines[i].strip()[:4]
        if head.upper() in ("IONS", "NEUT") and (head.isupper() or head.islower() or head[0].isupper()):
            i = skip_to_dashes(i + 1)
            while i < len(lines) and not lines[i].strip().startswith("--"):
                tok = lines[i].split()
                if tok:
                    mass, charge, T, ni0 = _floats(" ".join(tok[1:5]))
                    charge *= QE
                    mass *= AMU
                    if charge == 0.0:
                        ni0 *= ngas
                        mass = -abs(mass)
                    chem.species.append(Species(tok[0][:6], mass, charge, T, ni0))
                i += 1
            i += 1
            continue
        if head.upper() == "REAC":
            label = lines[i + 1].split()[0]
            lhs, rhs = label.split("->")
            names = lambda s: ["[" + t for t in s.split("[") if t]
            rea = [chem.index(n.strip("+")) for n in names(lhs)]
            pro = [chem.index(n.strip("+")) for n in names(rhs)]
            Eth, dE = _floats(lines[i + 2])[:2]
            sc = _floats(lines[i + 3])[:2]
            rt = _rtype(lines[i + 4].split()[0])
            i = skip_to_dashes(i + 5)
            E, S = [], []
            while i < len(lines) and not lines[i].strip().startswith("----"):
                v = _floats(lines[i])
                if len(v) >= 2:
                    E.append(v[0] * sc[0])
                    S.append(v[1] * sc[1])
                i += 1
            i += 1
            chem.reactions.append(Reaction(label, rea, pro, rt, Eth * sc[0], dE, np.array(E),
                                           np.array(S)))
            continue
        i += 1
    return chem


def write(chem: Chemistry, ngas: float) -> str:
    out = []
    ions = [s for s in chem.species[1:] if s.charge != 0.0]
    neus = [s for s in chem.species[1:] if s.charge == 0.0]
    for title, group in (("IONS", ions), ("NEUTRALS", neus)):
        if not group:
            continue
        out += [title, "name   mass  charge  temp.(eV)  n/n0", "-" * 37]
        for s in group:
            ni0 = s.ni0 / ngas if s.charge == 0.0 else s.ni0
            out.append(f"{s.name:<8s} {float(abs(s.mass) / AMU)!r} {float(s.charge / QE)!r} {float(s.T)!r} {float(ni0)!r}")
        out += ["-" * 37, ""]
    words = {1: "COLLISION", 2: "IONIZATION", 3: "EXCITATION", 4: "CHARGEEXCHANGE",
             5: "DISSOCIATION"}
    for r in chem.reactions:
        out += ["REACTION", r.label, f" {float(r.Eth)!r}  {float(r.dE)!r}", " 1.  1.", words[r.rtype],
                "-" * 60]
        out += [f"  {float(e)!r}  {float(s)!r}" for e, s in zip(r.E, r.S)]
        out += ["-" * 60, ""]
    return "\n".join(out) + "\n"


# ---------------------------------------------------------------------------
# analytic tables of the synthetic workloads (values are the builder's choice
# where the workload leaves them open, and are recorded in the written table)

def energy_grid(thresholds=(), emax=1.0e4, n=240):
    """0 eV, a geometric grid 1e-3..emax, and every threshold (so tables are
    exact at their onset)."""
    g = np.concatenate([[0.0], np.geomspace(1e-3, emax, n), np.asarray(thresholds, float)])
    return np.unique(g)


def lotz(E, I, q, a=4.0e-18, b=0.62, c=0.4):
    """Lotz (1968) single-shell ionization cross section in m^2 (a in m^2 eV^2)."""
    E = np.asarray(E, float)
    out = np.zeros_like(E)
    k = E > I
    u = E[k] / I
    out[k] = a * q * np.log(u) / (E[k] * I) * (1.0 - b * np.exp(-c * (u - 1.0)))
    return out


def excitation(E, Eth, s0):
    """Onset-limited excitation shape s0 * (1 - Eth/E) * ln(E/Eth) / (E/Eth)."""
    E = np.asarray(E, float)
    out = np.zeros_like(E)
    k = E > Eth
    u = E[k] / Eth
    out[k] = s0 * (1.0 - 1.0 / u) * np.log(u) / u * math.e
    return out


def argon(ngas: float, Ti_ion=0.03, Tgas=0.02585, sig_el=1.0e-19) -> Chemistry:
    """e-/Ar+ with background Ar: constant elastic, excitation at 11.5 eV and
    Lotz ionization at 15.76 eV (q=6, a=4.0e-14 cm^2 eV^2, b=0.62, c=0.4)."""
    chem = Chemistry([Species("[e]", ME, -QE, 0.0, 1.0),
                      Species("[Ar+]", 39.95 * AMU, QE, Ti_ion, 1.0),
                      Species("[Ar]", -39.95 * AMU, 0.0, Tgas, ngas)], [])
    E = energy_grid((11.5, 15.76))
    chem.reactions.append(Reaction("[e]+[Ar]->[e]+[Ar]", [1, 3], [1, 3], COLLISION, 0.0, 0.0, E,
                                   np.full_like(E, sig_el)))
    chem.reactions.append(Reaction("[e]+[Ar]->[e]+[Ar]", [1, 3], [1, 3], EXCITATION, 11.5, 0.0, E,
                                   excitation(E, 11.5, 1.0e-20)))
    chem.reactions.append(Reaction("[e]+[Ar]->[e]+[Ar+]+[e]", [1, 3], [1, 2, 1], IONIZATION, 15.76,
                                   0.0, E, lotz(E, 15.76, 6)))
    return chem


def hydrogen(ngas: float) -> Chemistry:
    """e-, H+, H2+, H- with background H2 (negative-ion-source workload):
    constant elastic, dissociation e+H2->e+H+H at 8.8 eV, and ionization
    e+H2->e+H2++e at 15.43 eV (Lotz, q=2)."""
    chem = Chemistry([Species("[e]", ME, -QE, 0.0, 1.0),
                      Species("[H+]", 1.00794 * AMU, QE, 0.1, 0.5),
                      Species("[H2+]", 2.01588 * AMU, QE, 0.1, 0.5),
                      Species("[H-]", 1.00794 * AMU, -QE, 0.1, 0.0),
                      Species("[H2]", -2.01588 * AMU, 0.0, 0.1, ngas),
                      Species("[H]", -1.00794 * AMU, 0.0, 0.8, 0.0)], [])
    E = energy_grid((8.8, 15.43))
    chem.reactions.append(Reaction("[e]+[H2]->[e]+[H2]", [1, 5], [1, 5], COLLISION, 0.0, 0.0, E,
                                   np.full_like(E, 1.0e-19)))
    chem.reactions.append(Reaction("[e]+[H2]->[e]+[H]+[H]", [1, 5], [1, 6, 6], DISSOCIATION, 8.8,
                                   0.0, E, excitation(E, 8.8, 9.0e-21)))
    chem.reactions.append(Reaction("[e]+[H2]->[e]+[H2+]+[e]", [1, 5], [1, 3, 1], IONIZATION, 15.43,
                                   0.0, E, lotz(E, 15.43, 2)))
    return chem
