"""Synthetic workloads of the hot path (host-side set-up, no GPU work).

Each builder returns a dict describing one plasma: grid, walls (bcnd labels
as generate_boundary.f90 draws them), the Gaussian Bz filter field
(Bfield.f90:54-57), the species with their uniform-load strips, the
background gas chemistry (reactions.argon / reactions.hydrogen), the time
step and the MG depth (log2 of the smaller cell count, main.f90:264-267).
``build_context`` turns one into a loaded context through the C ABI: the
device loader (load_part.f90 restated with a counter-based generator), the
reaction tables, and a first deposition.

Named configurations:
  argon-128   128 x 128 cells, e-/Ar+ at 200 ppc each (6,553,600 particles)
  argon-512   512 x 512 cells, 100 ppc (52,428,800)
  ion-source  1024 x 512 cells, e-/H+/H2+/H- with an exponential density
              profile and a grid electrode (about 80 M)
  argon-1024  1024 x 1024 cells, 100 ppc (209,715,200)
  argon-2048  2048 x 2048 cells, 120 ppc (1,006,632,960)
"""
from __future__ import annotations

import numpy as np

from . import reactions as RX

KB = 1.380649e-23
H_CELL = 5.0e-5            # dx = dy (m)
DT = 5.0e-12               # s
P_GAS, T_GAS = 0.3, 300.0  # Pa, K


def _bfilter(n1, n2, h, peak=0.01):
    """Bz(x) = peak * exp(-(x - L/2)^2 / (2 (L/8)^2)) on the nodes x = (ix-1) h."""
    L1 = n1 * h
    x = (np.arange(n1 + 3) - 1) * h
    Bi = np.zeros((n2 + 3, n1 + 3, 4))
    Bi[..., 2] = peak * np.exp(-(x - L1 / 2) ** 2 / (2 * (L1 / 8) ** 2))[None, :]
    Bi[..., 3] = np.abs(Bi[..., 2])
    return Bi


def argon_box(n=512, ppc=100, Te=3.0, Ti=0.03, n0=1.0e16, Bpeak=0.01):
    """e-/Ar+ in a grounded n x n box at ppc macroparticles per cell each,
    uniform density n0, Maxwellian Te / Ti, Bz filter, Ar at 0.3 Pa / 300 K
    with constant elastic, excitation (11.5 eV) and Lotz ionization
    (15.76 eV) cross sections."""
    h = H_CELL
    L1 = n * h
    bcnd = -np.ones((n + 3, n + 3), np.int32)
    bcnd[:2, :] = bcnd[n + 1:, :] = bcnd[:, :2] = bcnd[:, n + 1:] = 1
    ngas = P_GAS / (KB * T_GAS)
    chem = RX.argon(ngas)
    Nm = n0 * h * h * 1.0 / ppc
    M = 39.95 * RX.AMU
    vte = np.sqrt(2.0 * RX.QE * Te / RX.ME)          # vt0 of main.f90:468
    vti = np.sqrt(2.0 * RX.QE * Ti / M)
    vtg = np.sqrt(2.0 * RX.QE * 0.02585 / M)
    species = [dict(name="e", charge=-RX.QE, mass=RX.ME, Nm=Nm, vt=vte, loads=[(ppc, 0.0, L1)]),
               dict(name="i", charge=RX.QE, mass=M, Nm=Nm, vt=vti, loads=[(ppc, 0.0, L1)])]
    return dict(kind="e-Ar", name=f"argon-{n}", n=n, n1=n, n2=n, h=h, Bi=_bfilter(n, n, h, Bpeak), bcnd=bcnd,
                chem=chem, Nm=Nm, ppc=ppc, vte=vte, vti=vti, vtg=vtg, dt=DT, ng=int(np.log2(n)), ngas=ngas,
                species=species, vt0=[vte, vti, vtg], ngrid=1, tag_neg=0, n0=n0,
                desc=f"{n}x{n} cells, e-/Ar+ {ppc} ppc each ({2 * n * n * ppc} particles), B filter "
                     f"{Bpeak * 1e3:g} mT, Ar 0.3 Pa analytic MCC, dt 5e-12, MG 1e-10")


def ion_source(n1=1024, n2=512, ppc=124, strip=16):
    """Negative-ion-source-like domain: e-, H+, H2+, H- over background H2
    (analytic elastic, dissociation and ionization cross sections); density
    1e17 m^-3 over the driver half falling exponentially to 1e15 at the
    extraction wall (x = L), loaded in strips of `strip` columns with the
    strip's particles per cell; species fractions e- 0.9, H+ 0.5, H2+ 0.5,
    H- 0.1 of that density (quasi-neutral), one weight Nm for all, ppc the
    particles per cell of a species at unit fraction where the density is
    1e17; grounded box (label 1), a grid electrode of bars a few cells before
    the extraction wall (label 2), the extraction wall itself (label 3); Bz
    filter field of 10 mT at L/2."""
    h = H_CELL
    L1 = n1 * h
    bcnd = -np.ones((n2 + 3, n1 + 3), np.int32)
    bcnd[:2, :] = bcnd[n2 + 1:, :] = bcnd[:, :2] = 1
    bcnd[:, n1 + 1:] = 3
    c0 = n1 - 11 if (n1 - 11) % 2 == 1 else n1 - 10
    for r0 in range(5, n2 - 3, 8):
        bcnd[r0:r0 + 5, c0:c0 + 3] = 2
    n_hi, n_lo = 1.0e17, 1.0e15
    lam = (L1 / 2) / np.log(n_hi / n_lo)

    def density(xm):
        return n_hi if xm < L1 / 2 else n_hi * np.exp(-(xm - L1 / 2) / lam)

    Nm = n_hi * h * h * 1.0 / ppc
    ngas = P_GAS / (KB * T_GAS)
    chem = RX.hydrogen(ngas)
    temps = {"e": 2.0, "Hp": 0.1, "H2p": 0.1, "Hm": 0.1}
    masses = {"e": RX.ME, "Hp": 1.00794 * RX.AMU, "H2p": 2.01588 * RX.AMU, "Hm": 1.00794 * RX.AMU}
    charges = {"e": -RX.QE, "Hp": RX.QE, "H2p": RX.QE, "Hm": -RX.QE}
    frac = {"e": 0.9, "Hp": 0.5, "H2p": 0.5, "Hm": 0.1}
    species, total = [], 0
    for name in ("e", "Hp", "H2p", "Hm"):
        loads = []
        for c in range(0, n1, strip):
            k = int(round(ppc * frac[name] * density((c + strip / 2) * h) / n_hi))
            if k > 0:
                loads.append((k, c * h, (c + strip) * h))
                total += k * strip * n2
        species.append(dict(name=name, charge=charges[name], mass=masses[name], Nm=Nm,
                            vt=np.sqrt(2.0 * RX.QE * temps[name] / masses[name]), loads=loads))
    vt0 = [np.sqrt(2.0 * RX.QE * (2.0 if k == 0 else sp.T) / abs(sp.mass)) for k, sp in enumerate(chem.species)]
    return dict(kind="ion-source", name="ion-source", n=n1, n1=n1, n2=n2, h=h, Bi=_bfilter(n1, n2, h), bcnd=bcnd,
                chem=chem, Nm=Nm, ppc=ppc, dt=DT, ng=int(np.log2(min(n1, n2))), ngas=ngas, species=species,
                vt0=vt0, ngrid=3, tag_neg=4,
                desc=f"{n1}x{n2} cells, e-/H+/H2+/H- (about {total} particles), density 1e17 m^-3 over the "
                     "driver half falling exponentially to 1e15 at the extraction wall, grid electrode, "
                     "B filter 10 mT, H2 0.3 Pa analytic MCC, dt 5e-12, MG 1e-10")


CONFIGS = {
    "argon-128": lambda: argon_box(128, 200),
    "argon-512": lambda: argon_box(512, 100),
    "ion-source": lambda: ion_source(),
    "argon-1024": lambda: argon_box(1024, 100),
    "argon-2048": lambda: argon_box(2048, 120),
}


def load_counts(W):
    """Particles each species' uniform-load strips create (before all-wall
    cells are skipped, which the grounded boxes here do not have)."""
    return [sum(k * W["n2"] * int(round((xr - xl) / W["h"])) for k, xl, xr in s["loads"]) for s in W["species"]]


def build_context(W, precision, seed=20261202, capacity_margin=0.05, track_z=0, deposit=True):
    """A context holding workload W: walls, B field, the device-loaded
    particles (counter-based generator keyed by seed), the reaction tables,
    and (deposit=True) every species' first deposition."""
    import jd_lepic_amd as L
    from . import tables
    n1, n2 = W["n1"], W["n2"]
    sp = [L.Species(s["charge"], s["mass"], s["Nm"], int(c * (1.0 + capacity_margin)) + 4096)
          for s, c in zip(W["species"], load_counts(W))]
    ctx = L.Context(n1, n2, W["h"], W["h"], W["dt"], sp, ng=W["ng"], ngrid=W["ngrid"], precision=precision,
                    track_z=track_z, tag_neg=W["tag_neg"])
    ctx.set_boundary(W["bcnd"], np.zeros((n2 + 3, n1 + 3)))
    ctx.set_bfield(W["Bi"], (n1, n2), (W["h"], W["h"]))
    for p, s in enumerate(W["species"], start=1):
        for k, xl, xr in s["loads"]:
            ctx.load_uniform(p, k, s["vt"], xl, xr, seed)
    ctx.set_reactions(W["chem"], tables.ordering(W["chem"]), vt0=W["vt0"])
    if deposit:
        for p in range(1, len(W["species"]) + 1):
            ctx.charge_deposition(p)
    return ctx


def oracle_setup(W):
    """The oracle's MCC set-up dict for workload W (species charges, masses,
    weights, thermal speeds, neutral densities, nu_uplim of main.f90:494-495)."""
    chem = W["chem"]
    nt = chem.ntype_all
    nsp = len(W["species"])
    n1, n2 = W["n1"], W["n2"]
    return dict(ntype=nsp, ntype_all=nt, ns_coll=10, plt_src=0, n_cell=n1 * n2, n_rg=1, ixl_rg=[1],
                ixr_rg=[n1 + 1], iyl_rg=1, iyr_rg=n2 + 1, flag_ionrep=0,
                charge={k + 1: s.charge for k, s in enumerate(chem.species)}, mass=chem.masses(),
                Nm={k + 1: (W["species"][k]["Nm"] if k < nsp else 0.0) for k in range(nt)},
                vt0={k + 1: W["vt0"][k] for k in range(nt)},
                np_mx={k + 1: (s.ni0 if s.charge == 0.0 else 0.0) for k, s in enumerate(chem.species)},
                nu_uplim={k + 1: (6e7 if k == 0 else 5e6) for k in range(nt)})
