"""Shared MCC setup for oracle and GPU collision tests (e-/Ar+ in Ar)."""
import numpy as np

import oracle as O
from jd_lepic_amd import reactions as RX


def argon_case(n=16, N=20000, Te=3.0, Tgas=0.02585, ngas=7.2429e19, sig_el=1e-19, seed=1,
               only_elastic=False, ppc=None):
    chem = RX.argon(ngas, Tgas=Tgas, sig_el=sig_el)
    if only_elastic:
        for r in chem.reactions[1:]:
            r.S = np.zeros_like(r.S)
    h = 5e-5
    dom = O.Domain(n1=n, n2=n, h1=h, h2=h, dt=5e-12, bcnd=O.box_bcnd(n, n))
    masses = chem.masses()
    tab = O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                     masses, chem.ntype_all)
    vt0 = {k + 1: np.sqrt(2.0 * O.QE * (Te if k == 0 else s.T) / abs(s.mass))
           for k, s in enumerate(chem.species)}
    Vc = h * h * 1.0
    n0, ppc = 1e16, (ppc or max(N // (n * n), 1))
    Nm = {1: n0 * Vc / ppc, 2: n0 * Vc / ppc, 3: ngas * Vc / ppc}
    setup = dict(ntype=2, ntype_all=3, ns_coll=10, plt_src=0, n_cell=n * n, n_rg=1,
                 ixl_rg=[1], ixr_rg=[n + 1], iyl_rg=1, iyr_rg=n + 1, flag_ionrep=0,
                 charge={k + 1: s.charge for k, s in enumerate(chem.species)},
                 mass=masses, Nm=Nm, vt0=vt0,
                 np_mx={1: 0.0, 2: 0.0, 3: chem.species[2].ni0},
                 nu_uplim={1: 6e7, 2: 5e6, 3: 5e6})
    rng = np.random.default_rng(seed)
    L = n * h
    pos = np.column_stack([rng.uniform(h, L - h, N), rng.uniform(h, L - h, N), rng.uniform(0, 1, N)])
    ve = rng.normal(0, np.sqrt(O.QE * Te / O.ME), (N, 3))
    vi = rng.normal(0, np.sqrt(O.QE * 0.03 / (39.95 * O.AMU)), (N, 3))
    cap = N + N // 4 + 16
    e = np.zeros((cap, 6)); e[:N, :3] = pos; e[:N, 3:] = ve
    ion = np.zeros((cap, 6)); ion[:N, :3] = pos; ion[:N, 3:] = vi
    parts = {1: dict(vxp=e, np_tot=N, flag_dead=np.zeros(cap, np.int8)),
             2: dict(vxp=ion, np_tot=N, flag_dead=np.zeros(cap, np.int8))}
    np_red = [np.zeros((n + 3, n + 3)), np.zeros((n + 3, n + 3))]
    return dom, chem, tab, setup, parts, np_red


def step_sigma(E, Eth, s0):
    """sigma = s0 for E >= Eth, 0 below; the table carries a node just under
    the threshold so the linear interpolation ramps over 1e-12 of Eth only."""
    E = np.asarray(E, float)
    return np.where(E >= Eth, s0, 0.0)


def argon_step_chemistry(ngas, sig_el=1e-19, sig_exc=5e-20, sig_ion=5e-20, sig_cx=0.0, emax=500.0,
                         Tgas=0.0):
    """e-/Ar+ in Ar with constant cross sections above threshold (elastic from
    0, excitation from 11.5 eV, ionization from 15.76 eV) and, when sig_cx > 0,
    Ar+ + Ar charge exchange with a constant cross section.  With the target
    at rest (Tgas = 0) each production rate has the closed form
    n_gas * sigma * <|v| 1(E >= Eth)> over the incident particles."""
    chem = RX.Chemistry([RX.Species("[e]", RX.ME, -RX.QE, 0.0, 1.0),
                         RX.Species("[Ar+]", 39.95 * RX.AMU, RX.QE, 0.03, 1.0),
                         RX.Species("[Ar]", -39.95 * RX.AMU, 0.0, Tgas, ngas)], [])
    ths = (11.5, 15.76)
    E = RX.energy_grid(ths + tuple(t * (1.0 - 1e-12) for t in ths), emax=emax, n=120)
    chem.reactions.append(RX.Reaction("[e]+[Ar]->[e]+[Ar]", [1, 3], [1, 3], RX.COLLISION, 0.0, 0.0, E,
                                      np.full_like(E, sig_el)))
    chem.reactions.append(RX.Reaction("[e]+[Ar]->[e]+[Ar]", [1, 3], [1, 3], RX.EXCITATION, 11.5, 0.0, E,
                                      step_sigma(E, 11.5, sig_exc)))
    chem.reactions.append(RX.Reaction("[e]+[Ar]->[e]+[Ar+]+[e]", [1, 3], [1, 2, 1], RX.IONIZATION, 15.76,
                                      0.0, E, step_sigma(E, 15.76, sig_ion)))
    if sig_cx > 0.0:
        chem.reactions.append(RX.Reaction("[Ar+]+[Ar]->[Ar]+[Ar+]", [2, 3], [3, 2], RX.CHARGEEXCHANGE, 0.0,
                                          0.0, E, np.full_like(E, sig_cx)))
    return chem
