"""The negative-ion-source workload (1024x512 cells, e-/H+/H2+/H-, H2
chemistry, grid electrodes, exponential density) at reduced size: four
species in one context through the mover, the collisions (check mode), the
Poisson solve with biased electrodes, and the mixed-precision push."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import jd_lepic_amd as L  # noqa: E402
import oracle as O  # noqa: E402
import scenarios as S  # noqa: E402
from jd_lepic_amd import reactions as RX, tables  # noqa: E402

pytestmark = pytest.mark.gpu

H_MASS = 1.00794 * O.AMU
SPECIES = [  # charge, mass, thermal speed scale, Nm
    (-O.QE, O.ME, 8.0e5, 2000.0),
    (O.QE, H_MASS, 1.0e4, 2000.0),
    (O.QE, 2.01588 * O.AMU, 7.0e3, 1500.0),
    (-O.QE, H_MASS, 1.0e4, 500.0),
]
TAG_NEG = 4


def make_ctx(dom, caps, precision=L.FP64, ng=1, phi=None):
    sp = [L.Species(c, m, nm, cap) for (c, m, _, nm), cap in zip(SPECIES, caps)]
    ctx = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, sp, ng=ng, ngrid=dom.ngrid, precision=precision,
                    flag_B_pos=dom.flag_B_pos, tag_neg=TAG_NEG)
    ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)) if phi is None else phi)
    ctx.set_bfield(dom.Bi, dom.nB, dom.hB)
    return ctx


def test_four_species_mover_bit_exact():
    dom = S.h2_source_domain()
    parts = [S.exp_profile_particles(dom, 20000, vt, seed=31 + k) for k, (_, _, vt, _) in enumerate(SPECIES)]
    fds = [(np.random.default_rng(40 + k).random(len(p)) < 0.01).astype(np.int8) for k, p in enumerate(parts)]
    Ei = S.random_E(dom, 3e4, seed=33)
    with make_ctx(dom, [len(p) + 64 for p in parts]) as ctx:
        for k, p in enumerate(parts):
            ctx.upload(k + 1, p)
            ctx.set_flag_dead(k + 1, fds[k])
        ctx.set_grid(L.GRID_E, Ei)
        for k in range(4):
            ctx.part_mover(k + 1, deposit=True)
        got = [ctx.download(k + 1) for k in range(4)]
        tallies = [ctx.mover_tally(k + 1, dom.ngrid) for k in range(4)]
        rho = [ctx.get_grid(L.GRID_NP, k + 1) for k in range(4)]
    for k, (c, m, _, nm) in enumerate(SPECIES):
        ref = O.part_mover(dom, c, m, nm, Ei, parts[k], fds[k], tag_neg=1 if k + 1 == TAG_NEG else 0)
        assert got[k].shape[0] == ref["np_tot"] < len(parts[k]), k
        assert np.array_equal(got[k], ref["vxp"]), k
        np.testing.assert_allclose(tallies[k]["P_loss"], ref["P_loss"], rtol=1e-12, atol=0)
        assert np.array_equal(tallies[k]["p_mac"][0], ref["p_mac"][0])
        assert tallies[k]["cnt_dead"] == pytest.approx(ref["cnt_dead"], rel=1e-12, abs=0)
        rho_ref = O.charge_deposition(dom, nm, ref["vxp"])
        assert np.abs(rho[k] - rho_ref).max() <= 1e-12 * np.abs(rho_ref).max()
    # H- is the tagged negative species: its dead-flagged particles are counted
    assert tallies[TAG_NEG - 1]["cnt_dead"] != 0.0


def h2_setup(dom, chem, parts):
    n1, n2 = dom.n1, dom.n2
    masses = chem.masses()
    vt0 = {k + 1: np.sqrt(2.0 * O.QE * (2.0 if k == 0 else s.T) / abs(s.mass))
           for k, s in enumerate(chem.species)}
    return dict(ntype=4, ntype_all=6, ns_coll=10, plt_src=0, n_cell=n1 * n2, n_rg=1, ixl_rg=[1],
                ixr_rg=[n1 + 1], iyl_rg=1, iyr_rg=n2 + 1, flag_ionrep=0,
                charge={k + 1: s.charge for k, s in enumerate(chem.species)}, mass=masses,
                Nm={1: SPECIES[0][3], 2: SPECIES[1][3], 3: SPECIES[2][3], 4: SPECIES[3][3], 5: 1.0, 6: 1.0},
                vt0=vt0, np_mx={1: 0.0, 2: 0.0, 3: 0.0, 4: 0.0, 5: chem.species[4].ni0, 6: 0.0},
                nu_uplim={1: 6e7, 2: 5e6, 3: 5e6, 4: 5e6, 5: 5e6, 6: 5e6})


def test_h2_collisions_check_mode_replays_oracle():
    dom = S.h2_source_domain(32, 16)
    ngas = 1.0 / (1.380649e-23 * 300.0)       # 1 Pa of H2
    chem = RX.hydrogen(ngas)
    tab = O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                     chem.masses(), chem.ntype_all)
    N = 20000
    vx = [S.exp_profile_particles(dom, N, vt * (4.0 if k == 0 else 1.0), seed=51 + k)
          for k, (_, _, vt, _) in enumerate(SPECIES)]
    cap = N + N // 4
    parts = {}
    for k in range(4):
        a = np.zeros((cap, 6))
        a[:N] = vx[k]
        parts[k + 1] = dict(vxp=a, np_tot=N, flag_dead=np.zeros(cap, np.int8))
    setup = h2_setup(dom, chem, parts)
    with make_ctx(dom, [cap] * 4) as ctx:
        for k in range(4):
            ctx.upload(k + 1, vx[k])
        ctx.set_reactions(chem, tables.ordering(chem), vt0=[setup["vt0"][k] for k in range(1, 7)])
        for k in range(4):
            ctx.charge_deposition(k + 1)
        ctx.dens_red()
        np_red = [ctx.get_grid(L.GRID_NPRED, k + 1) for k in range(4)]
        out = ctx.collisions(11, mode=L.MCC_CHECK, iseed=424242)
        got = [ctx.download(k + 1) for k in range(4)]
    ref = O.collisions(dom, setup, tab, parts, np_red, 424242)
    assert out["seed"] == ref["seed"]
    assert np.array_equal(out["events"], ref["events"]) and ref["events"].sum() > 0
    for k in range(4):
        r = parts[k + 1]["vxp"][:parts[k + 1]["np_tot"]]
        assert got[k].shape == r.shape, k
        assert np.array_equal(got[k][:, :3], r[:, :3]), k
        np.testing.assert_allclose(got[k][:, 3:], r[:, 3:], rtol=1e-13, atol=1e-6)
    # ionization appended H2+ ions and electrons
    assert parts[3]["np_tot"] > N and parts[1]["np_tot"] > N


@pytest.mark.parametrize("tail", [0, 50])
def test_poisson_with_biased_electrodes(tail):
    dom = S.h2_source_domain()
    n1, n2, ng = dom.n1, dom.n2, 5
    phi = np.zeros((n2 + 3, n1 + 3))
    phi[dom.bcnd == 2] = 5.0                 # grid electrode
    phi[dom.bcnd == 3] = 10.0                # extraction wall
    rng = np.random.default_rng(61)
    b = rng.normal(0, 1e-3, (n2 + 1, n1 + 1)) * np.exp(-np.linspace(0, 4.6, n1 + 1))[None, :]
    b[dom.bcnd[1:n2 + 2, 1:n1 + 2] != -1] = 0.0
    sp = [L.Species(c, m, nm, 16) for (c, m, _, nm) in SPECIES]
    ctx = L.Context(n1, n2, dom.h1, dom.h2, dom.dt, sp, ng=ng, ngrid=dom.ngrid, mg_tail_nodes=tail)
    with ctx:
        ctx.set_boundary(dom.bcnd, phi)
        ctx.set_grid(L.GRID_RHS, b)
        info = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        got = ctx.get_grid(L.GRID_PHI)
    ref = O.iterative_solver(phi, b, dom.bcnd, (dom.h1, dom.h2), n1, n2, ng, eps=1e-10)
    assert info["vcycles"] == ref["vcycles"]
    assert np.array_equal(info["sweeps"], ref["sweeps"][:len(info["sweeps"])])
    assert np.abs(got - ref["phi"]).max() <= 1e-12 * np.abs(ref["phi"]).max()


def test_four_species_mixed_push_close_to_fp64():
    dom = S.h2_source_domain()
    parts = [S.exp_profile_particles(dom, 20000, vt, seed=71 + k) for k, (_, _, vt, _) in enumerate(SPECIES)]
    Ei = S.random_E(dom, 3e4, seed=72)
    out = {}
    for prec in (L.FP64, L.MIXED):
        with make_ctx(dom, [len(p) + 64 for p in parts], precision=prec) as ctx:
            for k, p in enumerate(parts):
                ctx.upload(k + 1, p)
            ctx.set_grid(L.GRID_E, Ei)
            for k in range(4):
                ctx.part_mover(k + 1, deposit=True)
            out[prec] = ([ctx.download(k + 1) for k in range(4)], [ctx.get_grid(L.GRID_NP, k + 1) for k in range(4)])
    for k in range(4):
        a, b = out[L.FP64][0][k], out[L.MIXED][0][k]
        assert abs(a.shape[0] - b.shape[0]) <= 2, k      # wall crossings decided in FP32 may differ
        n = min(a.shape[0], b.shape[0])
        h = dom.h1
        cell_a = np.floor(a[:n, :2] / h).astype(int)
        cell_b = np.floor(b[:n, :2] / h).astype(int)
        assert (cell_a == cell_b).all(axis=1).mean() > 0.999, k
        ra, rb = out[L.FP64][1][k], out[L.MIXED][1][k]
        assert np.abs(ra - rb).max() <= 2e-3 * np.abs(ra).max(), k
