"""GPU collisions through the C ABI.  Check mode against the oracle on the
same particles and ran2 seed: same candidate counts, same events per
reaction, same final ran2 state, same created particles; velocities within
1e-13 relative (log/cos/sin/acos of the device and host math libraries may
differ in the last bit).  Production mode: the elastic rate within 3 sigma
of n*sigma*<|v|>, and CM kinematics of ionization."""
import numpy as np
import pytest

import oracle as O
from mcc_setup import argon_case

L = pytest.importorskip("jd_lepic_amd")
from jd_lepic_amd import tables  # noqa: E402

pytestmark = pytest.mark.gpu


def gpu_ctx(dom, chem, parts, setup, precision=L.FP64):
    sp = [L.Species(chem.species[k].charge, chem.species[k].mass, setup["Nm"][k + 1],
                    parts[k + 1]["vxp"].shape[0]) for k in range(2)]
    ctx = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, sp, ng=1, precision=precision)
    ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)))
    for p in (1, 2):
        ctx.upload(p, parts[p]["vxp"][:parts[p]["np_tot"]])
    ctx.set_reactions(chem, tables.ordering(chem), vt0=[setup["vt0"][k] for k in (1, 2, 3)])
    return ctx


@pytest.mark.parametrize("Tgas", [0.02585, 0.0])
def test_check_mode_replays_oracle(Tgas):
    dom, chem, tab, setup, parts, _ = argon_case(n=16, N=20000, Te=20.0, Tgas=Tgas, seed=41)
    with gpu_ctx(dom, chem, parts, setup) as ctx:
        for p in (1, 2):
            ctx.charge_deposition(p)
        ctx.dens_red()
        np_red = [ctx.get_grid(L.GRID_NPRED, p) for p in (1, 2)]
        out = ctx.collisions(11, mode=L.MCC_CHECK, iseed=123456)
        ge, gi = ctx.download(1), ctx.download(2)
        cc = ctx.cnt_col()
        pl = ctx.mcc_ploss()
        gsrc = [ctx.get_grid(L.GRID_SRC, p) for p in (1, 2)]
        gsink = [ctx.get_grid(L.GRID_SINK, p) for p in (1, 2)]
    for p in (1, 2):
        parts[p]["src"] = np.zeros((dom.n2 + 3, dom.n1 + 3))
        parts[p]["sink"] = np.zeros((dom.n2 + 3, dom.n1 + 3))
    ref = O.collisions(dom, setup, tab, parts, np_red, 123456)
    # ss2D(1)/(2): a count per created / killed particle at the incident cell
    for p in (1, 2):
        assert np.array_equal(gsrc[p - 1], parts[p]["src"]) and np.array_equal(gsink[p - 1], parts[p]["sink"])
    assert parts[2]["src"].sum() == parts[2]["np_tot"] - 20000   # one count per created Ar+
    assert out["seed"] == ref["seed"]
    assert out["Nc"][0] == ref["Nc"][1]
    assert np.array_equal(out["events"], ref["events"]) and ref["events"].sum() > 0
    re, ri = parts[1]["vxp"][:parts[1]["np_tot"]], parts[2]["vxp"][:parts[2]["np_tot"]]
    assert ge.shape == re.shape and gi.shape == ri.shape
    assert np.array_equal(ge[:, :3], re[:, :3]) and np.array_equal(gi[:, :3], ri[:, :3])
    np.testing.assert_allclose(ge[:, 3:], re[:, 3:], rtol=1e-13, atol=1e-6)
    np.testing.assert_allclose(gi[:, 3:], ri[:, 3:], rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(cc, ref["cnt_col"], rtol=1e-12)
    np.testing.assert_allclose(pl[1:3], ref["P_loss3"][1:3], rtol=1e-10)


def test_production_elastic_rate_three_sigma():
    dom, chem, tab, setup, parts, _ = argon_case(n=32, N=400000, Te=3.0, Tgas=0.0, only_elastic=True,
                                                 seed=42)
    N = parts[1]["np_tot"]
    speed = np.linalg.norm(parts[1]["vxp"][:N, 3:], axis=1)
    with gpu_ctx(dom, chem, parts, setup) as ctx:
        for p in (1, 2):
            ctx.charge_deposition(p)
        ctx.dens_red()
        out = ctx.collisions(11, mode=L.MCC_PROD, seed=99)
        v = ctx.download(1)
    Nc = out["Nc"][0]
    nu_omp = float(np.float32(Nc) / np.float32(N)) / (10.0 * dom.dt)
    p = 7.2429e19 * 1e-19 * speed.mean() / nu_omp
    ev = out["events"][0]
    sd = np.sqrt(Nc * p * (1 - p)) + Nc * p * speed.std() / speed.mean() / np.sqrt(N)
    assert out["missed"] == 0
    assert abs(ev - Nc * p) < 3 * sd, (ev, Nc * p, sd)
    # elastic scattering off a target at rest keeps the speed in the CM frame
    M = 39.95 * O.AMU
    vcm0 = parts[1]["vxp"][:N, 3:] * (O.ME / (O.ME + M))
    r0 = np.linalg.norm(parts[1]["vxp"][:N, 3:] - vcm0, axis=1)
    r1 = np.linalg.norm(v[:, 3:] - vcm0, axis=1)
    changed = np.any(v[:, 3:] != parts[1]["vxp"][:N, 3:], axis=1)
    assert changed.sum() == ev
    np.testing.assert_allclose(r1[changed], r0[changed], rtol=1e-12)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_production_ionization_appends_in_order(precision):
    dom, chem, tab, setup, parts, _ = argon_case(n=16, N=50000, Te=40.0, Tgas=0.0, seed=43)
    N = parts[1]["np_tot"]
    with gpu_ctx(dom, chem, parts, setup, precision=precision) as ctx:
        for p in (1, 2):
            ctx.charge_deposition(p)
        ctx.dens_red()
        out = ctx.collisions(21, mode=L.MCC_PROD, seed=7)
        e, i = ctx.download(1), ctx.download(2)
    nion = out["events"][2]
    assert nion > 0
    assert e.shape[0] == N + nion and i.shape[0] == N + nion
    assert out["created"][0] == nion and out["created"][1] == nion
    # each new electron sits with its new ion
    tol = 1e-6 * dom.h1 if precision == L.MIXED else 0.0
    np.testing.assert_allclose(e[N:, :2], i[N:, :2], rtol=0, atol=tol)


# ---- constant cross sections above threshold, targets at rest ----

NGAS = 7.2429e19


def step_case(N_e, N_i, Te, Ti, sig_cx=0.0, emax=500.0, seed=1, n=32, precision=L.FP64):
    from mcc_setup import argon_step_chemistry
    chem = argon_step_chemistry(NGAS, sig_cx=sig_cx, emax=emax)
    h = 5e-5
    rng = np.random.default_rng(seed)
    Lb = n * h

    def load(N, T, m):
        pos = np.column_stack([rng.uniform(h, Lb - h, N), rng.uniform(h, Lb - h, N), rng.uniform(0, 1, N)])
        return np.column_stack([pos, rng.normal(0, np.sqrt(O.QE * T / m), (N, 3))])
    M = 39.95 * O.AMU
    e, i = load(N_e, Te, O.ME), load(N_i, Ti, M)
    sp = [L.Species(-O.QE, O.ME, 100.0, N_e + N_e // 4 + 1024), L.Species(O.QE, M, 100.0, N_i + N_e // 4 + 1024)]
    ctx = L.Context(n, n, h, h, 5e-12, sp, ng=1, precision=precision)
    ctx.set_boundary(O.box_bcnd(n, n), np.zeros((n + 3, n + 3)))
    ctx.upload(1, e)
    ctx.upload(2, i)
    vt0 = [np.sqrt(2 * O.QE * Te / O.ME), np.sqrt(2 * O.QE * Ti / M), 0.0]
    ctx.set_reactions(chem, tables.ordering(chem), vt0=vt0, nu_uplim=[1e10, 1e10, 1e10])
    for p in (1, 2):
        ctx.charge_deposition(p)
    ctx.dens_red()
    return ctx, chem, e, i


def expected_rate(v, m_inc, sig, Eth, Nc, N, dt=5e-12, ns_coll=10):
    """Per-candidate event probability n*sigma*<|v| 1(E >= Eth)>/nu_omp with the
    target at rest: relative speed |v|, E = mu |v|^2 / (2 qe)."""
    M = 39.95 * O.AMU
    mu = m_inc * M / (m_inc + M)
    speed = np.linalg.norm(v, axis=1)
    E = 0.5 * mu * speed * speed / O.QE
    nu_omp = float(np.float32(Nc) / np.float32(N)) / (ns_coll * dt)
    return NGAS * sig * np.mean(speed * (E >= Eth)) / nu_omp


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_production_rates_elastic_excitation_ionization_three_sigma(precision):
    """Production mode: the event count of each of the three electron
    reactions within 3 sigma of Nc * n*sigma*<|v| 1(E >= Eth)>/nu_omp
    (collisions.f90:485-647 selects reaction r of a candidate with
    probability nu_r/nu_omp)."""
    N = 2_000_000
    ctx, chem, e, _ = step_case(N, 1000, Te=20.0, Ti=0.03, seed=51, precision=precision)
    with ctx:
        out = ctx.collisions(11, mode=L.MCC_PROD, seed=2024)
    assert out["missed"] == 0
    Nc = out["Nc"][0]
    for c, (sig, Eth) in enumerate(((1e-19, 0.0), (5e-20, 11.5), (5e-20, 15.76))):
        p = expected_rate(e[:, 3:], O.ME, sig, Eth, Nc, N)
        ev = out["events"][c]
        sd = np.sqrt(Nc * p * (1 - p))
        assert p > 0.02 and abs(ev - Nc * p) < 3 * sd, (c, ev, Nc * p, sd)
    assert out["created"][0] == out["events"][2] == out["created"][1]


def test_production_two_colliding_species():
    """Electron ionization appends Ar+ before the Ar+ charge exchange runs in
    the same call: the ions' candidates come from the ions present when the
    call began (collisions.f90:184, 275), so no ion created in this call is
    exchanged; counts grow by the ionizations only; the exchange rate is
    within 3 sigma of its closed form."""
    N = 400_000
    ctx, chem, e, i = step_case(N, N, Te=40.0, Ti=30.0, sig_cx=2e-18, emax=2000.0, seed=52)
    with ctx:
        out = ctx.collisions(21, mode=L.MCC_PROD, seed=77)
        ge, gi = ctx.download(1), ctx.download(2)
    nion, ncx = out["events"][2], out["events"][3]
    assert out["missed"] == 0 and nion > 50 and ncx > 10
    assert ge.shape[0] == N + nion and gi.shape[0] == N + nion
    at_rest = np.all(gi[:, 3:] == 0.0, axis=1)        # exchanged ions take the resting target's velocity
    assert at_rest[:N].sum() == ncx
    assert not at_rest[N:].any()
    p = expected_rate(i[:, 3:], 39.95 * O.AMU, 2e-18, 0.0, out["Nc"][1], N)
    sd = np.sqrt(out["Nc"][1] * p * (1 - p))
    assert abs(ncx - out["Nc"][1] * p) < 3 * sd, (ncx, out["Nc"][1] * p, sd)


def test_check_mode_two_colliding_species_matches_oracle():
    """The same two-species chemistry in check mode: the GPU's ran2 replay
    against the oracle (events, seed, created particles, exchanged ions)."""
    N = 150_000
    ctx, chem, e, i = step_case(N, N, Te=40.0, Ti=30.0, sig_cx=2e-18, emax=2000.0, seed=53, n=16)
    with ctx:
        np_red = [ctx.get_grid(L.GRID_NPRED, p) for p in (1, 2)]
        out = ctx.collisions(11, mode=L.MCC_CHECK, iseed=4321)
        ge, gi = ctx.download(1), ctx.download(2)
    n = 16
    dom = O.Domain(n1=n, n2=n, h1=5e-5, h2=5e-5, dt=5e-12, bcnd=O.box_bcnd(n, n))
    tab = O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                     chem.masses(), chem.ntype_all)
    M = 39.95 * O.AMU
    setup = dict(ntype=2, ntype_all=3, ns_coll=10, plt_src=0, n_cell=n * n, n_rg=1, ixl_rg=[1], ixr_rg=[n + 1],
                 iyl_rg=1, iyr_rg=n + 1, flag_ionrep=0, charge={1: -O.QE, 2: O.QE, 3: 0.0},
                 mass=chem.masses(), Nm={1: 100.0, 2: 100.0, 3: 0.0},
                 vt0={1: np.sqrt(2 * O.QE * 40.0 / O.ME), 2: np.sqrt(2 * O.QE * 30.0 / M), 3: 0.0},
                 np_mx={1: 0.0, 2: 0.0, 3: NGAS}, nu_uplim={1: 1e10, 2: 1e10, 3: 1e10})
    cap = N + N // 4 + 1024
    parts = {}
    for p, arr in ((1, e), (2, i)):
        v = np.zeros((cap, 6))
        v[:N] = arr
        parts[p] = dict(vxp=v, np_tot=N, flag_dead=np.zeros(cap, np.int8))
    ref = O.collisions(dom, setup, tab, parts, np_red, 4321)
    assert out["seed"] == ref["seed"]
    assert np.array_equal(out["events"], ref["events"]) and ref["events"][3] > 0 and ref["events"][2] > 0
    re, ri = parts[1]["vxp"][:parts[1]["np_tot"]], parts[2]["vxp"][:parts[2]["np_tot"]]
    assert ge.shape == re.shape and gi.shape == ri.shape
    assert np.array_equal(ge[:, :3], re[:, :3]) and np.array_equal(gi[:, :3], ri[:, :3])
    np.testing.assert_allclose(ge[:, 3:], re[:, 3:], rtol=1e-13, atol=1e-6)
    np.testing.assert_allclose(gi[:, 3:], ri[:, 3:], rtol=1e-13, atol=1e-9)
