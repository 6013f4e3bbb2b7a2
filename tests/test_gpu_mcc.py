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
