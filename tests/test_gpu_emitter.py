"""The surface emitter on the GPU (part_flux_emitter.f90) through the C ABI.

Check mode replays the reference's ran2 stream and is compared with the
oracle's lo_flux_emitter (itself bit-exact against the reference's compiled
part_flux_emitter.f90): same count per species, same final seed, same
source map and chamfer/flat split, the residents untouched, the emitted
particles to the last bits of the device's log/sin/cos.  Production mode is
checked by the emitted count, the flux-Maxwellian normal speed, the
direction toward the plasma and repeatability per (seed, it).  Appends go
into fresh segments after a push (no repacking) or, when those would not
fit the arrays, densely after a pack."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import jd_lepic_amd as L  # noqa: E402
import oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu
N1, N2, H, DT, ZMAX = 64, 48, 5e-5, 5e-12, 0.32
MH = 1.00794 * L.AMU
XG1 = 39 * H
MASS = {1: L.ME, 2: MH, 3: MH, 4: MH}
CHARGE = {1: -L.QE, 2: L.QE, 3: -L.QE, 4: -L.QE}


def bcnd():
    bc = O.box_bcnd(N1, N2)
    for ix in range(40, 45):
        for iy in range(2, N2 + 1):
            if not 20 <= iy <= 30:
                bc[iy, ix] = 2
    return bc


def emitter(n_icp=1, chamfer=False, jne=10.0):
    ymax = N2 * H
    return dict(jne=jne, Lg=0.9 * ymax, Lh=0.1 * ymax, xg1=XG1, THm=1.0, n_holes=0 if n_icp == 1 else 1,
                n_icp=n_icp, ptype_fx=3, tag_b=4 if chamfer else 0, flag_chmfrd=1 if chamfer else 0, ns_Hm=1,
                chmfrd_array=(XG1, XG1 + 2 * H, 0.3 * ymax, 0.4 * ymax))


def residents(n, seed):
    """n plasma particles left of the electrode, at rest enough to stay."""
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(3 * H, 30 * H, n), rng.uniform(3 * H, (N2 - 3) * H, n),
                            rng.uniform(0, ZMAX, n), rng.normal(0, 1e3, (n, 3))])


def context(prec, res, cap=12000):
    sp = [L.Species(CHARGE[p], MASS[p], 1e3, cap) for p in (1, 2, 3, 4)]
    ctx = L.Context(N1, N2, H, H, DT, sp, zmax=ZMAX, ng=4, ngrid=2, precision=prec, tag_neg=3)
    ctx.set_boundary(bcnd(), np.zeros((N2 + 3, N1 + 3)))
    for p in (1, 2, 3, 4):
        ctx.upload(p, res[p])
    return ctx


def oracle_run(res, em, iseed, cap=12000):
    dom = O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=DT, bcnd=bcnd(), zmax=ZMAX)
    parts = {}
    for p in (1, 2, 3, 4):
        v = np.zeros((cap, 6))
        v[:res[p].shape[0]] = res[p]
        parts[p] = dict(vxp=v, np_tot=res[p].shape[0], src=np.zeros((N2 + 3, N1 + 3)))
    oe = dict(em, tag_neg=3, Nm={p: 1e3 for p in (1, 2, 3, 4)}, mass=MASS)
    out = O.flux_emitter(dom, oe, parts, iseed)
    assert out["rc"] == 0
    return out, parts


@pytest.mark.parametrize("n_icp,chamfer", [(1, False), (2, False), (1, True)])
def test_check_mode_matches_oracle(n_icp, chamfer):
    print(f"emitter check mode n_icp={n_icp} chamfer={chamfer}", flush=True)
    res = {p: residents(7 + p, p) for p in (1, 2, 3, 4)}
    em = emitter(n_icp, chamfer)
    want, parts = oracle_run(res, em, 8642)
    with context(L.FP64, res) as ctx:
        got = ctx.flux_emitter(em, it=1, mode=L.MCC_CHECK, iseed=8642)
        down = {p: ctx.download(p) for p in (1, 2, 3, 4)}
        src = {p: ctx.get_grid(L.GRID_SRC, p) for p in (1, 2, 3, 4)}
    assert got["seed"] == want["seed"] and got["Nh"] == want["Nh"] > 1000
    assert got["cnt_b"].tolist() == want["cnt_b"].tolist()
    assert got["Lgrd"] == want["Lgrd"] and got["dNh"] == want["dNh"]
    for p in (1, 2, 3, 4):
        n0, n = res[p].shape[0], parts[p]["np_tot"]
        assert down[p].shape[0] == n and got["created"][p - 1] == n - n0
        assert np.array_equal(down[p][:n0], res[p])
        np.testing.assert_allclose(down[p][n0:], parts[p]["vxp"][n0:n], rtol=1e-13, atol=1e-18)
        assert np.array_equal(src[p], parts[p]["src"])
        np.testing.assert_allclose(got["P_loss4"][p - 1], want["P_loss4"][p], rtol=1e-12)
    if chamfer:
        assert got["created"][3] > 0


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_production_current_and_velocities(prec):
    print(f"emitter production prec={prec}", flush=True)
    res = {p: residents(50, 10 + p) for p in (1, 2, 3, 4)}
    em = emitter(1, False, jne=30.0)
    with context(prec, res, cap=20000) as ctx:
        got = ctx.flux_emitter(em, it=3, mode=L.MCC_PROD, seed=17)
        new = ctx.download(3)[50:]
        src = ctx.get_grid(L.GRID_SRC, 3)
    n = new.shape[0]
    assert n == got["Nh"] == got["created"][2] and int(got["dNh"]) <= n <= int(got["dNh"]) + 1
    assert src.sum() == n and got["cnt_b"].tolist() == [n, 0]
    vt = np.sqrt(2 * L.QE * em["THm"] / MH)
    assert np.all(new[:, 3] < 0)                      # toward the plasma on the left
    assert abs(np.mean((new[:, 3] / vt) ** 2) - 1) < 5 / np.sqrt(n)
    assert abs(np.var(new[:, 4]) / (vt * vt / 2) - 1) < 5 * np.sqrt(2 / n)
    tol = 1e-12 if prec == L.FP64 else 2e-6 * H
    assert np.all(new[:, 0] <= XG1 + tol) and np.all(new[:, 0] >= XG1 + new[:, 3] * DT - tol)
    iy = (new[:, 1] / H).astype(int) + 1
    assert not np.any((iy >= 21) & (iy <= 28))        # not from the aperture
    e = 0.5 * 1e3 * MH * (new[:, 3:] ** 2).sum()
    np.testing.assert_allclose(got["P_loss4"][2], e, rtol=1e-9 if prec == L.FP64 else 1e-6)


def test_production_repeats_per_seed_and_step():
    res = {p: residents(20, 20 + p) for p in (1, 2, 3, 4)}
    em = emitter(2, False)
    outs = []
    for it in (5, 5, 6):
        with context(L.FP64, res) as ctx:
            ctx.flux_emitter(em, it=it, mode=L.MCC_PROD, seed=3)
            outs.append(ctx.download(3))
    assert np.array_equal(outs[0], outs[1]) and not np.array_equal(outs[0][20:50], outs[2][20:50])


@pytest.mark.parametrize("cap", [12000, 4000])
def test_append_after_push_and_mover_continues(cap):
    """After a push the species sits in partly filled segments: the emission
    lands in fresh segments (cap 12000) or, when those would overflow the
    arrays, after a pack (cap 4000); either way the species reads residents
    then emitted particles, and the next push runs over all of them."""
    print(f"emitter append cap={cap}", flush=True)
    res = {p: residents(1100, 30 + p) for p in (1, 2, 3, 4)}
    em = emitter(1, False, jne=12.5)
    with context(L.FP64, res, cap=cap) as ctx:
        ctx.calc_rhs()
        ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        ctx.calc_efield()
        for p in (1, 2, 3, 4):
            ctx.part_mover(p, deposit=True)
        before = ctx.download(3)
        got = ctx.flux_emitter(em, it=2, mode=L.MCC_CHECK, iseed=99)
        after = ctx.download(3)
        assert after.shape[0] == before.shape[0] + got["Nh"] and got["Nh"] > 2000
        assert np.array_equal(after[:before.shape[0]], before)
        ctx.part_mover(3, deposit=True)
        moved = ctx.download(3)
        np_dep = ctx.get_grid(L.GRID_NP, 3)
    assert 0 < moved.shape[0] <= after.shape[0] and np.all(np.isfinite(moved))
    # every surviving particle is deposited once: kq-weighted sum of np
    kq = np.where(bcnd() != -1, 2.0, 1.0)
    np.testing.assert_allclose((np_dep / kq).sum() * H * H * ZMAX / 1e3, moved.shape[0], rtol=1e-9)
