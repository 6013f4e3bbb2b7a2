"""Volume injection on the GPU (part_injection.f90 and the main loop's
species loop) through the C ABI.

Check mode replays the reference's ran2 stream and is compared with the
oracle's lo_part_injection on the same inputs: the same count per species,
the same final seed and source maps, residents untouched, injected particles
to the last bits of the device's log/sin/cos, the same energy and potential
tallies.  Production mode is checked by the counts, the region, the velocity
laws and pairing, repeatability, and the deposited charge."""
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
N1, N2, H, DT, ZMAX = 64, 48, 5e-5, 5e-12, 0.1
MH = 1.00794 * L.AMU
LX, LY = N1 * H, N2 * H
MASS = {1: L.ME, 2: MH, 3: MH}
CHARGE = {1: -L.QE, 2: L.QE, 3: -L.QE}
NM = 1e3


def bcnd():
    bc = O.box_bcnd(N1, N2)
    bc[10:14, 20:24] = 2
    return bc


def injection(opt, flag_c=0, **kw):
    vt = {1: np.sqrt(2 * L.QE * 3.0 / L.ME), 2: np.sqrt(2 * L.QE * 0.1 / MH), 3: np.sqrt(2 * L.QE * 0.1 / MH)}
    d = dict(opt_inj=opt, flag_c=flag_c, flag_nmn=0, flag_Tp=0, flag_diffsrc=0, ns_inj=1, I_inj=0.05,
             xl_pow=0.6 * LX, xr_pow=0.2 * LX, yl_pow=0.8 * LY, yr_pow=0.1 * LY, xa=0.5 * LX, dr=0.15 * LY,
             THm=2.0, th_B=0.3, ni0={1: 1.0, 2: 0.7, 3: 0.3}, vt0=vt)
    d.update(kw)
    return d


def residents(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(3 * H, LX - 3 * H, n), rng.uniform(3 * H, LY - 3 * H, n),
                            rng.uniform(0, ZMAX, n), rng.normal(0, 1e4, (n, 3))])


def context(prec, res, tag_neg=0, cap=30000):
    sp = [L.Species(CHARGE[p], MASS[p], NM, cap) for p in (1, 2, 3)]
    ctx = L.Context(N1, N2, H, H, DT, sp, zmax=ZMAX, ng=4, ngrid=2, precision=prec, tag_neg=tag_neg)
    ctx.set_boundary(bcnd(), np.zeros((N2 + 3, N1 + 3)))
    for p in (1, 2, 3):
        ctx.upload(p, res[p])
    return ctx


def gpu_args(inj):
    return dict(inj, ni0=[inj["ni0"][p] for p in (1, 2, 3)], vt0=[inj["vt0"][p] for p in (1, 2, 3)])


def oracle_run(res, inj, iseed, tag_neg=0, N_inj=None, phi=None, cap=30000):
    dom = O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=DT, bcnd=bcnd(), zmax=ZMAX)
    parts = {}
    for p in (1, 2, 3):
        v = np.zeros((cap, 6))
        v[:res[p].shape[0]] = res[p]
        parts[p] = dict(vxp=v, np_tot=res[p].shape[0], src=np.zeros((N2 + 3, N1 + 3)))
    oi = dict(inj, tag_neg=tag_neg, Nm={p: NM for p in (1, 2, 3)}, mass=MASS, charge=CHARGE)
    out = O.part_injection(dom, oi, parts, iseed, N_inj=N_inj, phi=phi)
    assert out["rc"] == 0
    return out, parts


def close(got, want):
    """Within a few ulp of the column's scale: the device's log/sin/cos may
    differ from the host's in the last bit, and a beam offset added to a
    drawn vz can cancel most of it."""
    for c in range(6):
        scale = np.abs(want[:, c]).max() if want.shape[0] else 0.0
        np.testing.assert_allclose(got[:, c], want[:, c], rtol=1e-13, atol=1e-14 * scale + 1e-18)


CASES = {
    "slit": dict(opt=1),
    "disk": dict(opt=1, flag_c=1),
    "gauss_slit": dict(opt=-1),
    "gauss_disk": dict(opt=-1, flag_c=1),
    "beam": dict(opt=3),
    "flux": dict(opt=4, flag_nmn=1, THm=0.5, xr_pow=0.0, xl_pow=0.5 * LX),
    "losses": dict(opt=2),
    "tp_diffsrc": dict(opt=1, flag_Tp=1, flag_diffsrc=1),
}


@pytest.mark.parametrize("case", list(CASES))
def test_check_mode_matches_oracle(case):
    print(f"injection check mode case={case}", flush=True)
    kw = dict(CASES[case])
    opt = kw.pop("opt")
    inj = injection(opt, **kw)
    tag_neg = 3 if case == "losses" else 0
    lost = {1: 11, 2: 900, 3: 300} if case == "losses" else None
    res = {p: residents(5 + p, p) for p in (1, 2, 3)}
    with context(L.FP64, res, tag_neg=tag_neg) as ctx:
        phi = None
        if inj["flag_diffsrc"]:
            X, Y = np.meshgrid(np.arange(N1 + 3) * H, np.arange(N2 + 3) * H)
            phi = 5.0 * np.sin(X / LX * 3.0) * np.cos(Y / LY * 2.0) + 2.0
            ctx.set_grid(L.GRID_PHI, phi)
        n_inj = [lost[p] for p in (1, 2, 3)] if lost else None
        got = ctx.part_injection(gpu_args(inj), N_inj=n_inj, it=1, mode=L.MCC_CHECK, iseed=4321)
        down = {p: ctx.download(p) for p in (1, 2, 3)}
        src = {p: ctx.get_grid(L.GRID_SRC, p) for p in (1, 2, 3)}
    want, parts = oracle_run(res, inj, 4321, tag_neg=tag_neg, N_inj=lost, phi=phi)
    assert got["seed"] == want["seed"]
    assert sum(got["injected"]) > 500
    for p in (1, 2, 3):
        n0, n = res[p].shape[0], parts[p]["np_tot"]
        assert got["injected"][p - 1] == want["injected"][p] == n - n0
        assert down[p].shape[0] == n and np.array_equal(down[p][:n0], res[p])
        close(down[p][n0:], parts[p]["vxp"][n0:n])
        assert np.array_equal(src[p], parts[p]["src"])
        np.testing.assert_allclose(got["P_loss4"][p - 1], want["P_loss4"][p], rtol=1e-12)
        np.testing.assert_allclose(got["P_loss1"][p - 1], want["P_loss1"][p], rtol=1e-12)


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_production_counts_region_and_pairing(prec):
    print(f"injection production prec={prec}", flush=True)
    inj = injection(1)
    res = {p: residents(40, 10 + p) for p in (1, 2, 3)}
    with context(prec, res) as ctx:
        got = ctx.part_injection(gpu_args(inj), it=4, mode=L.MCC_PROD, seed=8)
        new = {p: ctx.download(p)[40:] for p in (1, 2, 3)}
        src = {p: ctx.get_grid(L.GRID_SRC, p) for p in (1, 2, 3)}
    bc = bcnd()
    for p in (1, 2, 3):
        v = new[p]
        n = v.shape[0]
        dN = inj["I_inj"] * DT * inj["ni0"][p] / (L.QE * NM)
        assert n == got["injected"][p - 1] and int(dN) <= n <= int(dN) + 1 and src[p].sum() == n
        tol = 0.0 if prec == L.FP64 else 1e-6 * H
        assert np.all((v[:, 0] >= inj["xr_pow"] - tol) & (v[:, 0] <= inj["xl_pow"] + tol))
        assert np.all((v[:, 1] >= inj["yr_pow"] - tol) & (v[:, 1] <= inj["yl_pow"] + tol))
        ix, iy = (v[:, 0] / H).astype(int) + 1, (v[:, 1] / H).astype(int) + 1
        assert not np.any((bc[iy, ix] >= 1) & (bc[iy, ix + 1] >= 1) & (bc[iy + 1, ix + 1] >= 1) &
                          (bc[iy + 1, ix] >= 1))
        # the pairing of part_injection's velocity draws
        assert np.array_equal(v[0::2, 3], v[0::2, 5])
        m = (n // 2) * 2
        assert np.array_equal(v[1:m:2, 5], v[0:m:2, 4])
        vt = inj["vt0"][p]
        assert abs(v[:, 4].var() / (vt * vt / 2) - 1) < 5 * np.sqrt(2 / n)
        e = 0.5 * NM * MASS[p] * (v[:, 3:] ** 2).sum()
        np.testing.assert_allclose(got["P_loss4"][p - 1], e, rtol=1e-9 if prec == L.FP64 else 1e-6)


@pytest.mark.parametrize("case", ["flux", "gauss_disk", "beam"])
def test_production_velocity_laws(case):
    print(f"injection production laws case={case}", flush=True)
    kw = dict(CASES[case])
    opt = kw.pop("opt")
    inj = injection(opt, **kw, I_inj=0.2)
    res = {p: residents(10, 20 + p) for p in (1, 2, 3)}
    with context(L.FP64, res, cap=120000) as ctx:
        ctx.part_injection(gpu_args(inj), it=2, mode=L.MCC_PROD, seed=5)
        new = {p: ctx.download(p)[10:] for p in (1, 2, 3)}
    v = new[1]
    n = v.shape[0]
    vt = inj["vt0"][1]
    if case == "flux":
        g = np.linspace(0.0, 4 * vt, 20001)                       # electrons: vb = 0 with flag_nmn = 1
        f = g * np.exp(-g ** 2 / vt ** 2)
        mean = np.trapezoid(g * f, g) / np.trapezoid(f, g)
        sd = np.sqrt(np.trapezoid(g * g * f, g) / np.trapezoid(f, g) - mean ** 2)
        assert abs(v[:, 3].mean() - mean) < 5 * sd / np.sqrt(n)
        assert np.all(v[:, 0] >= 0) and np.all(v[:, 0] <= v[:, 3] * DT * (1 + 1e-12))
        assert new[2].shape[0] > 0                                 # flag_nmn = 1: ions too
    elif case == "gauss_disk":
        r2 = ((v[:, 0] - inj["xa"]) ** 2 + (v[:, 1] - LY / 2) ** 2) / inj["dr"] ** 2
        assert np.all(r2 <= 4.0)
        want = (1 - 5 * np.exp(-4)) / (1 - np.exp(-4))
        assert abs(r2.mean() / want - 1) < 5 / np.sqrt(n)
    else:
        assert new[2].shape[0] == 0 and new[3].shape[0] == 0
        vb = np.sqrt(2 * L.QE * inj["THm"] / L.ME)
        assert abs(v[::2, 5].mean() - vb) < 5 * vt / np.sqrt(n)


def test_production_repeats_per_seed_and_step():
    inj = injection(1)
    res = {p: residents(20, 30 + p) for p in (1, 2, 3)}
    outs = []
    for it in (7, 7, 8):
        with context(L.FP64, res) as ctx:
            ctx.part_injection(gpu_args(inj), it=it, mode=L.MCC_PROD, seed=1)
            outs.append(ctx.download(1))
    assert np.array_equal(outs[0], outs[1]) and not np.array_equal(outs[0][20:60], outs[2][20:60])


@pytest.mark.parametrize("cap", [30000, 6140])
def test_injection_after_push_deposits_once(cap):
    """lepic_step order: push + deposit, then injection with deposit = 1
    adds the new particles' charge (fresh segments, cap 30000, or the dense
    fall-back after a pack, the electrons at cap 6140): np counts every
    particle once."""
    print(f"injection deposit cap={cap}", flush=True)
    inj = injection(1, I_inj=0.1)
    res = {p: residents(3000, 40 + p) for p in (1, 2, 3)}
    with context(L.FP64, res, cap=cap) as ctx:
        ctx.calc_rhs()
        ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        ctx.calc_efield()
        for p in (1, 2, 3):
            ctx.part_mover(p, deposit=True)
        before = {p: ctx.download(p) for p in (1, 2, 3)}
        got = ctx.part_injection(gpu_args(inj), it=1, mode=L.MCC_PROD, seed=2, deposit=True)
        after = {p: ctx.download(p) for p in (1, 2, 3)}
        dens = {p: ctx.get_grid(L.GRID_NP, p) for p in (1, 2, 3)}
    kq = np.where(bcnd() != -1, 2.0, 1.0)
    for p in (1, 2, 3):
        assert after[p].shape[0] == before[p].shape[0] + got["injected"][p - 1]
        assert np.array_equal(after[p][:before[p].shape[0]], before[p])
        np.testing.assert_allclose((dens[p] / kq).sum() * H * H * ZMAX / NM, after[p].shape[0], rtol=1e-9)


def test_step_reinjects_losses():
    """opt_inj = 2 inside lepic_step: the ions lost in the step's push come
    back as themselves."""
    inj = injection(2)
    res = {p: residents(4000, 50 + p) for p in (1, 2, 3)}
    rng = np.random.default_rng(5)
    res[2][:, 0] = rng.uniform(0.01 * H, 0.2 * H, 4000)      # ions next to the left wall, moving into it
    res[2][:2000, 3] = -3e6
    with context(L.FP64, res) as ctx:
        ctx.charge_deposition(1)
        ctx.charge_deposition(2)
        ctx.reset_tallies()
        r = ctx.step(1, mcc_mode=-1, nsort=0, inj=gpu_args(inj))
        lost = ctx.mover_tally(2, ngrid=2)
        n2 = ctx.count(2)
    assert r["injected"] > 0
    lost_ions = int(round(lost["p_mac"][0].sum()))
    assert lost_ions > 0 and n2 == 4000            # every lost ion came back


def test_step_reinjects_losses_of_skipped_steps():
    """opt_inj = 2 with ns_inj = 2: the ions lost in the push of a step without
    injection are re-injected at the next injection step together with that
    step's own losses (N_inj accumulates, main.f90:1044-1070)."""
    inj = dict(injection(2), ns_inj=2)
    res = {p: residents(4000, 60 + p) for p in (1, 2, 3)}
    rng = np.random.default_rng(6)
    res[2][:, 0] = rng.uniform(0.01 * H, 0.2 * H, 4000)
    res[2][:2000, 3] = -3e6
    with context(L.FP64, res) as ctx:
        ctx.charge_deposition(1)
        ctx.charge_deposition(2)
        ctx.reset_tallies()
        r1 = ctx.step(1, mcc_mode=-1, nsort=0, inj=gpu_args(inj))
        n_mid = ctx.count(2)
        lost1 = int(round(ctx.mover_tally(2, ngrid=2)["p_mac"][0].sum()))
        r2 = ctx.step(2, mcc_mode=-1, nsort=0, inj=gpu_args(inj))
        lost2 = int(round(ctx.mover_tally(2, ngrid=2)["p_mac"][0].sum()))
        n_end = ctx.count(2)
    assert r1["injected"] == 0 and lost1 > 0 and n_mid == 4000 - lost1
    assert r2["injected"] > 0
    assert n_end == 4000, (lost1, lost2, n_end)     # losses of both steps came back
