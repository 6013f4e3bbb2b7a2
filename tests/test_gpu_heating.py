"""Electron heating on the GPU (eheating, e_heating.f90:1-112).

Check mode replays the reference's ran2 stream in particle order and is
compared with the oracle's lo_eheating (itself bit-exact against the
reference's compiled e_heating.f90): the same particles heated, the same
final seed, positions untouched and velocities to the last bits of the
device's libm.  Production mode is checked by what the heating must do: the
heated fraction inside the region is nu*dt within 4 sigma, nothing outside
changes, heated in-plane velocities carry variance vt^2/2 of their strip,
the returned power is the energy change of the stored velocities, and the
same (seed, it) gives the same result."""
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
N, H, DT, NM = 64, 5e-5, 5e-12, 1e3
LEN = N * H


def particles(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0.02 * LEN, 0.98 * LEN, n), rng.uniform(0.02 * LEN, 0.98 * LEN, n),
                            np.zeros(n), rng.normal(0, 3e5, (n, 3))])


def heat_params(flag_c, flag_gpp, eheat_type, nudt):
    return dict(flag_c=flag_c, flag_gpp=flag_gpp, eheat_type=eheat_type, xa=0.4 * LEN, dr=0.2 * LEN,
                yl_pow=0.2 * LEN, yr_pow=0.7 * LEN, nudt=nudt, Pabs_cor=0.8,
                vt=np.array([8e5, 1.1e6, 1.4e6]))


def context(prec, v):
    sp = [L.Species(-L.QE, L.ME, NM, v.shape[0] + 4096), L.Species(L.QE, 40 * L.AMU, NM, 4096)]
    ctx = L.Context(N, N, H, H, DT, sp, ng=4, precision=prec, track_z=0)
    ctx.set_boundary(O.box_bcnd(N, N), np.zeros((N + 3, N + 3)))
    ctx.upload(1, v)
    return ctx


def inside(v, h):
    if h["flag_c"] == 0:
        return (np.abs(v[:, 0] - h["xa"]) <= h["dr"]) & (v[:, 1] >= h["yl_pow"]) & (v[:, 1] <= h["yr_pow"])
    return (v[:, 0] - h["xa"]) ** 2 + (v[:, 1] - LEN / 2) ** 2 <= h["dr"] ** 2


@pytest.mark.parametrize("flag_c,flag_gpp,eheat_type", [(0, 0, 1), (1, 0, 1), (1, 1, 0), (0, 1, 0)])
def test_check_mode_matches_oracle(flag_c, flag_gpp, eheat_type):
    print(f"heating check mode flag_c={flag_c} flag_gpp={flag_gpp} eheat_type={eheat_type}", flush=True)
    v = particles(20000, 3)
    h = heat_params(flag_c, flag_gpp, eheat_type, 0.3)
    oh = dict(h, ymax=LEN, dt=DT, Nm=NM, mass=L.ME)
    want, dPw, sw = O.eheating(v, oh, 987654)
    with context(L.FP64, v) as ctx:
        dP, s = ctx.eheating(h, mode=L.MCC_CHECK, iseed=987654)
        got = ctx.download(1)
    assert s == sw
    assert np.array_equal(got[:, :3], want[:, :3])
    changed = np.any(want != v, axis=1)
    assert changed.sum() > 100 and np.array_equal(np.any(got != v, axis=1), changed)
    # device log/sin/cos against glibc's: within a few ulp
    np.testing.assert_allclose(got[:, 3:], want[:, 3:], rtol=1e-13, atol=1e-6)
    np.testing.assert_allclose(dP, dPw, rtol=1e-10)


def test_check_mode_rejects_mixed():
    v = particles(100, 1)
    with context(L.MIXED, v) as ctx:
        with pytest.raises(L.LepicError):
            ctx.eheating(heat_params(0, 0, 1, 0.3), mode=L.MCC_CHECK)


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_production_statistics(prec):
    print(f"heating production prec={prec}", flush=True)
    v = particles(400000, 7)
    h = heat_params(0, 0, 1, 0.25)
    with context(prec, v) as ctx:
        before = ctx.download(1)
        dP, _ = ctx.eheating(h, it=5, mode=L.MCC_PROD, seed=99)
        got = ctx.download(1)
    assert np.array_equal(got[:, :3], before[:, :3])
    ins = inside(before, h)
    changed = np.any(got[:, 3:] != before[:, 3:], axis=1)
    assert not np.any(changed & ~ins)
    m, p = ins.sum(), 0.25
    assert abs(changed.sum() - p * m) < 4 * np.sqrt(m * p * (1 - p)), (changed.sum(), p * m)
    strip = np.minimum((3 * before[:, 1] / LEN).astype(int), 2)
    for k in range(3):
        sel = changed & (strip == k)
        vt = h["vt"][k]
        for c in (3, 4, 5):
            var = got[sel, c].var()
            # sample variance of ~n Gaussians: relative error sqrt(2/n)
            assert abs(var / (vt * vt / 2) - 1) < 5 * np.sqrt(2 / sel.sum()), (k, c, var / (vt * vt / 2))
    e_new = 0.5 * NM * L.ME * (got[changed, 3:] ** 2).sum()
    e_old = 0.5 * NM * L.ME * (before[changed, 3:] ** 2).sum()
    np.testing.assert_allclose(dP, e_new - e_old, rtol=1e-9 if prec == L.FP64 else 1e-6)


@pytest.mark.parametrize("prec", [L.FP64, L.MIXED])
def test_production_gaussian_profile_and_determinism(prec):
    """flag_gpp = 1 on the disk: the heated fraction at distance r is
    nudt*9*Pabs_cor*exp(-9 r^2/dr^2); the same (seed, it) repeats, another
    it draws anew."""
    print(f"heating production gpp prec={prec}", flush=True)
    v = particles(400000, 11)
    h = heat_params(1, 1, 0, 0.1)
    outs = []
    for it in (3, 3, 4):
        with context(prec, v) as ctx:
            before = ctx.download(1)
            ctx.eheating(h, it=it, mode=L.MCC_PROD, seed=5)
            outs.append(ctx.download(1))
    assert np.array_equal(outs[0], outs[1]) and not np.array_equal(outs[0], outs[2])
    changed = np.any(outs[0][:, 3:] != before[:, 3:], axis=1)
    r2 = ((before[:, 0] - h["xa"]) ** 2 + (before[:, 1] - LEN / 2) ** 2) / h["dr"] ** 2
    prob = np.where(r2 <= 1.0, h["nudt"] * 9 * h["Pabs_cor"] * np.exp(-9 * r2), 0.0)
    for lo, hi in [(0.0, 0.05), (0.05, 0.2), (0.2, 0.5), (0.5, 1.0)]:
        sel = (r2 >= lo) & (r2 < hi)
        exp_n, var_n = prob[sel].sum(), (prob[sel] * (1 - prob[sel])).sum()
        assert abs(changed[sel].sum() - exp_n) < 4 * np.sqrt(var_n) + 1, (lo, changed[sel].sum(), exp_n)
    assert not np.any(changed & (r2 > 1.0))


def test_step_heats_before_the_push():
    """lepic_step calls the heating on it % ns_heat == 0 and reports its power."""
    v = particles(50000, 13)
    h = heat_params(0, 0, 0, 0.5)
    with context(L.FP64, v) as ctx:
        ctx.charge_deposition(1)
        r1 = ctx.step(1, mcc_mode=-1, heat=h, ns_heat=2)
        r2 = ctx.step(2, mcc_mode=-1, heat=h, ns_heat=2)
    assert r1["P_heat"] == 0.0 and r2["P_heat"] > 0.0
