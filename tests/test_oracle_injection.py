"""The oracle's volume injection (lo_part_injection, the restatement of
part_injection.f90 and the main loop's call of it).  part_injection.f90
includes mpif.h, which this image lacks, so the reference routine cannot be
compiled: the oracle is pinned by what the routine fixes in closed form --
the number injected per species for each opt_inj, the injection region, the
skipped wall cells, the velocity laws (Maxwellian; the z beam; the shifted
Maxwellian flux off x = 0), the pairing of its velocity draws (vx = vz on
every other particle, whose vy becomes the next particle's vz), and the
energy, potential and source tallies."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

N1, N2, H, DT = 64, 48, 5e-5, 5e-12
MH = 1.00794 * O.AMU
LX, LY = N1 * H, N2 * H


def domain():
    bc = O.box_bcnd(N1, N2)
    bc[10:14, 20:24] = 2            # a wall block inside the slit region
    return O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=DT, bcnd=bc, zmax=0.1)


def inj(opt, flag_c=0, **kw):
    vt = {1: np.sqrt(2 * O.QE * 3.0 / O.ME), 2: np.sqrt(2 * O.QE * 0.1 / MH), 3: np.sqrt(2 * O.QE * 0.1 / MH)}
    d = dict(opt_inj=opt, flag_c=flag_c, flag_nmn=0, flag_Tp=0, flag_diffsrc=0, ns_inj=1, tag_neg=0,
             I_inj=0.16, xl_pow=0.6 * LX, xr_pow=0.2 * LX, yl_pow=0.8 * LY, yr_pow=0.1 * LY, xa=0.5 * LX,
             dr=0.15 * LY, THm=2.0, th_B=0.3, ni0={1: 1.0, 2: 0.7, 3: 0.3}, vt0=vt,
             Nm={1: 1e3, 2: 1e3, 3: 1e3}, mass={1: O.ME, 2: MH, 3: MH}, charge={1: -O.QE, 2: O.QE, 3: O.QE})
    d.update(kw)
    return d


def parts(ntype=3, cap=40000):
    return {p: dict(vxp=np.zeros((cap, 6)), np_tot=0, src=np.zeros((N2 + 3, N1 + 3))) for p in range(1, ntype + 1)}


def dN(i, p):
    return i["I_inj"] * i["ns_inj"] * DT * i["ni0"][p] / (O.QE * i["Nm"][1])


def test_fixed_current_counts_region_and_walls():
    d, i, P = domain(), inj(1), parts()
    out = O.part_injection(d, i, P, 1234)
    assert out["rc"] == 0
    for p in (1, 2, 3):
        n = P[p]["np_tot"]
        assert int(dN(i, p)) <= n <= int(dN(i, p)) + 1 and n > 1000, (p, n, dN(i, p))
        v = P[p]["vxp"][:n]
        assert np.all((v[:, 0] >= i["xr_pow"]) & (v[:, 0] <= i["xl_pow"]))
        assert np.all((v[:, 1] >= i["yr_pow"]) & (v[:, 1] <= i["yl_pow"]))
        ix, iy = (v[:, 0] / H).astype(int) + 1, (v[:, 1] / H).astype(int) + 1
        walls = d.bcnd
        allw = (walls[iy, ix] >= 1) & (walls[iy, ix + 1] >= 1) & (walls[iy + 1, ix + 1] >= 1) & (walls[iy + 1, ix] >= 1)
        assert not np.any(allw)                        # no particle in a wall cell
        # uniform over the region outside the walls: mean x within 5 sigma
        lo, hi = i["xr_pow"], i["xl_pow"]
        assert abs(v[:, 0].mean() - (lo + hi) / 2) < 5 * (hi - lo) / np.sqrt(12 * n) + 0.02 * (hi - lo)
        # source map: one count per particle at its cell
        src = np.zeros_like(P[p]["src"])
        np.add.at(src, (iy, ix), 1.0)
        assert np.array_equal(src, P[p]["src"])
        e = 0.5 * i["Nm"][p] * i["mass"][p] * (v[:, 3:] ** 2).sum()
        assert abs(out["P_loss4"][p] / e - 1) < 1e-12 and out["injected"][p] == n


def test_velocity_pairing_and_maxwellian():
    """Odd particles take the second pair of draws for vx and vy and set vz =
    vx; even particles take the first pair and the previous vy as vz."""
    d, i, P = domain(), inj(1), parts()
    O.part_injection(d, i, P, 99)
    n = P[1]["np_tot"]
    v = P[1]["vxp"][:n]
    assert np.array_equal(v[0::2, 3], v[0::2, 5])
    m = (n // 2) * 2
    assert np.array_equal(v[1:m:2, 5], v[0:m:2, 4])
    vt = i["vt0"][1]
    for c in (3, 4, 5):
        assert abs(v[:, c].var() / (vt * vt / 2) - 1) < 5 * np.sqrt(2 / n)


def test_ion_loss_reinjection_counts():
    """opt_inj = 2: each lost ion comes back as itself, electrons follow the
    ions' total weighted by ni0(1), the tagged negative ion its own share."""
    d = domain()
    i = inj(2, tag_neg=3, ni0={1: 0.8, 2: 1.0, 3: 0.2})
    P = parts()
    lost = {1: 17, 2: 1500, 3: 400}
    O.part_injection(d, i, P, 7, N_inj=lost)
    tmp = lost[2] + lost[3] - lost[3]
    assert P[2]["np_tot"] == lost[2]
    assert int(0.8 * tmp) <= P[1]["np_tot"] <= int(0.8 * tmp) + 1
    assert int(0.2 * tmp) <= P[3]["np_tot"] <= int(0.2 * tmp) + 1


def test_beam_along_z_injects_electrons_only():
    d, i, P = domain(), inj(3), parts()
    O.part_injection(d, i, P, 5)
    assert P[2]["np_tot"] == 0 and P[3]["np_tot"] == 0
    n = P[1]["np_tot"]
    v = P[1]["vxp"][:n]
    vb = np.sqrt(2 * O.QE * i["THm"] / O.ME)
    assert abs(v[:, 5].mean() - vb) < 5 * i["vt0"][1] / np.sqrt(2 * n)


def test_gaussian_disk_profile():
    """opt_inj < 0 on the disk: (x, y) Gaussian about (xa, ymax/2) with
    component variance dr^2/2, cut at radius 2 dr."""
    d, i, P = domain(), inj(-1, flag_c=1), parts()
    O.part_injection(d, i, P, 11)
    n = P[1]["np_tot"]
    v = P[1]["vxp"][:n]
    dx, dy = v[:, 0] - i["xa"], v[:, 1] - LY / 2
    r2 = (dx * dx + dy * dy) / i["dr"] ** 2
    assert np.all(r2 <= 4.0)
    # E[r^2 | r^2 <= 4] for r^2 ~ Exp(mean 1): (1 - 5 e^-4) / (1 - e^-4)
    want = (1 - 5 * np.exp(-4)) / (1 - np.exp(-4))
    assert abs(r2.mean() / want - 1) < 5 / np.sqrt(n)
    assert abs(dx.mean()) < 5 * i["dr"] / np.sqrt(2 * n)


def test_shifted_maxwellian_flux_off_the_left_wall():
    """opt_inj = 4 with flag_nmn = 1: every species leaves x = 0 with a normal
    speed drawn from v exp(-(v - vb cos th_B)^2/vt^2) on [max(0, vb cos th_B -
    4 vt), vb cos th_B + 4 vt], electrons with vb = 0."""
    d = O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=DT, bcnd=O.box_bcnd(N1, N2), zmax=0.1)
    i = inj(4, flag_nmn=1, THm=0.5, xr_pow=0.0, xl_pow=0.5 * LX)
    P = parts()
    out = O.part_injection(d, i, P, 21)
    assert out["rc"] == 0
    for p, vb in ((1, 0.0), (2, np.sqrt(2 * O.QE * i["THm"] / MH))):
        n = P[p]["np_tot"]
        v = P[p]["vxp"][:n]
        vt = i["vt0"][p]
        vbx = vb * np.cos(i["th_B"])
        g = np.linspace(max(0.0, vbx - 4 * vt), vbx + 4 * vt, 20001)
        f = g * np.exp(-(g - vbx) ** 2 / vt ** 2)
        mean = np.trapezoid(g * f, g) / np.trapezoid(f, g)
        sd = np.sqrt(np.trapezoid(g * g * f, g) / np.trapezoid(f, g) - mean ** 2)
        assert abs(v[:, 3].mean() - mean) < 5 * sd / np.sqrt(n), p
        assert np.all(v[:, 0] >= 0) and np.all(v[:, 0] <= v[:, 3] * DT * (1 + 1e-12))
        assert abs(v[:, 4].mean() - vb * np.sin(i["th_B"])) < 5 * vt / np.sqrt(2 * n)


def test_potential_tally_is_exact_on_a_linear_phi():
    """flag_diffsrc = 1: P_loss(1) -= Nm q phi(x, y), phi bilinearly
    interpolated; on a linear phi the interpolation is exact."""
    d, i, P = domain(), inj(1, flag_diffsrc=1), parts()
    X, Y = np.meshgrid((np.arange(N1 + 3) - 1) * H, (np.arange(N2 + 3) - 1) * H)
    phi = 3.0 + 200.0 * X - 50.0 * Y
    out = O.part_injection(d, i, P, 3, phi=phi)
    for p in (1, 2):
        v = P[p]["vxp"][:P[p]["np_tot"]]
        want = -(i["Nm"][p] * i["charge"][p] * (3.0 + 200.0 * v[:, 0] - 50.0 * v[:, 1])).sum()
        assert abs(out["P_loss1"][p] - want) <= 1e-9 * np.abs(i["Nm"][p] * i["charge"][p] * 3.0) * len(v)


def test_bounded_retries():
    """A region that lies entirely in walls: the oracle returns -1 where the
    reference would retry forever."""
    d = domain()
    i = inj(1, xl_pow=21.5 * H, xr_pow=19.2 * H, yl_pow=11.5 * H, yr_pow=9.2 * H)
    assert O.part_injection(d, i, parts(), 1, max_tries=1000)["rc"] == -1
