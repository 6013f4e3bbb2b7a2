"""The oracle's surface emitter (lo_flux_emitter) against the reference's own
part_flux_emitter (part_flux_emitter.f90 compiled by oracle/Makefile), bit
for bit, and against what the emission must do: the emitted count is the
current js*zmax*Lgrd*ns_Hm*dt/(Nm*qe) up to the rounding draw, every
particle leaves an electrode cell toward the plasma, and the normal speed
follows the flux Maxwellian (<v_n^2> = vt^2)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle as O  # noqa: E402
import refharness as R  # noqa: E402

N1, N2, H = 64, 48, 5e-5
MH = 1.00794 * O.AMU
XG1 = 39 * H                   # node 40: the electrode's upstream face


def geometry():
    """Box walls (label 1); an electrode slab (label 2) over node columns
    40..44 with an aperture over node rows 20..30."""
    bc = O.box_bcnd(N1, N2)
    for ix in range(40, 45):
        for iy in range(2, N2 + 1):
            if not 20 <= iy <= 30:
                bc[iy, ix] = 2
    return O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=5e-12, bcnd=bc, zmax=0.32)


def emitter(n_icp=1, chamfer=False, jne=10.0):
    ymax = N2 * H
    return dict(jne=jne, Lg=0.9 * ymax, Lh=0.1 * ymax, xg1=XG1, THm=1.0, n_holes=0 if n_icp == 1 else 1,
                n_icp=n_icp, ptype_fx=3, tag_neg=3, tag_b=4 if chamfer else 0, flag_chmfrd=1 if chamfer else 0,
                ns_Hm=1, chmfrd_array=(XG1, XG1 + 2 * H, 0.3 * ymax, 0.4 * ymax),
                Nm={1: 1e3, 2: 1e3, 3: 1e3, 4: 1e3}, mass={1: O.ME, 2: MH, 3: MH, 4: MH})


def particles(seed, n0=5, extra=8000):
    rng = np.random.default_rng(seed)
    parts = {}
    for p in (1, 2, 3, 4):
        v = np.zeros((n0 + extra, 6))
        v[:n0] = rng.uniform(0.1, 0.2, (n0, 6))
        parts[p] = dict(vxp=v, np_tot=n0, src=np.zeros((N2 + 3, N1 + 3)))
    return parts


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("n_icp,chamfer", [(1, False), (2, False), (1, True)])
def test_emitter_matches_reference(n_icp, chamfer):
    dom, em = geometry(), emitter(n_icp, chamfer)
    parts = particles(1)
    ref, ss1, cnt_b, ploss4, seed_r, Lgrd_r = R.flux_emitter(dom, em, {p: dict(d) for p, d in parts.items()},
                                                             8642)
    out = O.flux_emitter(dom, em, parts, 8642)
    assert out["rc"] == 0 and out["seed"] == seed_r and out["Lgrd"] == Lgrd_r
    assert out["cnt_b"].tolist() == cnt_b.tolist() and out["Nh"] > 1000
    for p in (1, 2, 3, 4):
        assert parts[p]["np_tot"] == ref[p]["np_tot"]
        n = parts[p]["np_tot"]
        assert np.array_equal(parts[p]["vxp"][:n], ref[p]["vxp"][:n]), p
        assert np.array_equal(parts[p]["src"], ss1[p - 1])
        assert out["P_loss4"][p] == ploss4[p - 1]
    if chamfer:
        assert parts[4]["np_tot"] > 5 and cnt_b[1] == parts[4]["np_tot"] - 5


def test_emitter_current_and_flux_maxwellian():
    dom, em = geometry(), emitter(1, False, jne=40.0)
    parts = particles(2, extra=10000)
    out = O.flux_emitter(dom, em, parts, 777)
    assert out["rc"] == 0
    Lgrd = em["n_icp"] * (em["Lg"] - em["Lh"] * em["n_holes"])
    dNh = 10 * em["jne"] * dom.zmax * Lgrd * dom.dt / (em["Nm"][3] * O.QE)
    assert out["Lgrd"] == Lgrd and abs(out["dNh"] / dNh - 1) < 1e-14
    n = parts[3]["np_tot"] - 5
    assert n == out["Nh"] and int(dNh) <= n <= int(dNh) + 1
    assert parts[3]["src"].sum() == n
    v = parts[3]["vxp"][5:5 + n]
    vt = np.sqrt(2 * O.QE * em["THm"] / MH)
    # the plasma is on the left of the slab: every particle moves to -x
    assert np.all(v[:, 3] < 0)
    m2 = np.mean((v[:, 3] / vt) ** 2)
    assert abs(m2 - 1.0) < 5 / np.sqrt(n)
    # tangential pair Maxwellian with component variance vt^2/2
    assert abs(np.var(v[:, 4]) / (vt * vt / 2) - 1) < 5 * np.sqrt(2 / n)
    # launched from the electrode face, spread over one ns_Hm*dt flight
    assert np.all(v[:, 0] <= XG1) and np.all(v[:, 0] >= XG1 + v[:, 3] * dom.dt * (1 + 1e-12))
    # y outside the aperture rows
    iy = (v[:, 1] / H).astype(int) + 1
    assert not np.any((iy >= 21) & (iy <= 28))
    # energy tally
    e = 0.5 * em["Nm"][3] * MH * (v[:, 3:] ** 2).sum()
    assert abs(out["P_loss4"][3] / e - 1) < 1e-12


def test_emitter_gives_up_without_an_electrode():
    dom = O.Domain(n1=N1, n2=N2, h1=H, h2=H, dt=5e-12, bcnd=O.box_bcnd(N1, N2), zmax=0.32)
    parts = particles(3)
    out = O.flux_emitter(dom, emitter(), parts, 5, max_tries=1000)
    assert out["rc"] == -1
