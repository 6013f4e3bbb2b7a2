"""The oracle's particle routines against the reference's own Fortran
routines (oracle/_ref), bit for bit, plus closed forms that do not need it."""
import numpy as np
import pytest

import oracle as O
import refharness as R
import scenarios as S

need_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (no reference checkout)")

E_CHG, E_MASS = -O.QE, O.ME
I_CHG, I_MASS = O.QE, 39.95 * O.AMU


def test_ran2_park_miller_known_answer():
    # Park & Miller (CACM 1988): seed 1 reaches 1043618065 after 10000 draws
    _, s = O.ran2(1, 10000)
    assert s == 1043618065


@need_ref
def test_ran2_and_gauss_match_reference():
    a, sa = O.ran2(123456, 5000)
    b, sb = R.ran2(123456, 5000)
    assert sa == sb and np.array_equal(a, b)
    rnd = np.random.default_rng(3).random((2000, 2))
    assert np.array_equal(O.load_gauss(1.3e6, rnd), R.load_gauss(1.3e6, rnd))


def test_load_gauss_moments():
    # vp = vt sqrt(-ln(1-r)) and a uniform angle: each component N(0, vt^2/2)
    vals, _ = O.ran2(99, 40000)
    v = O.load_gauss(2.0, vals.reshape(-1, 2))
    var = v.var(axis=0)
    assert np.allclose(var, 2.0, rtol=0.05)
    assert abs(v.mean()) < 0.03


@need_ref
@pytest.mark.parametrize("Bmode", ["map", "const", "coarse", None])
@pytest.mark.parametrize("species", ["e", "ion"])
def test_part_mover_matches_reference(Bmode, species):
    dom = S.box_domain(16, 12, Bmode=Bmode, Bpeak=0.05, electrode=(Bmode == "map"))
    chg, m = (E_CHG, E_MASS) if species == "e" else (I_CHG, I_MASS)
    vt = 1.0e6 if species == "e" else 3.0e3
    parts = S.particles(dom, 3000, vt, seed=5)
    Ei = S.random_E(dom, 5e4, seed=6)
    fd = (np.random.default_rng(8).random(len(parts)) < 0.02).astype(np.int8)
    Nm = 1234.5
    a = O.part_mover(dom, chg, m, Nm, Ei, parts, fd, tag_neg=1)
    b = R.part_mover(dom, chg, m, Nm, Ei, parts, fd, tag_neg=1)
    assert a["np_tot"] == b["np_tot"]
    assert a["np_tot"] < len(parts)  # some particles did leave
    assert np.array_equal(a["vxp"], b["vxp"])
    assert np.array_equal(a["P_loss"], b["P_loss"])
    assert np.array_equal(a["p_mac"], b["p_mac"])
    assert a["cnt_dead"] == b["cnt_dead"]
    assert np.array_equal(a["ss2D"], b["ss2D"])
    assert np.array_equal(a["flag_dead"], b["flag_dead"])


@need_ref
def test_flag_b_pos_matches_reference():
    dom = S.box_domain(8, 8, Bmode="map", Bpeak=0.05, flag_B_pos=1)
    parts = S.particles(dom, 500, 3e3, seed=9)
    Ei = S.random_E(dom, 1e4, seed=10)
    a = O.part_mover(dom, I_CHG, I_MASS, 1.0, Ei, parts)
    b = R.part_mover(dom, I_CHG, I_MASS, 1.0, Ei, parts)
    assert np.array_equal(a["vxp"], b["vxp"])


def test_leapfrog_uniform_field_closed_form():
    # no B: v' = v + (q/m) E dt exactly as 2*k1*E, x' = x + dt v'
    dom = S.box_domain(16, 16, Bmode=None)
    Ei = np.zeros((19, 19, 2))
    Ei[..., 0], Ei[..., 1] = 1.5e3, -2.5e3
    parts = S.particles(dom, 400, 1e5, seed=11, edge_frac=0.0)
    out = O.part_mover(dom, E_CHG, E_MASS, 1.0, Ei, parts)
    k1 = dom.dt * E_CHG / (2.0 * E_MASS)
    v = parts[:, 3:5] + np.array([2.0 * k1 * 1.5e3, 2.0 * k1 * -2.5e3])
    keep = np.ones(len(parts), bool)
    assert out["np_tot"] == len(parts)
    # bilinear weights sum to 1 up to rounding
    np.testing.assert_allclose(out["vxp"][:, 3:5], v, rtol=1e-13, atol=1e-9)
    np.testing.assert_allclose(out["vxp"][:, 0], parts[:, 0] + dom.dt * out["vxp"][:, 3], rtol=0, atol=1e-18)
    assert keep.all()


def test_boris_pure_b_rotation_closed_form():
    # E=0, constant B: |v_perp| kept; phase advance per step 2 atan(wc dt / 2)
    dom = S.box_domain(16, 16, Bmode="const", Bpeak=0.01)
    Bi = dom.Bi.copy()
    Bi[1, 1, :3] = (0.0, 0.0, 0.01)
    Bi[1, 1, 3] = 0.01
    dom.Bi = Bi
    Ei = np.zeros((19, 19, 2))
    L = 16 * dom.h1
    parts = np.array([[L / 2, L / 2, 0.1, 3e5, -1e5, 2e5]])
    out = O.part_mover(dom, E_CHG, E_MASS, 1.0, Ei, parts)
    v0, v1 = parts[0, 3:6], out["vxp"][0, 3:6]
    assert np.isclose(np.hypot(*v1[:2]), np.hypot(*v0[:2]), rtol=1e-14)
    assert v1[2] == v0[2]
    wc = O.QE * 0.01 / E_MASS
    dphi = np.arctan2(v1[1], v1[0]) - np.arctan2(v0[1], v0[0])
    expect = 2.0 * np.arctan(wc * dom.dt / 2.0)   # electrons rotate counterclockwise about +z
    assert np.isclose(dphi, expect, rtol=1e-10)


@need_ref
def test_charge_deposition_matches_reference():
    dom = S.box_domain(16, 12, Bmode=None, electrode=True)
    parts = S.particles(dom, 5000, 1e5, seed=12, edge_frac=0.0)
    kq = O.kq(dom)
    a = O.charge_deposition(dom, 987.0, parts, kq)
    b = R.charge_deposition(dom, 987.0, parts, kq)
    assert np.array_equal(a, b)


def test_charge_deposition_conserves_charge():
    # with kq = 1 everywhere, sum(np)*h1*h2*zmax = Nm * N (weights sum to 1)
    dom = S.box_domain(10, 10, Bmode=None)
    parts = S.particles(dom, 2000, 1.0, seed=13, edge_frac=0.0)
    ones = np.ones((13, 13))
    a = O.charge_deposition(dom, 5.0, parts, ones)
    assert np.isclose(a.sum() * dom.h1 * dom.h2 * dom.zmax, 5.0 * 2000, rtol=1e-12)


@need_ref
@pytest.mark.parametrize("electrode", [False, True])
def test_efield_matches_reference(electrode):
    dom = S.box_domain(16, 12, Bmode=None, electrode=electrode)
    rng = np.random.default_rng(14)
    phi = rng.normal(0, 10.0, (15, 19))
    E0 = rng.normal(0, 1.0, (15, 19, 2))   # the routine leaves ghost values alone
    a = O.calc_efield(16, 12, (dom.h1, dom.h2), phi, dom.bcnd, E0)
    b = R.calc_efield(16, 12, (dom.h1, dom.h2), phi, dom.bcnd, E0)
    assert np.array_equal(a, b)


def test_efield_quadratic_phi_closed_form():
    # central differences are exact for quadratics at interior nodes
    n1 = n2 = 12
    h = (1e-3, 2e-3)
    bc = O.box_bcnd(n1, n2)
    x = (np.arange(n1 + 3) - 1) * h[0]
    y = (np.arange(n2 + 3) - 1) * h[1]
    X, Y = np.meshgrid(x, y)
    phi = 3.0 * X ** 2 - 2.0 * X * Y + 5.0 * Y ** 2 + X
    E = O.calc_efield(n1, n2, h, phi, bc)
    inner = bc == -1
    np.testing.assert_allclose(E[..., 0][inner], -(6.0 * X - 2.0 * Y + 1.0)[inner], rtol=1e-9)
    np.testing.assert_allclose(E[..., 1][inner], -(-2.0 * X + 10.0 * Y)[inner], rtol=1e-9)


@need_ref
def test_part_sorting_matches_reference():
    dom = S.box_domain(16, 12, Bmode=None)
    parts = S.particles(dom, 4000, 1.0, seed=15, edge_frac=0.0)
    a, pa = O.part_sorting(16, 12, (dom.h1, dom.h2), parts)
    b, pb = R.part_sorting(16, 12, (dom.h1, dom.h2), parts)
    assert np.array_equal(a, b) and np.array_equal(pa, pb)


def test_part_sorting_is_stable_counting_sort():
    dom = S.box_domain(8, 6, Bmode=None)
    parts = S.particles(dom, 1500, 1.0, seed=16, edge_frac=0.0)
    key = (parts[:, 0] / dom.h1).astype(int) + 8 * (parts[:, 1] / dom.h2).astype(int)
    order = np.argsort(key, kind="stable")
    v, Plist = O.part_sorting(8, 6, (dom.h1, dom.h2), parts)
    assert np.array_equal(v, parts[order])
    counts = np.bincount(key, minlength=8 * 6)
    assert np.array_equal(Plist[1:8 * 6 + 1], np.cumsum(counts))
