"""Pins of the oracle's null-collision MCC (collisions.f90 needs MPI and
cannot be compiled here, so the exact event sequence is parity unpinned).
Pinned instead: the table build (closed forms), the RNG consumption per
candidate, the centre-of-mass kinematics of every event type, and the
elastic rate against n*sigma*<v> within 3 sigma."""
import numpy as np

import oracle as O
from mcc_setup import argon_case


def test_ordering_union_grid_and_sigv_max():
    dom, chem, tab, setup, parts, np_red = argon_case()
    allE = np.unique(np.concatenate([r.E for r in chem.reactions]))
    assert np.array_equal(tab["sig_Er"], allE)
    for c, r in enumerate(chem.reactions):
        np.testing.assert_allclose(tab["sig"][c], np.interp(allE, r.E, r.S), rtol=1e-12, atol=0)
        mu = O.ME * 39.95 * O.AMU / (O.ME + 39.95 * O.AMU)
        vr = np.sqrt(2.0 * allE * O.QE / mu)
        keep = allE >= 0.1
        assert np.isclose(tab["sigv_mx"][1][c], (vr * tab["sig"][c])[keep].max(), rtol=1e-15)
    assert tab["p_ncol"][1] == 3 and tab["p_ncol"][2] == 0
    assert list(tab["sig_list"][1][1:4]) == [1, 2, 3]


def _draws(log, ncand_first_odd=True):
    n = 1  # the Nc rounding draw of the electrons (ions have no reactions)
    for k, (ip, c) in enumerate(log):
        n += 1 + (4 if k % 2 == 0 else 2) + 1 + (4 if c > 0 else 0)
    return n


def test_rng_consumption_per_candidate():
    dom, chem, tab, setup, parts, np_red = argon_case(N=4000)
    seed0 = 123456
    out = O.collisions(dom, setup, tab, parts, np_red, seed0, log_cap=100000)
    log = out["log"]
    assert len(log) == out["Nc"][1] > 0
    assert (log[:, 1] > 0).any()
    _, s = O.ran2(seed0, _draws(log))
    assert s == out["seed"]


def _picked_once(log):
    ips, counts = np.unique(log[:, 0], return_counts=True)
    return set(ips[counts == 1].tolist())


def test_event_kinematics_neutral_at_rest():
    # Tgas = 0 puts the target at rest, so v_cm = m_e v / (m_e + M) and vr = |v|
    dom, chem, tab, setup, parts, np_red = argon_case(N=20000, Te=30.0, Tgas=0.0)
    v0 = parts[1]["vxp"][:, 3:].copy()
    n_e0, n_i0 = parts[1]["np_tot"], parts[2]["np_tot"]
    out = O.collisions(dom, setup, tab, parts, np_red, 777, log_cap=100000)
    log = out["log"]
    once = _picked_once(log)
    me, M = O.ME, 39.95 * O.AMU
    mu = me * M / (me + M)
    n_new = 0
    checked = {1: 0, 2: 0, 3: 0}
    for ip, c in log:
        if c == 2 or c == 3 or c == 1:
            pass
        if c == 0:
            continue
        if c == 3:
            n_new += 1
        if ip not in once:
            continue
        v = v0[ip - 1]
        vcm = me * v / (me + M)
        vn = parts[1]["vxp"][ip - 1, 3:]
        speed2 = np.dot(v, v)
        if c == 1:
            assert np.isclose(np.linalg.norm(vn - vcm), np.sqrt(speed2) * M / (me + M), rtol=1e-12)
        elif c == 2:
            Ee = 0.5 * mu * speed2 - 11.5 * O.QE
            sm = 1 / me + 1 / M
            assert np.isclose(np.dot(vn - vcm, vn - vcm), 2 * Ee / (sm * me * me), rtol=1e-12)
        checked[c] += 1
    assert checked[1] > 10 and checked[2] > 0
    # ionization: products appended in order, at the incident position,
    # zero CM momentum and CM energy 0.5 mu vr^2 - Eth
    assert parts[1]["np_tot"] == n_e0 + n_new and parts[2]["np_tot"] == n_i0 + n_new
    k = 0
    for ip, c in log:
        if c != 3:
            continue
        if ip in once:
            v = v0[ip - 1]
            vcm = me * v / (me + M)
            e1 = parts[1]["vxp"][ip - 1]
            e2 = parts[1]["vxp"][n_e0 + k]
            ion = parts[2]["vxp"][n_i0 + k]
            assert np.array_equal(e2[:3], e1[:3]) and np.array_equal(ion[:3], e1[:3])
            p = me * (e1[3:] - vcm) + me * (e2[3:] - vcm) + M * (ion[3:] - vcm)
            assert np.linalg.norm(p) < 1e-12 * me * np.linalg.norm(v)
            E = 0.5 * me * np.sum((e1[3:] - vcm) ** 2) + 0.5 * me * np.sum((e2[3:] - vcm) ** 2) \
                + 0.5 * M * np.sum((ion[3:] - vcm) ** 2)
            assert np.isclose(E, 0.5 * mu * np.dot(v, v) - 15.76 * O.QE, rtol=1e-11)
        k += 1
    assert k == n_new and n_new > 0


def test_elastic_rate_three_sigma():
    # constant sigma, target at rest: P(collision | candidate) = n sigma |v| / nu_max_OMP
    dom, chem, tab, setup, parts, np_red = argon_case(N=200000, Te=3.0, Tgas=0.0,
                                                       only_elastic=True, seed=5)
    N = parts[1]["np_tot"]
    speed = np.linalg.norm(parts[1]["vxp"][:N, 3:], axis=1)
    out = O.collisions(dom, setup, tab, parts, np_red, 4242, log_cap=1 << 20)
    Nc = out["Nc"][1]
    nu_omp = float(np.float32(Nc) / np.float32(N)) / (10.0 * dom.dt)
    p = 7.2429e19 * 1e-19 * speed.mean() / nu_omp
    ev = out["events"][0]
    sd = np.sqrt(Nc * p * (1 - p)) + Nc * p * speed.std() / speed.mean() / np.sqrt(N)
    assert abs(ev - Nc * p) < 3 * sd, (ev, Nc * p, sd)
