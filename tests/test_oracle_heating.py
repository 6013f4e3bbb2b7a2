"""The oracle's electron heating (lo_eheating) against the reference's own
eheating (e_heating.f90 compiled by oracle/Makefile), bit for bit, and
against what the heating must do statistically."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle as O  # noqa: E402
import refharness as R  # noqa: E402

H, N1 = 5e-5, 64


def case(flag_c, flag_gpp, eheat_type, n=20000, seed=3, nudt=0.3):
    rng = np.random.default_rng(seed)
    L = N1 * H
    v = np.column_stack([rng.uniform(0, L, n), rng.uniform(0, L, n), np.zeros(n),
                         rng.normal(0, 3e5, (n, 3))])
    heat = dict(flag_c=flag_c, flag_gpp=flag_gpp, eheat_type=eheat_type, xa=0.4 * L, dr=0.2 * L,
                yl_pow=0.2 * L, yr_pow=0.7 * L, ymax=L, nudt=nudt, dt=5e-12, Pabs_cor=0.8, Nm=1e3,
                mass=O.ME, vt=np.array([8e5, 1.1e6, 1.4e6]))
    return v, heat


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("flag_c,flag_gpp,eheat_type", [(0, 0, 1), (1, 0, 1), (1, 1, 0), (0, 1, 0)])
def test_eheating_matches_reference(flag_c, flag_gpp, eheat_type):
    v, heat = case(flag_c, flag_gpp, eheat_type)
    got, dP, s = O.eheating(v, heat, 987654)
    ref, dPr, sr = R.eheating(v, heat, 987654)
    assert np.array_equal(got, ref) and dP == dPr and s == sr
    assert not np.array_equal(got, v)


def test_eheating_statistics():
    """Uniform rate: the heated fraction of the particles inside the slit is
    nudt; heated in-plane velocities (eheat_type 1) are Maxwellian with
    component variance vt^2/2 of their strip (vz takes the second value of a
    pair drawn for an earlier particle, possibly of another strip, as
    e_heating.f90:80-86 does); outside nothing changes."""
    v, heat = case(0, 0, 1, n=200000, nudt=0.25)
    got, dP, _ = O.eheating(v, heat, 4242)
    L = N1 * H
    inside = (np.abs(v[:, 0] - heat["xa"]) <= heat["dr"]) & (v[:, 1] >= heat["yl_pow"]) & (v[:, 1] <= heat["yr_pow"])
    changed = np.any(got != v, axis=1)
    assert not np.any(changed & ~inside)
    m, p = inside.sum(), 0.25
    assert abs(changed.sum() - p * m) < 4 * np.sqrt(m * p * (1 - p))
    strip = (3 * v[:, 1] / L).astype(int)
    for k in range(3):
        sel = changed & (strip == k)
        var = got[sel, 3:5].var(axis=0).mean()
        vt = heat["vt"][k]
        assert abs(var / (vt * vt / 2) - 1) < 0.06, (k, var / (vt * vt / 2))
    e_new = 0.5 * heat["Nm"] * heat["mass"] * (got[changed, 3:] ** 2).sum()
    e_old = 0.5 * heat["Nm"] * heat["mass"] * (v[changed, 3:] ** 2).sum()
    assert abs(dP - (e_new - e_old)) <= 1e-9 * abs(e_new)
