"""Charge and energy totals through the C ABI (lepic_species_totals,
lepic_field_energy) against numpy on the same data: the particles as the
library returns them (so the mixed form's FP32 velocities enter exactly)
and the E grid as set.  Only the summation order differs: 1e-12 relative.
After a push that loses particles at the walls the totals follow the
survivors, and the lost kinetic energy is the mover's wall tally."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

L = pytest.importorskip("jd_lepic_amd")
pytestmark = pytest.mark.gpu

ZMAX = 0.37


def context(dom, nm, cap, precision):
    ctx = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, [L.Species(-O.QE, O.ME, nm, cap)], zmax=ZMAX,
                    ngrid=dom.ngrid, precision=precision)
    ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)))
    return ctx


def totals_np(v, nm):
    return v.shape[0], 0.5 * nm * O.ME * float(np.sum(v[:, 3:] ** 2)), nm * -O.QE * v.shape[0]


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_species_totals_before_and_after_a_push(precision):
    print(f"totals precision={precision}", flush=True)
    dom = S.box_domain(64, 40, Bmode=None)
    parts = S.particles(dom, 50000, 2e6, seed=41, edge_frac=0.05)
    Ei = S.random_E(dom, 1e4, seed=42)
    nm = 123.0
    with context(dom, nm, 60000, precision) as ctx:
        ctx.upload(1, parts)
        t0, v0 = ctx.species_totals(1), ctx.download(1)
        ctx.set_grid(L.GRID_E, Ei)
        ctx.reset_tallies()
        ctx.part_mover(1, deposit=False)
        t1 = ctx.species_totals(1)           # read from the mover's partly filled segments
        lost = ctx.mover_tally(1, dom.ngrid)["P_loss"][0]
        v1 = ctx.download(1)
    for t, v in ((t0, v0), (t1, v1)):
        n, ke, q = totals_np(v, nm)
        assert t["count"] == n
        assert abs(t["kinetic"] - ke) <= 1e-12 * ke
        assert abs(t["charge"] - q) <= 1e-12 * abs(q)
    assert t1["count"] < t0["count"]
    # energy of the lost particles, taken at their pushed velocities
    assert lost > 0 and lost < t0["kinetic"]


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_field_energy(precision):
    dom = S.box_domain(48, 36, Bmode=None, electrode=True)
    Ei = S.random_E(dom, 3e4, seed=43)
    with context(dom, 1.0, 16, precision) as ctx:
        ctx.set_grid(L.GRID_E, Ei)
        got = ctx.field_energy()
        E = ctx.get_grid(L.GRID_E)
    m = dom.bcnd == -1
    want = 0.5 * L.EPS0 * float(np.sum(E[m] ** 2)) * dom.h1 * dom.h2 * ZMAX
    assert want > 0 and abs(got - want) <= 1e-12 * want
