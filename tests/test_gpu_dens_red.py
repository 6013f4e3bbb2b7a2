"""dens_red (calc_rho.f90:134-199) on the GPU against the oracle: np_red is the
deposited density on nodes 1..n+1 with zero ghosts (one partition: an exact
copy), np_mx the maximum over interior nodes (bcnd == -1), for every species,
in FP64 and mixed contexts, on a grid with an internal electrode (wall nodes
that carry density but are not interior)."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

L = pytest.importorskip("jd_lepic_amd")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_dens_red_matches_oracle(precision):
    dom = S.box_domain(48, 40, Bmode=None, electrode=True)
    e = S.particles(dom, 120000, 1.0, seed=61, edge_frac=0.05)
    i = S.particles(dom, 60000, 1.0, seed=62, edge_frac=0.05)
    sp = [L.Species(-O.QE, O.ME, 300.0, 130000), L.Species(O.QE, 39.95 * O.AMU, 600.0, 70000)]
    with L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, sp, ng=1, ngrid=dom.ngrid,
                   precision=precision) as ctx:
        ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)))
        ctx.upload(1, e)
        ctx.upload(2, i)
        for p in (1, 2):
            ctx.charge_deposition(p)
        nps = [ctx.get_grid(L.GRID_NP, p) for p in (1, 2)]
        mx = ctx.dens_red()
        reds = [ctx.get_grid(L.GRID_NPRED, p) for p in (1, 2)]
    ref_red, ref_mx = O.dens_red(dom.n1, dom.n2, dom.bcnd, nps)
    for p in range(2):
        assert np.array_equal(reds[p], ref_red[p])
        assert mx[p] == ref_mx[p] > 0.0
        # wall nodes of the electrode carry density but do not set the maximum
        walls = dom.bcnd[1:dom.n2 + 2, 1:dom.n1 + 2] >= 1
        assert nps[p][1:dom.n2 + 2, 1:dom.n1 + 2][walls].max() > 0.0
    # the deposited density itself against the oracle's deposition
    if precision == L.FP64:
        for p, (parts, Nm) in enumerate(((e, 300.0), (i, 600.0))):
            ref = O.charge_deposition(dom, Nm, parts)
            assert np.abs(nps[p] - ref).max() <= 1e-12 * np.abs(ref).max()
