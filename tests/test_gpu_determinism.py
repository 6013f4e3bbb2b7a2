"""Run-to-run reproducibility.  The deposition accumulates fixed-point
integers (integer addition is associative), so the density does not depend
on the order in which the atomics land: two runs on the same input give the
same rho bit for bit, and so do whole steps (rhs -> Poisson -> E -> sort ->
push + deposit) in FP64 and in mixed precision, down to the V-cycle sweep
counts that the density feeds."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

L = pytest.importorskip("jd_lepic_amd")
pytestmark = pytest.mark.gpu


def ctx_for(dom, precision, cap, ng):
    sp = [L.Species(-O.QE, O.ME, 250.0, cap), L.Species(O.QE, 39.95 * O.AMU, 250.0, cap)]
    ctx = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, sp, ng=ng, ngrid=dom.ngrid, precision=precision)
    ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)))
    ctx.set_bfield(dom.Bi, dom.nB, dom.hB)
    return ctx


def test_fp64_deposit_bit_identical_across_runs():
    dom = S.box_domain(96, 64, Bmode="map", Bpeak=0.02)
    parts = S.particles(dom, 1_500_000, 1.0e6, seed=71, edge_frac=0.0)
    Ei = S.random_E(dom, 3e4, seed=72)
    out = []
    for run in range(2):
        with ctx_for(dom, L.FP64, 1_600_000, 6) as ctx:
            ctx.upload(1, parts)
            ctx.set_grid(L.GRID_E, Ei)
            ctx.charge_deposition(1)
            r0 = ctx.get_grid(L.GRID_NP, 1)
            ctx.charge_deposition(1)
            assert np.array_equal(ctx.get_grid(L.GRID_NP, 1), r0)    # same context, again
            ctx.part_mover(1, deposit=True)
            out.append((r0, ctx.get_grid(L.GRID_NP, 1), ctx.download(1)))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    ref = O.charge_deposition(dom, 250.0, parts)
    assert np.abs(out[0][0] - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_steps_bit_identical_across_runs(precision):
    dom = S.box_domain(128, 96, Bmode="map", Bpeak=0.01)
    N = 800_000
    e = S.particles(dom, N, 1.0e6, seed=73, edge_frac=0.01)
    i = e.copy()
    i[:, 3:] = np.random.default_rng(74).normal(0, 300.0, (N, 3))
    runs = []
    for run in range(2):
        with ctx_for(dom, precision, N + 1024, 6) as ctx:
            ctx.upload(1, e)
            ctx.upload(2, i)
            for p in (1, 2):
                ctx.charge_deposition(p)
            infos = [ctx.step(it, mcc_mode=-1, nsort=2) for it in range(1, 5)]
            runs.append(dict(vc=[r["vcycles"] for r in infos], res=[r["res"] for r in infos],
                             phi=ctx.get_grid(L.GRID_PHI), rho=[ctx.get_grid(L.GRID_NP, p) for p in (1, 2)],
                             parts=[ctx.download(p) for p in (1, 2)]))
    a, b = runs
    assert a["vc"] == b["vc"] and a["res"] == b["res"]
    assert np.array_equal(a["phi"], b["phi"])
    for p in range(2):
        assert np.array_equal(a["rho"][p], b["rho"][p])
        assert np.array_equal(a["parts"][p], b["parts"][p])
