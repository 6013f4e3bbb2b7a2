"""Averaged moments on the GPU (part_moments, calc_avg, phi/E averages)
against the oracle on the same particles, densities and sink maps."""
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
N, H, DT = 32, 5e-5, 5e-12


def setup(prec):
    sp = [L.Species(-L.QE, L.ME, 2e3, 300000), L.Species(L.QE, 40 * L.AMU, 2e3, 300000)]
    ctx = L.Context(N, N, H, H, DT, sp, ng=5, precision=prec, track_z=0)
    bc = O.box_bcnd(N, N)
    ctx.set_boundary(bc, np.zeros((N + 3, N + 3)))
    ctx.load_uniform(1, 30, 2e6, 0.0, N * H, 5)      # hot electrons: some reach the walls
    ctx.load_uniform(2, 30, 2e3, 0.0, N * H, 5)
    return ctx, bc


@pytest.mark.parametrize("prec,size_na", [(L.FP64, 1), (L.FP64, 2), (L.MIXED, 1), (L.MIXED, 4)])
def test_moments_match_oracle(prec, size_na):
    ctx, bc = setup(prec)
    masses, Nms = [L.ME, 40 * L.AMU], [2e3, 2e3]
    dom = O.Domain(n1=N, n2=N, h1=H, h2=H, dt=DT, bcnd=bc)
    want = None
    with ctx:
        ctx.calc_rhs()
        ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        ctx.calc_efield()
        phi, E = ctx.get_grid(L.GRID_PHI).copy(), ctx.get_grid(L.GRID_E).copy()
        rng = np.random.default_rng(size_na)
        for rep in range(2):
            for p in (1, 2):
                ctx.part_mover(p, deposit=True)
            parts = [ctx.download(p) for p in (1, 2)]
            nps = [ctx.get_grid(L.GRID_NP, p).copy() for p in (1, 2)]
            sinks = [ctx.get_grid(L.GRID_SINK, p).copy() for p in (1, 2)]
            if rep == 0:
                assert sinks[0].sum() > 0         # the case exercises the sink map
            # source maps as injection/emission/ionization would leave them
            srcs = [np.floor(rng.uniform(0, 3, nps[0].shape)) for _ in (1, 2)]
            for p in (1, 2):
                ctx.set_grid(L.GRID_SRC, srcs[p - 1], p)
            pm = [O.part_moments(dom, Nms[p], parts[p], size_na) for p in range(2)]
            want = O.calc_avg(dom, masses, Nms, nps, pm, sinks, size_na, want, srcs=srcs)
            ctx.moments_accumulate(size_na)
            assert ctx.get_grid(L.GRID_SINK, 1).sum() == 0.0   # cleared after the average
            assert ctx.get_grid(L.GRID_SRC, 2).sum() == 0.0
        got, phi_avg, count = ctx.moments()
    assert count == 2
    np.testing.assert_array_equal(phi_avg[..., 0], 2 * phi)
    np.testing.assert_array_equal(phi_avg[..., 1:].ravel(), 2 * E.ravel())
    # FP64: same expressions, atomics change the summation order only; the
    # mixed form's positions are cell + FP32 offset, exact to 6e-8 cell
    tol = 1e-11 if prec == L.FP64 else 2e-5
    for q in range(7):
        scale = np.abs(want[..., q]).max()
        if scale == 0.0:
            assert np.abs(got[..., q]).max() == 0.0, q
            continue
        err = np.abs(got[..., q] - want[..., q]).max() / scale
        assert err <= tol, (q, err)
