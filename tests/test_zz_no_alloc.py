"""lepic_part_sorting_noalloc (static device scratch, no allocation) against
lepic_part_sorting on the same particles, through the C ABI: the sorted
particles and the cell-end list Plist must be identical bit for bit, on two
calls in a row (the second with a smaller species, over the scratch the
first call left behind), in FP64 and mixed contexts.  Inputs beyond the
static scratch (more than 512 segments) are rejected with
LEPIC_ERR_CAPACITY and leave the particles untouched."""
import ctypes as C

import numpy as np
import pytest

import scenarios as S

L = pytest.importorskip("jd_lepic_amd")
pytestmark = pytest.mark.gpu

ERR_CAPACITY = -4


def ctx_pair(dom, precision, cap):
    sp = [L.Species(-L.QE, L.ME, 100.0, cap)]
    out = []
    for _ in range(2):
        c = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, sp, ng=1, precision=precision)
        c.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)))
        out.append(c)
    return out


def sort_via(ctx, fn, with_plist=True):
    pl = np.full(ctx.n1 + ctx.n1 * ctx.n2 + 2, -7, np.int32) if with_plist else None
    ptr = pl.ctypes.data_as(C.POINTER(C.c_int32)) if with_plist else None
    rc = fn(ctx._ctx, 1, ptr)
    return rc, pl


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_noalloc_sort_matches_allocating_sort(precision):
    dom = S.box_domain(96, 80)
    first = S.particles(dom, 400_000, 1.0e6, seed=81, edge_frac=0.02)
    second = S.particles(dom, 150_000, 1.0e6, seed=82, edge_frac=0.02)
    a, b = ctx_pair(dom, precision, 400_000)
    try:
        lib = a.lib
        for parts in (first, second):
            a.upload(1, parts)
            b.upload(1, parts)
            # a push scatters the particles across segments before the sort
            a.set_grid(L.GRID_E, S.random_E(dom, 3e4, seed=83))
            b.set_grid(L.GRID_E, S.random_E(dom, 3e4, seed=83))
            a.part_mover(1, deposit=False)
            b.part_mover(1, deposit=False)
            rc_a, pl_a = sort_via(a, lib.lepic_part_sorting)
            rc_b, pl_b = sort_via(b, lib.lepic_part_sorting_noalloc)
            assert rc_a == 0 and rc_b == 0, (rc_a, rc_b, b.lib.lepic_last_error(b._ctx).decode())
            assert np.array_equal(pl_a, pl_b)
            assert a.count(1) == b.count(1)
            assert np.array_equal(a.download(1), b.download(1))
            # and without the list
            rc_a, _ = sort_via(a, lib.lepic_part_sorting, with_plist=False)
            rc_b, _ = sort_via(b, lib.lepic_part_sorting_noalloc, with_plist=False)
            assert rc_a == 0 and rc_b == 0
            assert np.array_equal(a.download(1), b.download(1))
    finally:
        a.close()
        b.close()


def test_noalloc_sort_rejects_species_beyond_scratch():
    dom = S.box_domain(64, 64)
    parts = S.particles(dom, 600_000, 1.0e6, seed=84, edge_frac=0.0)
    a, b = ctx_pair(dom, L.MIXED, 600_000)
    try:
        b.upload(1, parts)
        before = b.download(1)
        rc, _ = sort_via(b, b.lib.lepic_part_sorting_noalloc)
        assert rc == ERR_CAPACITY
        assert "static scratch" in b.lib.lepic_last_error(b._ctx).decode()
        assert np.array_equal(b.download(1), before)
    finally:
        a.close()
        b.close()
