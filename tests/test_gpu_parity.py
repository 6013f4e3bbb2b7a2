"""GPU parity: every phase through the C ABI against the oracle on the same
seeded inputs.  FP64 push, E field, rhs, sort: bit-exact.  Deposition: the
GPU sums in another order, so 1e-12 relative to the largest node.  Poisson
FP64: same V-cycle and sweep counts, phi within 1e-12 relative.  Mixed
precision: FP32 tolerances stated per test.  MCC check mode: identical events
and ran2 state; velocities within 1e-13 relative (device and host libm may
differ in the last bits of log/cos/sin/acos)."""
import numpy as np
import pytest

import oracle as O
import scenarios as S

L = pytest.importorskip("jd_lepic_amd")
pytestmark = pytest.mark.gpu

E_CHG, E_MASS = -O.QE, O.ME
I_CHG, I_MASS = O.QE, 39.95 * O.AMU


def make_ctx(dom, species, ng=1, precision=L.FP64, track_z=1, phi=None, tail=0):
    ctx = L.Context(dom.n1, dom.n2, dom.h1, dom.h2, dom.dt, species, zmax=dom.zmax, ng=ng,
                    ngrid=dom.ngrid, precision=precision, flag_B_pos=dom.flag_B_pos, track_z=track_z,
                    mg_tail_nodes=tail)
    ctx.set_boundary(dom.bcnd, np.zeros((dom.n2 + 3, dom.n1 + 3)) if phi is None else phi)
    if dom.Bi is not None:
        ctx.set_bfield(dom.Bi, dom.nB, dom.hB)
    return ctx


@pytest.mark.parametrize("Bmode", ["map", "const", "coarse", None])
@pytest.mark.parametrize("species", ["e", "ion"])
def test_mover_fp64_bit_exact(Bmode, species):
    dom = S.box_domain(40, 28, Bmode=Bmode, Bpeak=0.05, electrode=(Bmode == "map"))
    chg, m = (E_CHG, E_MASS) if species == "e" else (I_CHG, I_MASS)
    vt = 1.0e6 if species == "e" else 3.0e3
    parts = S.particles(dom, 30000, vt, seed=21)
    Ei = S.random_E(dom, 5e4, seed=22)
    fd = (np.random.default_rng(23).random(len(parts)) < 0.02).astype(np.int8)
    Nm = 1234.5
    with make_ctx(dom, [L.Species(chg, m, Nm, 40000)]) as ctx:
        ctx.upload(1, parts)
        ctx.set_flag_dead(1, fd)
        ctx.set_grid(L.GRID_E, Ei)
        ctx.part_mover(1, deposit=True)
        got = ctx.download(1)
        t = ctx.mover_tally(1, dom.ngrid)
        rho = ctx.get_grid(L.GRID_NP, 1)
        sink = ctx.get_grid(L.GRID_SINK, 1)
        flags = ctx.get_flag_dead(1, len(parts))
    ref = O.part_mover(dom, chg, m, Nm, Ei, parts, fd)
    assert got.shape[0] == ref["np_tot"] < len(parts)
    assert np.array_equal(got, ref["vxp"])
    np.testing.assert_allclose(t["P_loss"], ref["P_loss"], rtol=1e-12, atol=0)
    assert np.array_equal(t["p_mac"][0], ref["p_mac"][0])
    np.testing.assert_allclose(t["p_mac"][1], ref["p_mac"][1], rtol=1e-12, atol=0)
    assert np.array_equal(sink, ref["ss2D"])
    assert not flags.any()
    rho_ref = O.charge_deposition(dom, Nm, ref["vxp"])
    assert np.abs(rho - rho_ref).max() <= 1e-12 * np.abs(rho_ref).max()


def test_deposition_only_fp64():
    dom = S.box_domain(64, 48, Bmode=None, electrode=True)
    parts = S.particles(dom, 200000, 1.0, seed=24, edge_frac=0.0)
    parts = parts[np.lexsort(((parts[:, 0] / dom.h1).astype(int), (parts[:, 1] / dom.h2).astype(int)))]
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 77.0, 200000)]) as ctx:
        ctx.upload(1, parts)
        ctx.charge_deposition(1)
        rho = ctx.get_grid(L.GRID_NP, 1)
    ref = O.charge_deposition(dom, 77.0, parts)
    assert np.abs(rho - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("electrode", [False, True])
def test_efield_bit_exact(electrode):
    dom = S.box_domain(48, 36, Bmode=None, electrode=electrode)
    phi = np.random.default_rng(25).normal(0, 10.0, (39, 51))
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 16)]) as ctx:
        ctx.set_grid(L.GRID_PHI, phi)
        ctx.calc_efield()
        E = ctx.get_grid(L.GRID_E)
    ref = O.calc_efield(48, 36, (dom.h1, dom.h2), phi, dom.bcnd, np.zeros((39, 51, 2)))
    assert np.array_equal(E, ref)


def test_rhs_bit_exact():
    dom = S.box_domain(32, 32, Bmode=None)
    rng = np.random.default_rng(26)
    nps = [rng.random((35, 35)) * 1e16, rng.random((35, 35)) * 1e16]
    sp = [L.Species(E_CHG, E_MASS, 1.0, 16), L.Species(I_CHG, I_MASS, 1.0, 16)]
    with make_ctx(dom, sp) as ctx:
        for p in (1, 2):
            ctx.set_grid(L.GRID_NP, nps[p - 1], ptype=p)
        ctx.calc_rhs()
        rhs = ctx.get_grid(L.GRID_RHS)
    ref = O.calc_rhs(32, 32, [E_CHG, I_CHG], nps)
    assert np.array_equal(rhs, ref)


def poisson_case(n1, n2, seed):
    """Grounded box, top wall at 10 V, and an electrode block at 25 V whose
    edges sit on odd node indices (so injection keeps it on coarse levels;
    a plate on even indices vanishes from the coarse masks and the
    reference's V-cycle then diverges, in the oracle as well)."""
    dom = S.box_domain(n1, n2, Bmode=None)
    bc = dom.bcnd.copy()
    phi = np.zeros((n2 + 3, n1 + 3))
    phi[n2 + 1:, :] = 10.0
    a1, a2 = n1 // 4 + 1, n2 // 4 + 1
    bc[a2:a2 + n2 // 8 + 1, a1:a1 + n1 // 8 + 1] = 2
    phi[bc == 2] = 25.0
    dom.bcnd, dom.ngrid = bc, 2
    rng = np.random.default_rng(seed)
    b = rng.normal(0, 1e-3, (n2 + 1, n1 + 1))
    b[bc[1:n2 + 2, 1:n1 + 2] != -1] = 0.0
    return dom, phi, b


# tail = mg_tail_nodes: 0 keeps the library default (levels up to 5000 nodes in
# the single-workgroup kernel); 100 puts every level but the last few on the
# multi-block sweep kernels, including the ng=1 SOR-only solve
@pytest.mark.parametrize("n,ng,tail", [(32, 5, 0), (128, 7, 0), (256, 8, 0), (64, 1, 0),
                                       (128, 7, 100), (256, 8, 100), (64, 1, 100)])
def test_poisson_fp64_matches_oracle(n, ng, tail):
    dom, phi, b = poisson_case(n, n, 27 + n)
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 16)], ng=ng, phi=phi, tail=tail) as ctx:
        ctx.set_grid(L.GRID_RHS, b)
        info = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        got = ctx.get_grid(L.GRID_PHI)
    ref = O.iterative_solver(phi, b, dom.bcnd, (dom.h1, dom.h2), n, n, ng, eps=1e-10)
    assert info["vcycles"] == ref["vcycles"]
    assert np.array_equal(info["sweeps"], ref["sweeps"][:len(info["sweeps"])])
    scale = np.abs(ref["phi"]).max()
    assert np.abs(got - ref["phi"]).max() <= 1e-12 * scale
    assert info["res"] <= 1e-10


@pytest.mark.parametrize("tail", [0, 100])
def test_poisson_fp64_warm_start_sequence(tail):
    """Three solves in a row, each starting from the previous phi with a new
    right-hand side (the time loop's pattern), against the oracle doing the
    same; the state carried between solves must not leak into the next."""
    n, ng = 128, 7
    dom, phi, b = poisson_case(n, n, 5)
    rng = np.random.default_rng(9)
    ref_phi = phi.copy()
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 16)], ng=ng, phi=phi, tail=tail) as ctx:
        for step in range(3):
            bb = b * (1.0 + 0.05 * step) + rng.normal(0, 1e-5, b.shape) * (dom.bcnd[1:n + 2, 1:n + 2] == -1)
            ctx.set_grid(L.GRID_RHS, bb)
            info = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
            got = ctx.get_grid(L.GRID_PHI)
            ref = O.iterative_solver(ref_phi, bb, dom.bcnd, (dom.h1, dom.h2), n, n, ng, eps=1e-10)
            ref_phi = ref["phi"]
            assert info["vcycles"] == ref["vcycles"], step
            assert np.array_equal(info["sweeps"], ref["sweeps"][:len(info["sweeps"])]), step
            assert np.abs(got - ref_phi).max() <= 1e-12 * np.abs(ref_phi).max(), step


# 192 x 96: levels of 193 x 97 ... nodes, whose last tile of a row or column
# takes the side's odd node, and (tail=100) small-tile levels whose grid
# transfers run inside the sweep launches
@pytest.mark.parametrize("tail", [0, 100])
@pytest.mark.parametrize("n1,n2,ng", [(128, 128, 7), (192, 96, 6)])
def test_poisson_mixed_within_1e6_of_fp64(n1, n2, ng, tail):
    dom, phi, b = poisson_case(n1, n2, 31)
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 16)], ng=ng, precision=L.MIXED, phi=phi,
                  tail=tail) as ctx:
        ctx.set_grid(L.GRID_RHS, b)
        info = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        got = ctx.get_grid(L.GRID_PHI)
        # a second solve from the converged phi with a changed rhs
        b2 = b * 1.1
        ctx.set_grid(L.GRID_RHS, b2)
        info2 = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        got2 = ctx.get_grid(L.GRID_PHI)
    ref = O.iterative_solver(phi, b, dom.bcnd, (dom.h1, dom.h2), n1, n2, ng, eps=1e-10)
    assert info["res"] <= 1e-10
    assert np.abs(got - ref["phi"]).max() <= 1e-6 * np.abs(ref["phi"]).max()
    ref2 = O.iterative_solver(ref["phi"], b2, dom.bcnd, (dom.h1, dom.h2), n1, n2, ng, eps=1e-10)
    assert info2["res"] <= 1e-10
    assert np.abs(got2 - ref2["phi"]).max() <= 1e-6 * np.abs(ref2["phi"]).max()


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_sort_matches_oracle(precision):
    dom = S.box_domain(64, 40, Bmode=None)
    parts = S.particles(dom, 100000, 1.0, seed=32, edge_frac=0.0)
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 100000)], precision=precision) as ctx:
        ctx.upload(1, parts)
        before = ctx.download(1)
        pl = ctx.part_sorting(1, want_plist=True)
        got = ctx.download(1)
    ref, pref = O.part_sorting(64, 40, (dom.h1, dom.h2), before)
    assert np.array_equal(got, ref) and np.array_equal(pl, pref)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
@pytest.mark.parametrize("n1,n2", [(32, 30), (2048, 1024), (2048, 2048)])
def test_sort_digit_passes(n1, n2, precision):
    """The radix sort splits the key (cell index, plus the empty-slot key
    above it) into the fewest digits of at most 11 bits: one pass for 960
    cells, two of 11 bits for 2^21 cells, three (8, 8, 7 bits) for the
    2^22 cells of 2048 x 2048."""
    print(f"sort passes {n1}x{n2} precision={precision}", flush=True)
    dom = S.box_domain(n1, n2, Bmode=None)
    parts = S.particles(dom, 60000, 1.0, seed=37, edge_frac=0.0)
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 60000)], precision=precision) as ctx:
        ctx.upload(1, parts)
        before = ctx.download(1)
        pl = ctx.part_sorting(1, want_plist=True)
        got = ctx.download(1)
    ref, pref = O.part_sorting(n1, n2, (dom.h1, dom.h2), before)
    assert np.array_equal(got, ref) and np.array_equal(pl, pref)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_sort_of_nearly_sorted_particles(precision):
    """Particles sorted once and then moved by up to a cell and a half, as
    between two sorts of a run: each segment's cells then span fewer than
    256 compact keys, the segment sort's single-digit path; the first three
    segments' particles share one cell (a key span of zero)."""
    dom = S.box_domain(64, 40, Bmode=None)
    parts = S.particles(dom, 100000, 1.0, seed=41, edge_frac=0.0)
    sp = [L.Species(E_CHG, E_MASS, 1.0, 100000)]
    with make_ctx(dom, sp, precision=precision) as ctx:
        ctx.upload(1, parts)
        ctx.part_sorting(1)
        moved = ctx.download(1).copy()
        rng = np.random.default_rng(43)
        L1, L2 = dom.n1 * dom.h1, dom.n2 * dom.h2
        moved[:, 0] = np.clip(moved[:, 0] + rng.uniform(-1.5, 1.5, len(moved)) * dom.h1, 0.1 * dom.h1, L1 - 0.1 * dom.h1)
        moved[:, 1] = np.clip(moved[:, 1] + rng.uniform(-1.5, 1.5, len(moved)) * dom.h2, 0.1 * dom.h2, L2 - 0.1 * dom.h2)
        moved[:3072, 0] = 10.5 * dom.h1
        moved[:3072, 1] = 7.5 * dom.h2
        ctx.upload(1, moved)
        before = ctx.download(1)
        pl = ctx.part_sorting(1, want_plist=True)
        got = ctx.download(1)
    ref, pref = O.part_sorting(64, 40, (dom.h1, dom.h2), before)
    assert np.array_equal(got, ref) and np.array_equal(pl, pref)


@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_sort_of_unpacked_segments(precision):
    """The mover leaves partly filled segments; the sort reads them as they
    are and must give what sorting the packed species gives (and, in FP64,
    what the oracle's mover + sort give)."""
    dom = S.box_domain(64, 40, Bmode=None)
    parts = S.particles(dom, 120000, 2e6, seed=35, edge_frac=0.05)
    Ei = S.random_E(dom, 1e4, seed=36)
    sp = [L.Species(E_CHG, E_MASS, 1.0, 120000)]
    with make_ctx(dom, sp, precision=precision) as a, make_ctx(dom, sp, precision=precision) as b:
        for c in (a, b):
            c.upload(1, parts)
            c.set_grid(L.GRID_E, Ei)
            c.part_mover(1, deposit=False)
        moved = b.download(1)                 # packs b
        assert moved.shape[0] < parts.shape[0]    # some particles were lost at the walls
        pa = a.part_sorting(1, want_plist=True)   # a is still segmented
        pb = b.part_sorting(1, want_plist=True)
        got_a, got_b = a.download(1), b.download(1)
    assert np.array_equal(got_a, got_b) and np.array_equal(pa, pb)
    if precision == L.FP64:
        ref = O.part_mover(dom, E_CHG, E_MASS, 1.0, Ei, parts)
        sref, pref = O.part_sorting(64, 40, (dom.h1, dom.h2), ref["vxp"])
        assert np.array_equal(got_a, sref) and np.array_equal(pa, pref)


def test_mixed_push_close_to_fp64():
    dom = S.box_domain(64, 48, Bmode="map", Bpeak=0.02)
    parts = S.particles(dom, 50000, 1e6, seed=33, edge_frac=0.0)
    parts[:, 0] = np.clip(parts[:, 0], 3 * dom.h1, (dom.n1 - 3) * dom.h1)
    parts[:, 1] = np.clip(parts[:, 1], 3 * dom.h2, (dom.n2 - 3) * dom.h2)
    Ei = S.random_E(dom, 3e4, seed=34)
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 10.0, 50000)], precision=L.MIXED, track_z=1) as ctx:
        ctx.upload(1, parts)
        ctx.set_grid(L.GRID_E, Ei)
        ctx.part_mover(1, deposit=True)
        got = ctx.download(1)
        rho = ctx.get_grid(L.GRID_NP, 1)
    ref = O.part_mover(dom, E_CHG, E_MASS, 10.0, Ei, parts)
    assert got.shape[0] == ref["np_tot"]
    np.testing.assert_allclose(got[:, :2], ref["vxp"][:, :2], rtol=0, atol=2e-6 * dom.h1)
    vs = np.abs(ref["vxp"][:, 3:]).max()
    np.testing.assert_allclose(got[:, 3:], ref["vxp"][:, 3:], rtol=0, atol=2e-6 * vs)
    rho_ref = O.charge_deposition(dom, 10.0, ref["vxp"])
    assert np.abs(rho - rho_ref).max() <= 1e-4 * np.abs(rho_ref).max()


def test_load_uniform_statistics():
    dom = S.box_domain(64, 32, Bmode=None, electrode=True)
    vt = 1.0e6
    with make_ctx(dom, [L.Species(E_CHG, E_MASS, 1.0, 64 * 32 * 20)]) as ctx:
        ctx.load_uniform(1, 20, vt, 0.0, 64 * dom.h1, seed=5)
        p = ctx.download(1)
    walls = (dom.bcnd >= 1)
    cells_all_wall = walls[1:-2, 1:-2] & walls[1:-2, 2:-1] & walls[2:-1, 1:-2] & walls[2:-1, 2:-1]
    assert p.shape[0] == 20 * (64 * 32 - cells_all_wall.sum())
    ix = (p[:, 0] / dom.h1).astype(int)
    iy = (p[:, 1] / dom.h2).astype(int)
    key = ix + 64 * iy
    assert np.all(np.diff(key) >= 0)                     # loaded cell-sorted
    assert np.allclose(p[:, 3:].var(axis=0), vt * vt / 2, rtol=0.03)
