"""Parity at the benchmark workloads (jd_lepic_amd.workloads), one step of
the hot path in main.f90:689-1090 order.

argon-128 (128 x 128 cells, e-/Ar+ at 200 ppc, 6.55 M particles), FP64:
  * every phase through its C-ABI entry point against the oracle on the
    same state (the GPU's own input to that phase): rhs, E field, push and
    sort bit-exact; phi within 1e-12 with the same V-cycle and per-visit
    sweep counts; deposition within 1e-12; dens_red exact; MCC check mode
    with the same events, ran2 state and particles (velocities within 1e-13,
    device and host libm);
  * the whole step through lepic_step against the oracle's chain of the same
    routines from the same deposited state.
argon-512 (52.4 M) and the ion-source domain (1024 x 512, four species,
about 79 M), FP64 per phase against the oracle, and the mixed-precision
form against the oracle on the mixed context's own state: densities 2e-5,
phi 1e-6, E 1e-5 of the largest magnitude; positions 1e-5 of a cell and
velocities 1e-5 relative after the push, the same survivors; production MCC
rates within 3 sigma.
argon-1024 (210 M) and argon-2048 (1.007 B particles), mixed precision, at
full size: particle counts against the wall-loss and MCC tallies, charge
and energy bookkeeping, the sort's cell-end list against the sorted
particles, stability of the sort on a device-side sample against a stable
host argsort, MCC rates within 3 sigma of the cross sections evaluated on a
sample of the electrons."""
import numpy as np
import pytest

import oracle as O

L = pytest.importorskip("jd_lepic_amd")
from jd_lepic_amd import tables, workloads as WL  # noqa: E402

pytestmark = pytest.mark.gpu


def domain(W):
    return O.Domain(n1=W["n1"], n2=W["n2"], h1=W["h"], h2=W["h"], dt=W["dt"], bcnd=W["bcnd"], ngrid=W["ngrid"],
                    Bi=W["Bi"], nB=(W["n1"], W["n2"]), hB=(W["h"], W["h"]))


def oracle_table(W):
    chem = W["chem"]
    return O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                      chem.masses(), chem.ntype_all)


def rel(a, b):
    s = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else float(np.max(np.abs(a)))


def say(msg):
    print(f"[configs] {msg}", flush=True)


def expected_events(chem, table, v, m_inc, Nc, N, ngas, dt, ns_coll=10):
    """Per-reaction expected event counts of Nc candidates drawn from the
    particles v (or a uniform sample of them): Nc * n_gas * <sigma_r(E) |v|> /
    nu_omp, with the target's thermal motion neglected (for electrons on a
    300 K gas it changes the rate by ~1e-4)."""
    masses = chem.masses()
    tgt = abs(masses[chem.reactions[0].reactants[1]])
    mu = m_inc * tgt / (m_inc + tgt)
    speed = np.linalg.norm(v, axis=1)
    E = 0.5 * mu * speed * speed / O.QE
    nu_omp = float(np.float32(Nc) / np.float32(N)) / (ns_coll * dt)
    out = []
    for c in range(table["ncol"]):
        sig = np.interp(E, table["sig_Er"], table["sig"][c])
        out.append(Nc * ngas * np.mean(sig * speed) / nu_omp)
    return np.array(out)


# ---------------------------------------------------------------------------
# FP64: each phase against the oracle on the same state
# ---------------------------------------------------------------------------
def fp64_phases(W, it=11, iseed=123456):
    dom = domain(W)
    n1, n2, h = W["n1"], W["n2"], W["h"]
    nsp = len(W["species"])
    charges = [s["charge"] for s in W["species"]]
    say(f"{W['name']}: load + deposit")
    with WL.build_context(W, L.FP64) as ctx:
        nps = [ctx.get_grid(L.GRID_NP, p) for p in range(1, nsp + 1)]
        parts0 = [ctx.download(p) for p in range(1, nsp + 1)]
        for p in range(nsp):
            ref = O.charge_deposition(dom, W["species"][p]["Nm"], parts0[p])
            assert rel(nps[p], ref) <= 1e-12, f"deposition of species {p + 1}"
        # rhs (calc_rho.f90:37-77): bit-exact
        ctx.calc_rhs()
        rhs = ctx.get_grid(L.GRID_RHS)
        assert np.array_equal(rhs, O.calc_rhs(n1, n2, charges, nps))
        # Poisson (iterative_solver.f90 / mg.f90 / sors.f90) from a cold start
        say(f"{W['name']}: Poisson")
        info = ctx.iterative_solver(eps=1e-10, start=L.START_COLD, log_cap=1 << 16)
        phi = ctx.get_grid(L.GRID_PHI)
        ref = O.iterative_solver(np.zeros((n2 + 3, n1 + 3)), rhs, W["bcnd"], (h, h), n1, n2, W["ng"], eps=1e-10)
        assert info["vcycles"] == ref["vcycles"]
        assert np.array_equal(info["sweeps"], ref["sweeps"][:len(info["sweeps"])])
        assert rel(phi, ref["phi"]) <= 1e-12
        # E field (Efield.f90): bit-exact from the GPU's phi
        ctx.calc_efield()
        E = ctx.get_grid(L.GRID_E)
        assert np.array_equal(E, O.calc_efield(n1, n2, (h, h), phi, W["bcnd"], np.zeros((n2 + 3, n1 + 3, 2))))
        # sort (sorting.f90:92-251): vxp and Plist bit-exact
        say(f"{W['name']}: sort")
        sorted_parts = []
        for p in range(1, nsp + 1):
            before = ctx.download(p)
            pl = ctx.part_sorting(p, want_plist=True)
            after = ctx.download(p)
            v_ref, pl_ref = O.part_sorting(n1, n2, (h, h), before)
            assert np.array_equal(after, v_ref), f"sorted particles of species {p}"
            assert np.array_equal(pl, pl_ref), f"Plist of species {p}"
            sorted_parts.append(after)
        # dens_red + MCC in check mode (collisions.f90)
        say(f"{W['name']}: dens_red + MCC (check mode)")
        mx = ctx.dens_red()
        reds = [ctx.get_grid(L.GRID_NPRED, p) for p in range(1, nsp + 1)]
        ref_red, ref_mx = O.dens_red(n1, n2, W["bcnd"], nps)
        for p in range(nsp):
            assert np.array_equal(reds[p], ref_red[p]) and mx[p] == ref_mx[p]
        out = ctx.collisions(it, mode=L.MCC_CHECK, iseed=iseed)
        after_mcc = [ctx.download(p) for p in range(1, nsp + 1)]
        flags = [ctx.get_flag_dead(p, after_mcc[p - 1].shape[0]) for p in range(1, nsp + 1)]
        parts = {}
        for p in range(nsp):
            cap = sorted_parts[p].shape[0] + sorted_parts[p].shape[0] // 10 + 1024
            v = np.zeros((cap, 6))
            v[:sorted_parts[p].shape[0]] = sorted_parts[p]
            parts[p + 1] = dict(vxp=v, np_tot=sorted_parts[p].shape[0], flag_dead=np.zeros(cap, np.int8))
        ref = O.collisions(dom, WL.oracle_setup(W), oracle_table(W), parts, reds, iseed)
        assert out["seed"] == ref["seed"]
        assert np.array_equal(out["events"], ref["events"]) and ref["events"].sum() > 0
        for p in range(nsp):
            r = parts[p + 1]["vxp"][:parts[p + 1]["np_tot"]]
            g = after_mcc[p]
            assert g.shape == r.shape
            assert np.array_equal(g[:, :3], r[:, :3])
            np.testing.assert_allclose(g[:, 3:], r[:, 3:], rtol=1e-13, atol=1e-13 * np.abs(r[:, 3:]).max())
            assert np.array_equal(flags[p], parts[p + 1]["flag_dead"][:r.shape[0]])
        # push + deposit (part_expmover.f90) of every species: bit-exact
        say(f"{W['name']}: push + deposit")
        for p in range(1, nsp + 1):
            ctx.reset_tallies()
            ctx.part_mover(p, deposit=True)
            got = ctx.download(p)
            rho = ctx.get_grid(L.GRID_NP, p)
            sp = W["species"][p - 1]
            ref = O.part_mover(dom, sp["charge"], sp["mass"], sp["Nm"], E, after_mcc[p - 1], flags[p - 1],
                               tag_neg=W["tag_neg"] if W["tag_neg"] == p else 0)
            assert got.shape[0] == ref["np_tot"]
            assert np.array_equal(got, ref["vxp"]), f"pushed particles of species {p}"
            t = ctx.mover_tally(p, W["ngrid"])
            assert np.array_equal(t["p_mac"][0], ref["p_mac"][0])
            assert rel(rho, O.charge_deposition(dom, sp["Nm"], ref["vxp"])) <= 1e-12


def test_argon128_fp64_phases_match_oracle():
    fp64_phases(WL.CONFIGS["argon-128"]())


def test_argon128_fp64_step_matches_oracle_chain():
    """lepic_step at it = 11 (sort and MCC fire) against the oracle's routines
    chained in the same order from the same deposited state."""
    W = WL.CONFIGS["argon-128"]()
    dom = domain(W)
    n1, n2, h = W["n1"], W["n2"], W["h"]
    charges = [s["charge"] for s in W["species"]]
    with WL.build_context(W, L.FP64) as ctx:
        nps = [ctx.get_grid(L.GRID_NP, p) for p in (1, 2)]
        parts0 = [ctx.download(p) for p in (1, 2)]
        r = ctx.step(11, mcc_mode=L.MCC_CHECK, iseed=98765, start=L.START_COLD)
        phi = ctx.get_grid(L.GRID_PHI)
        rho = [ctx.get_grid(L.GRID_NP, p) for p in (1, 2)]
        got = [ctx.download(p) for p in (1, 2)]
    rhs = O.calc_rhs(n1, n2, charges, nps)
    ref = O.iterative_solver(np.zeros((n2 + 3, n1 + 3)), rhs, W["bcnd"], (h, h), n1, n2, W["ng"], eps=1e-10)
    assert r["vcycles"] == ref["vcycles"]
    assert rel(phi, ref["phi"]) <= 1e-12
    E = O.calc_efield(n1, n2, (h, h), ref["phi"], W["bcnd"], np.zeros((n2 + 3, n1 + 3, 2)))
    sorted_parts = [O.part_sorting(n1, n2, (h, h), v)[0] for v in parts0]
    reds, _ = O.dens_red(n1, n2, W["bcnd"], nps)
    parts = {}
    for p in (1, 2):
        v = sorted_parts[p - 1]
        cap = v.shape[0] + v.shape[0] // 10 + 1024
        vv = np.zeros((cap, 6))
        vv[:v.shape[0]] = v
        parts[p] = dict(vxp=vv, np_tot=v.shape[0], flag_dead=np.zeros(cap, np.int8))
    mref = O.collisions(dom, WL.oracle_setup(W), oracle_table(W), parts, reds, 98765)
    assert r["iseed"] == mref["seed"]
    for p in (1, 2):
        sp = W["species"][p - 1]
        n = parts[p]["np_tot"]
        ref = O.part_mover(dom, sp["charge"], sp["mass"], sp["Nm"], E, parts[p]["vxp"][:n],
                           parts[p]["flag_dead"][:n])
        g = got[p - 1]
        assert g.shape == ref["vxp"].shape
        scale = np.abs(ref["vxp"]).max(axis=0)
        assert np.all(np.abs(g - ref["vxp"]) <= 1e-12 * scale)
        assert rel(rho[p - 1], O.charge_deposition(dom, sp["Nm"], ref["vxp"])) <= 1e-12


def test_argon512_fp64_phases_match_oracle():
    fp64_phases(WL.CONFIGS["argon-512"]())


def test_ion_source_fp64_phases_match_oracle():
    fp64_phases(WL.CONFIGS["ion-source"]())


# ---------------------------------------------------------------------------
# mixed precision against the oracle on the mixed context's state
# ---------------------------------------------------------------------------
def mixed_phases(W, it=21, seed=11):
    dom = domain(W)
    n1, n2, h = W["n1"], W["n2"], W["h"]
    nsp = len(W["species"])
    charges = [s["charge"] for s in W["species"]]
    say(f"{W['name']} mixed: load + deposit")
    with WL.build_context(W, L.MIXED) as ctx:
        nps = [ctx.get_grid(L.GRID_NP, p) for p in range(1, nsp + 1)]
        parts0 = [ctx.download(p) for p in range(1, nsp + 1)]
        for p in range(nsp):
            ref = O.charge_deposition(dom, W["species"][p]["Nm"], parts0[p])
            assert rel(nps[p], ref) <= 2e-5, f"mixed deposition of species {p + 1}"
        ctx.calc_rhs()
        rhs = ctx.get_grid(L.GRID_RHS)
        info = ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        phi = ctx.get_grid(L.GRID_PHI)
        ref = O.iterative_solver(np.zeros((n2 + 3, n1 + 3)), rhs, W["bcnd"], (h, h), n1, n2, W["ng"], eps=1e-10)
        assert rel(phi, ref["phi"]) <= 1e-6 and info["res"] <= 1e-10
        ctx.calc_efield()
        E = ctx.get_grid(L.GRID_E)
        assert rel(E, O.calc_efield(n1, n2, (h, h), ref["phi"], W["bcnd"], np.zeros((n2 + 3, n1 + 3, 2)))) <= 1e-5
        say(f"{W['name']} mixed: sort")
        for p in range(1, nsp + 1):
            pl = ctx.part_sorting(p, want_plist=True)
            v = ctx.download(p)
            cell = (np.floor(v[:, 1] / h).astype(np.int64)) * n1 + np.floor(v[:, 0] / h).astype(np.int64)
            # cells from the FP32 offsets can round onto the next cell's edge
            # in FP64; the cell-end list must bracket every particle
            assert np.all(np.diff(pl.astype(np.int64)) >= 0) and pl[-1] == v.shape[0]
            assert np.mean(np.diff(cell) >= 0) > 0.99999
        say(f"{W['name']} mixed: MCC (production)")
        ctx.dens_red()
        N_e = ctx.count(1)
        e_sample = ctx.download_range(1, 0, N_e)
        out = ctx.collisions(it, mode=L.MCC_PROD, seed=seed)
        assert out["missed"] == 0
        table = tables.ordering(W["chem"])
        me = W["species"][0]["mass"]
        exp = expected_events(W["chem"], table, e_sample[:, 3:], me, out["Nc"][0], N_e, W["ngas"], W["dt"])
        for c in range(table["ncol"]):
            if table["col_info"][c][2] != 1:        # electron reactions only (reactant 1 = e-)
                continue
            sd = np.sqrt(max(exp[c], 1.0))
            assert abs(out["events"][c] - exp[c]) < 3 * sd + 1, (c, out["events"][c], exp[c])
        say(f"{W['name']} mixed: push + deposit")
        for p in range(1, nsp + 1):
            before = ctx.download(p)
            flags = ctx.get_flag_dead(p, before.shape[0])
            ctx.part_mover(p, deposit=True)
            got = ctx.download(p)
            rho = ctx.get_grid(L.GRID_NP, p)
            sp = W["species"][p - 1]
            ref = O.part_mover(dom, sp["charge"], sp["mass"], sp["Nm"], E, before, flags,
                               tag_neg=W["tag_neg"] if W["tag_neg"] == p else 0)
            assert abs(got.shape[0] - ref["np_tot"]) <= 2, (got.shape[0], ref["np_tot"])
            if got.shape[0] == ref["np_tot"]:
                r = ref["vxp"]
                assert np.max(np.abs(got[:, :2] - r[:, :2])) <= 1e-5 * h
                vs = np.abs(r[:, 3:]).max()
                assert np.max(np.abs(got[:, 3:] - r[:, 3:])) <= 1e-5 * vs
            assert rel(rho, O.charge_deposition(dom, sp["Nm"], got)) <= 2e-5


def test_argon512_mixed_against_oracle():
    mixed_phases(WL.CONFIGS["argon-512"]())


def test_ion_source_mixed_against_oracle():
    mixed_phases(WL.CONFIGS["ion-source"]())


# ---------------------------------------------------------------------------
# full size: properties that do not need the oracle to run the whole state
# ---------------------------------------------------------------------------
def _void_rows(v):
    """Rows (x, y, vx) as opaque 24-byte keys for exact matching."""
    a = np.ascontiguousarray(v[:, [0, 1, 3]])
    return a.view(np.dtype((np.void, 24))).ravel()


def full_size_properties(W, it=11, seed=5, sample=1 << 20):
    n1, n2, h = W["n1"], W["n2"], W["h"]
    nsp = len(W["species"])
    kq = np.where(W["bcnd"] != -1, 2.0, 1.0)
    say(f"{W['name']}: load")
    with WL.build_context(W, L.MIXED) as ctx:
        counts0 = [ctx.count(p) for p in range(1, nsp + 1)]
        assert counts0 == WL.load_counts(W)
        for p in range(nsp):
            t = ctx.species_totals(p + 1)
            assert t["count"] == counts0[p]
            assert t["charge"] == pytest.approx(counts0[p] * W["species"][p]["Nm"] * W["species"][p]["charge"],
                                                rel=1e-12)
        # step 1: rhs -> Poisson -> E -> push + deposit (no sort, no MCC)
        say(f"{W['name']}: step 1")
        ctx.calc_rhs()
        info = ctx.iterative_solver(eps=1e-10, start=L.START_COLD)
        assert info["res"] <= 1e-10 and info["vcycles"] > 0
        ctx.calc_efield()
        ctx.reset_tallies()
        e_before = sum(ctx.species_totals(q)["kinetic"] for q in range(1, nsp + 1))
        for q in range(1, nsp + 1):
            ctx.part_mover(q, deposit=True)
        for q in range(1, nsp + 1):
            t = ctx.mover_tally(q, W["ngrid"])
            lost = int(round(t["p_mac"][0].sum()))
            assert ctx.count(q) == counts0[q - 1] - lost
            tot = ctx.species_totals(q)
            sp = W["species"][q - 1]
            assert tot["charge"] == pytest.approx(ctx.count(q) * sp["Nm"] * sp["charge"], rel=1e-12)
            # the density integrates to the particle count (kq = 2 on wall
            # nodes doubles their share)
            rho = ctx.get_grid(L.GRID_NP, q)
            total = (rho / kq).sum() * h * h * 1.0 / sp["Nm"]
            assert total == pytest.approx(ctx.count(q), rel=1e-5)
        e_after = sum(ctx.species_totals(q)["kinetic"] for q in range(1, nsp + 1))
        lost_e = sum(ctx.mover_tally(q, W["ngrid"])["P_loss"][0] for q in range(1, nsp + 1))
        assert abs(e_after + lost_e - e_before) < 1e-2 * e_before
        counts1 = [ctx.count(q) for q in range(1, nsp + 1)]
        # step 2 (it % 10 == 1): rhs -> Poisson -> E -> sort -> MCC
        say(f"{W['name']}: step 2 sort")
        ctx.calc_rhs()
        info = ctx.iterative_solver(eps=1e-10, start=L.START_PREV)
        assert info["res"] <= 1e-10
        ctx.calc_efield()
        # sort with the cell-end list; stability on a device-side sample: the
        # particles of a contiguous pre-sort slice against a stable host
        # argsort of their cells
        N = counts1[0]
        first = N // 3
        pre = ctx.download_range(1, first, sample)
        pl = ctx.part_sorting(1, want_plist=True).astype(np.int64)
        assert np.all(np.diff(pl) >= 0) and pl[-1] == N
        for q in range(2, nsp + 1):
            ctx.part_sorting(q)
        cell_pre = np.floor(pre[:, 1] / h).astype(np.int64) * n1 + np.floor(pre[:, 0] / h).astype(np.int64)
        lo, hi = int(cell_pre.min()), int(cell_pre.max())
        a, b = int(pl[lo]), int(pl[hi + 1])
        assert b - a < 8 * sample
        post = ctx.download_range(1, a, b - a)
        kp = _void_rows(post)
        order_post = np.argsort(kp)
        j = np.searchsorted(kp[order_post], _void_rows(pre))
        pos = order_post[np.minimum(j, len(kp) - 1)]
        assert np.array_equal(post[pos], pre)                 # every sampled particle found once
        pos = pos + a
        assert np.array_equal(np.searchsorted(pl, pos, side="right") - 1, cell_pre)
        order = np.argsort(cell_pre, kind="stable")
        same = np.diff(cell_pre[order]) == 0
        assert np.all(np.diff(pos[order])[same] > 0)          # stable within each cell
        # production MCC rates on a sample of the electrons
        say(f"{W['name']}: step 2 MCC")
        ctx.dens_red()
        e_sample = ctx.download_range(1, N // 2, min(N // 2, 4 * sample))
        out = ctx.collisions(it, mode=L.MCC_PROD, seed=seed)
        assert out["missed"] == 0
        table = tables.ordering(W["chem"])
        exp = expected_events(W["chem"], table, e_sample[:, 3:], W["species"][0]["mass"], out["Nc"][0], N,
                              W["ngas"], W["dt"])
        for c in range(table["ncol"]):
            sd = np.sqrt(max(exp[c], 1.0)) + exp[c] * 3.0 / np.sqrt(e_sample.shape[0])
            assert abs(out["events"][c] - exp[c]) < 3 * sd + 1, (c, out["events"][c], exp[c])
        for q in range(nsp):
            assert ctx.count(q + 1) == counts1[q] + out["created"][q]


def test_argon1024_full_size_properties():
    full_size_properties(WL.CONFIGS["argon-1024"]())


def test_argon2048_full_size_properties():
    full_size_properties(WL.CONFIGS["argon-2048"]())
