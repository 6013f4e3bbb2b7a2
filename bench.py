"""Benchmark of the per-timestep hot path on one MI355X.

Workload (one GPU): 512 x 512 cells, dx = dy = 5e-5 m, grounded box, e- and
Ar+ (39.95 u) at 100 macroparticles per cell each (52,428,800 particles),
n0 = 1e16 m^-3, Maxwellian Te = 3 eV and Ti = 0.03 eV, Bz(x) Gaussian filter
of 10 mT peak at L/2 and width L/8, background Ar at 0.3 Pa / 300 K with
analytic cross sections (constant elastic 1e-19 m^2, excitation at 11.5 eV,
Lotz ionization at 15.76 eV), dt = 5e-12 s, MG Poisson (9 levels) to 1e-10,
sort and null-collision MCC every 10 steps.  Synthetic inputs from the
device loader (counter-based RNG, recorded seed).

One step = rhs -> Poisson -> E -> [sort] -> [MCC] -> push+deposit of both
species (main.f90:689-1090 order).  The headline value is particles/s of
push + deposit + MCC in the mixed-precision form with the learned warm start
(FP32 push/deposit with FP64 grid sums, FP32 MG smoothing in FP64
refinement, Poisson started from the network's guess when it lowers the
initial residual; the network is trained on the run's own first solves,
untimed); plain mixed and plain FP64 are measured on the same workload
beside it.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import jd_lepic_amd as L  # noqa: E402
from jd_lepic_amd import reactions as RX, tables  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per particle of one mixed mover launch at this workload, measured
# with rocprofv3 PMC counters (FETCH_SIZE and WRITE_SIZE, each scaled by what
# it reads for a 1 GiB device copy): 51.4 B against 48 algorithmic bytes
# (profiles/mover_mixed_traffic_512x512_100ppc.md)
TRAFFIC_BYTES_PER_PARTICLE = 51.4


class Hip:
    """HIP events on the library's stream (kernel timing) via libamdhip64."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        for n in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize", "hipEventElapsedTime",
                  "hipEventDestroy", "hipDeviceSynchronize"):
            getattr(self.lib, n).restype = C.c_int
        self.lib.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def event(self):
        e = C.c_void_p()
        assert self.lib.hipEventCreate(C.byref(e)) == 0
        return e

    def record(self, e, stream):
        assert self.lib.hipEventRecord(e, stream) == 0

    def elapsed(self, a, b):
        self.lib.hipEventSynchronize(b)
        ms = C.c_float()
        assert self.lib.hipEventElapsedTime(C.byref(ms), a, b) == 0
        return ms.value

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0

    def copy_once(self, nbytes):
        """One device-to-device copy of nbytes (hipMemcpy's copy kernel)."""
        a, b = C.c_void_p(), C.c_void_p()
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        assert self.lib.hipMalloc(C.byref(a), nbytes) == 0 and self.lib.hipMalloc(C.byref(b), nbytes) == 0
        assert self.lib.hipMemcpy(b, a, nbytes, 3) == 0          # hipMemcpyDeviceToDevice
        self.sync()
        self.lib.hipFree(a)
        self.lib.hipFree(b)
        return nbytes


def workload(n=512, ppc=100):
    """e-/Ar+ in a grounded n x n box at ppc per cell each, Bz filter field."""
    h = 5.0e-5
    L1 = n * h
    x = (np.arange(n + 3) - 1) * h
    Bz = 0.01 * np.exp(-(x - L1 / 2) ** 2 / (2 * (L1 / 8) ** 2))   # Bfield.f90:54-57
    Bi = np.zeros((n + 3, n + 3, 4))
    Bi[..., 2] = Bz[None, :]
    Bi[..., 3] = np.abs(Bi[..., 2])
    bcnd = -np.ones((n + 3, n + 3), np.int32)
    bcnd[:2, :] = bcnd[n + 1:, :] = bcnd[:, :2] = bcnd[:, n + 1:] = 1
    ngas = 0.3 / (1.380649e-23 * 300.0)
    chem = RX.argon(ngas)
    n0, Vc = 1.0e16, h * h * 1.0
    Nm = n0 * Vc / ppc
    vte = np.sqrt(2.0 * L.QE * 3.0 / L.ME)
    vti = np.sqrt(2.0 * L.QE * 0.03 / (39.95 * L.AMU))
    vtg = np.sqrt(2.0 * L.QE * 0.02585 / (39.95 * L.AMU))
    species = [dict(name="e", charge=-L.QE, mass=L.ME, Nm=Nm, vt=vte, loads=[(ppc, 0.0, L1)]),
               dict(name="i", charge=L.QE, mass=39.95 * L.AMU, Nm=Nm, vt=vti, loads=[(ppc, 0.0, L1)])]
    return dict(kind="e-Ar", n=n, n1=n, n2=n, h=h, Bi=Bi, bcnd=bcnd, chem=chem, Nm=Nm, ppc=ppc, vte=vte,
                vti=vti, vtg=vtg, dt=5.0e-12, ng=int(np.log2(n)), ngas=ngas, species=species,
                vt0=[vte, vti, vtg], ngrid=1, tag_neg=0,
                desc=f"{n}x{n} cells, e-/Ar+ {ppc} ppc each ({2 * n * n * ppc} particles), B filter 10 mT, "
                     "Ar 0.3 Pa analytic MCC, dt 5e-12, MG 1e-10")


def workload_ion_source(n1=1024, n2=512, ppc=124, strip=16):
    """Negative-ion-source-like domain: e-, H+, H2+, H- over background H2
    (analytic elastic, dissociation and ionization cross sections); density
    1e17 m^-3 over the driver half falling exponentially to 1e15 at the
    extraction wall (x = L), loaded in strips of `strip` columns with the
    strip's particles per cell; species fractions e- 0.9, H+ 0.5, H2+ 0.5,
    H- 0.1 of that density (quasi-neutral), one weight Nm for all, ppc the
    particles per cell of a species at unit fraction where the density is
    1e17; grounded box (label 1), a grid electrode of bars a few cells before
    the extraction wall (label 2), the extraction wall itself (label 3); Bz
    filter field of 10 mT at L/2."""
    h = 5.0e-5
    L1 = n1 * h
    bcnd = -np.ones((n2 + 3, n1 + 3), np.int32)
    bcnd[:2, :] = bcnd[n2 + 1:, :] = bcnd[:, :2] = 1
    bcnd[:, n1 + 1:] = 3
    c0 = n1 - 11 if (n1 - 11) % 2 == 1 else n1 - 10
    for r0 in range(5, n2 - 3, 8):
        bcnd[r0:r0 + 5, c0:c0 + 3] = 2
    x = (np.arange(n1 + 3) - 1) * h
    Bi = np.zeros((n2 + 3, n1 + 3, 4))
    Bi[..., 2] = 0.01 * np.exp(-(x - L1 / 2) ** 2 / (2 * (L1 / 8) ** 2))[None, :]
    Bi[..., 3] = np.abs(Bi[..., 2])
    n_hi, n_lo = 1.0e17, 1.0e15
    lam = (L1 / 2) / np.log(n_hi / n_lo)

    def density(xm):
        return n_hi if xm < L1 / 2 else n_hi * np.exp(-(xm - L1 / 2) / lam)

    Nm = n_hi * h * h * 1.0 / ppc
    ngas = 0.3 / (1.380649e-23 * 300.0)
    chem = RX.hydrogen(ngas)
    temps = {"e": 2.0, "Hp": 0.1, "H2p": 0.1, "Hm": 0.1}
    masses = {"e": L.ME, "Hp": 1.00794 * L.AMU, "H2p": 2.01588 * L.AMU, "Hm": 1.00794 * L.AMU}
    charges = {"e": -L.QE, "Hp": L.QE, "H2p": L.QE, "Hm": -L.QE}
    frac = {"e": 0.9, "Hp": 0.5, "H2p": 0.5, "Hm": 0.1}
    species, total = [], 0
    for name in ("e", "Hp", "H2p", "Hm"):
        loads = []
        for c in range(0, n1, strip):
            k = int(round(ppc * frac[name] * density((c + strip / 2) * h) / n_hi))
            if k > 0:
                loads.append((k, c * h, (c + strip) * h))
                total += k * strip * n2
        species.append(dict(name=name, charge=charges[name], mass=masses[name], Nm=Nm,
                            vt=np.sqrt(2.0 * L.QE * temps[name] / masses[name]), loads=loads))
    vt0 = [np.sqrt(2.0 * L.QE * (2.0 if k == 0 else sp.T) / abs(sp.mass)) for k, sp in enumerate(chem.species)]
    return dict(kind="ion-source", n=n1, n1=n1, n2=n2, h=h, Bi=Bi, bcnd=bcnd, chem=chem, Nm=Nm, ppc=ppc,
                dt=5.0e-12, ng=int(np.log2(min(n1, n2))), ngas=ngas, species=species, vt0=vt0, ngrid=3,
                tag_neg=4,
                desc=f"{n1}x{n2} cells, e-/H+/H2+/H- (about {total} particles), density 1e17 m^-3 over the "
                     "driver half falling exponentially to 1e15 at the extraction wall, grid electrode, "
                     "B filter 10 mT, H2 0.3 Pa analytic MCC, dt 5e-12, MG 1e-10")


def build_ctx(W, precision, seed=20261202):
    n1, n2 = W["n1"], W["n2"]
    sp = []
    for s in W["species"]:
        nload = sum(k * n2 * int(round((xr - xl) / W["h"])) for k, xl, xr in s["loads"])
        sp.append(L.Species(s["charge"], s["mass"], s["Nm"], int(nload * 1.05) + 4096))
    ctx = L.Context(n1, n2, W["h"], W["h"], W["dt"], sp, ng=W["ng"], ngrid=W["ngrid"], precision=precision,
                    track_z=0, tag_neg=W["tag_neg"])
    ctx.set_boundary(W["bcnd"], np.zeros((n2 + 3, n1 + 3)))
    ctx.set_bfield(W["Bi"], (n1, n2), (W["h"], W["h"]))
    for p, s in enumerate(W["species"], start=1):
        for k, xl, xr in s["loads"]:
            ctx.load_uniform(p, k, s["vt"], xl, xr, seed)
    ctx.set_reactions(W["chem"], tables.ordering(W["chem"]), vt0=W["vt0"])
    for p in range(1, len(W["species"]) + 1):
        ctx.charge_deposition(p)
    return ctx


def run_mode(W, precision, steps, warmup, hip, seed=7, learned=False, learn_steps=40, train_iters=300):
    """One mode of the step loop: learn_steps untimed steps, warmup untimed
    steps, then `steps` timed steps, so every mode times the same steps of
    the run.  With learned=True the first learn_steps solves record (drhs,
    dphi) samples, the network is trained on them (its time reported apart),
    and every later solve starts from the network's guess when that lowers
    the initial residual (the library skips the network for a growing
    back-off after a guess that gains too little)."""
    ctx = build_ctx(W, precision)
    ctx.kernel_timing(True)      # mover kernels bracketed by events inside the library
    stream = C.c_void_p()
    ctx.lib.lepic_stream(ctx._ctx, C.byref(stream))
    nsp = len(W["species"])
    ev = [hip.event() for _ in range(5 + 2 * nsp)]
    ev_end = ev[4 + 2 * nsp]
    acc = dict(field=0.0, sort=0.0, mcc=0.0, push=0.0, vcycles=0, push_kernel_ms=[[] for _ in range(nsp)],
               net_used=0, net_probed=0, res0_gain=[], train_s=0.0, loss=None)
    pre = learn_steps
    if learned:
        ctx.net_create(capacity=64, seed=seed)
        ctx.net_recording(True)
    # the first solve starts cold and converges from zero; later ones warm-start
    for it in range(1, pre + warmup + steps + 1):
        timed = it > pre + warmup
        if learned and it == pre + 1:
            ctx.net_recording(False)
            hip.sync()
            t = time.perf_counter()
            acc["loss"] = ctx.net_train(train_iters, batch=4, lr=2e-3)
            acc["train_s"] = time.perf_counter() - t
        if it == pre + warmup + 1:
            hip.sync()
            t0 = time.perf_counter()
        start = L.START_COLD if it == 1 else (L.START_NET if learned and it > pre else L.START_PREV)
        hip.record(ev[0], stream)
        ctx.calc_rhs()
        info = ctx.iterative_solver(eps=1e-10, ncycl=20000, start=start)
        ctx.calc_efield()
        hip.record(ev[1], stream)
        if it % 10 == 1:
            for p in range(1, nsp + 1):
                ctx.part_sorting(p)
        hip.record(ev[2], stream)
        if it > 1 and it % 10 == 1:
            ctx.dens_red()
            ctx.collisions(it, mode=L.MCC_PROD, seed=seed)
        hip.record(ev[3], stream)
        for p in range(1, nsp + 1):
            hip.record(ev[4 + 2 * (p - 1)], stream)
            ctx.part_mover(p, deposit=True)
            hip.record(ev[5 + 2 * (p - 1)], stream)
        hip.record(ev_end, stream)
        if timed:
            acc["field"] += hip.elapsed(ev[0], ev[1])
            acc["sort"] += hip.elapsed(ev[1], ev[2])
            acc["mcc"] += hip.elapsed(ev[2], ev[3])
            acc["push"] += hip.elapsed(ev[3], ev_end)
            acc["vcycles"] += info["vcycles"]
            if learned and info["res0_alt"] >= 0.0:
                # the network was evaluated: gain = residual of the previous-phi
                # start over that of the network's guess
                acc["net_probed"] += 1
                used = info["start_used"] == L.START_NET
                acc["net_used"] += used
                acc["res0_gain"].append(info["res0_alt"] / info["res0"] if used else info["res0"] / info["res0_alt"])
            for p in range(nsp):
                acc["push_kernel_ms"][p].append(ctx.kernel_ms(p + 1))
    hip.sync()
    wall = time.perf_counter() - t0
    counts = [ctx.count(p) for p in range(1, nsp + 1)]
    ctx.close()
    return acc, wall, counts


def _rel(a, b):
    """max |a - b| over max |b| (b the FP64 result)."""
    s = float(np.max(np.abs(b)))
    return float(np.max(np.abs(a - b)) / s) if s > 0 else 0.0


def _stocks(c, nsp):
    """Kinetic energy, field energy and charge of every species now."""
    t = [c.species_totals(q) for q in range(1, nsp + 1)]
    return dict(kinetic=sum(x["kinetic"] for x in t), field=c.field_energy(),
                charge=sum(x["charge"] for x in t), charge_e=abs(t[0]["charge"]),
                count=[x["count"] for x in t])


def accuracy_and_sor(W, steps=10, sor_solves=3):
    """Untimed checks beside the step loop.

    Mixed against FP64 from the same initial particles (the device loader
    with the same seed): max-norm differences, relative to the FP64 field's
    largest magnitude, of each species' density after the first deposition,
    and of the densities, phi and E after `steps` steps, and of the density
    and temperature maps (part_moments/calc_avg) averaged over those steps on
    nodes above a tenth of the peak density; and per precision the energy
    budget over those steps (kinetic + field energy at the start and end, the
    kinetic energy the walls took from the movers' tallies, and the relative
    drift of their sum) and the total charge.  The loader puts every species
    at the same positions, so the first right-hand side is zero up to
    rounding: its difference is given relative to one species' charge
    density instead.

    Then the Poisson problems of the last `sor_solves` steps (the FP64 run's
    right-hand sides) solved by red-black SOR alone (ng = 1, FP64; every call
    iterates to eps with the 10^6-sweep limit of sors.f90:63-64) beside the
    9-level MG in FP64 and in mixed precision: ms per solve and the phi
    difference."""
    n1, n2 = W["n1"], W["n2"]
    names = [sp["name"] for sp in W["species"]]
    nsp = len(names)
    ctx = {p: build_ctx(W, p) for p in (L.FP64, L.MIXED)}
    out, rhs_seq = {}, []
    stock0 = {}
    for c in ctx.values():
        c.moments_reset()
        c.reset_tallies()
    for it in range(1, steps + 1):
        for p, c in ctx.items():
            c.calc_rhs()
            if p == L.FP64:
                rhs_seq.append(c.get_grid(L.GRID_RHS))
            c.iterative_solver(eps=1e-10, ncycl=20000, start=L.START_COLD if it == 1 else L.START_PREV)
            c.calc_efield()
        if it == 1:
            stock0 = {p: _stocks(c, nsp) for p, c in ctx.items()}
            g = {p: [c.get_grid(L.GRID_RHS)] + [c.get_grid(L.GRID_NP, q) for q in range(1, nsp + 1)]
                 for p, c in ctx.items()}
            qscale = L.QE * float(np.max(np.abs(g[L.FP64][1])))
            out["first_deposit"] = {f"np_{nm}": _rel(g[L.MIXED][q], g[L.FP64][q]) for q, nm in enumerate(names, 1)}
            out["first_deposit"]["rhs_over_species_charge"] = \
                float(np.max(np.abs(g[L.MIXED][0] - g[L.FP64][0]))) / qscale
        for p, c in ctx.items():
            if it % 10 == 1:
                for s in range(1, nsp + 1):
                    c.part_sorting(s)
            for s in range(1, nsp + 1):
                c.part_mover(s, deposit=True)
            c.moments_accumulate(1)
    last, budget = {}, {}
    for p, c in ctx.items():
        last[p] = {f"np_{nm}": c.get_grid(L.GRID_NP, q) for q, nm in enumerate(names, 1)}
        last[p]["mom"] = c.moments()[0]
        c.calc_rhs()
        c.iterative_solver(eps=1e-10, ncycl=20000, start=L.START_PREV)
        c.calc_efield()
        last[p].update(phi=c.get_grid(L.GRID_PHI), E=c.get_grid(L.GRID_E))
        s0, s1 = stock0[p], _stocks(c, nsp)
        wall = sum(c.mover_tally(q)["P_loss"][0] for q in range(1, nsp + 1))
        e0 = s0["kinetic"] + s0["field"]
        budget["fp64" if p == L.FP64 else "mixed"] = {
            "kinetic_J": [s0["kinetic"], s1["kinetic"]], "field_J": [s0["field"], s1["field"]],
            "wall_loss_J": wall,
            "energy_drift_rel": (s1["kinetic"] + s1["field"] + wall - e0) / e0,
            "charge_C": [s0["charge"], s1["charge"]],
            "charge_change_over_species_charge": (s1["charge"] - s0["charge"]) / s0["charge_e"],
            "lost": [a - b for a, b in zip(s0["count"], s1["count"])]}
    out["energy_charge_budget"] = budget
    out["after_steps"] = {k: _rel(last[L.MIXED][k], last[L.FP64][k])
                          for k in [f"np_{nm}" for nm in names] + ["phi", "E"]}
    out["after_steps"]["phi_max_abs_V"] = float(np.max(np.abs(last[L.FP64]["phi"])))
    mom = {}
    for s, name in enumerate(names):
        d64, dmx = last[L.FP64]["mom"][s], last[L.MIXED]["mom"][s]
        m = d64[..., 0] > 0.1 * d64[..., 0].max()
        for j, q in ((0, "density"), (1, "temperature")):
            mj = m & (d64[..., j] != 0)
            mom[f"{q}_{name}"] = float(np.max(np.abs(dmx[..., j][mj] - d64[..., j][mj]) / np.abs(d64[..., j][mj])))
    out["moments_max_rel_on_dense_nodes"] = mom
    out["steps"] = steps
    for c in ctx.values():
        c.close()

    def solve_seq(ng, prec, k):
        c = L.Context(n1, n2, W["h"], W["h"], W["dt"], [L.Species(-L.QE, L.ME, 1.0, 16)], ng=ng,
                      ngrid=W["ngrid"], precision=prec)
        c.set_boundary(W["bcnd"], np.zeros((n2 + 3, n1 + 3)))
        ms, sw, vc = [], [], []
        for j, b in enumerate(rhs_seq[-k:]):
            c.set_grid(L.GRID_RHS, b)
            c.sync()
            t = time.perf_counter()
            info = c.iterative_solver(eps=1e-10, ncycl=20000, start=L.START_COLD if j == 0 else L.START_PREV)
            c.sync()
            ms.append(1e3 * (time.perf_counter() - t))
            sw.append(int(np.sum(info["sweeps"])))
            vc.append(int(info["vcycles"]))
        phi = c.get_grid(L.GRID_PHI)
        c.close()
        return ms, sw, vc, phi

    k = min(sor_solves, len(rhs_seq))
    if k == 0:
        return out, None
    ms1, sw1, _, phi1 = solve_seq(1, L.FP64, k)
    ms9, _, vc9, phi9 = solve_seq(W["ng"], L.FP64, k)
    msm, _, vcm, phim = solve_seq(W["ng"], L.MIXED, k)
    sor = {"solves": k, "sor_ng1_fp64_ms": ms1, "sor_ng1_sweeps": sw1,
           "mg_fp64_ms": ms9, "mg_fp64_vcycles": vc9, "mg_mixed_ms": msm, "mg_mixed_vcycles": vcm,
           "phi_sor_vs_mg_fp64": _rel(phi1, phi9), "phi_mg_mixed_vs_fp64": _rel(phim, phi9),
           "note": "first solve cold, later ones warm-started from the previous phi; "
                   "right-hand sides of the last steps of the FP64 run"}
    return out, sor


def cpu_baseline(W, budget_s=10.0):
    """The oracle (scalar C restatement, one core) on a bounded sample of the
    same workload: 64 x 64 cells of it at 100 ppc for both species."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n, ppc, h = 64, W["ppc"], W["h"]
    rng = np.random.default_rng(3)
    N = n * n * ppc
    pos = np.column_stack([rng.uniform(0, n * h, N), rng.uniform(0, n * h, N), np.zeros(N)])
    ve = rng.normal(0, W["vte"] / np.sqrt(2), (N, 3))
    vi = rng.normal(0, W["vti"] / np.sqrt(2), (N, 3))
    bc = O.box_bcnd(n, n)
    x = (np.arange(n + 3) - 1) * h
    Bi = np.zeros((n + 3, n + 3, 4))
    Bi[..., 2] = 0.01 * np.exp(-(x - n * h / 2) ** 2 / (2 * (n * h / 8) ** 2))[None, :]
    Bi[..., 3] = np.abs(Bi[..., 2])
    dom = O.Domain(n1=n, n2=n, h1=h, h2=h, dt=W["dt"], bcnd=bc, Bi=Bi, nB=(n, n), hB=(h, h))
    Ei = np.random.default_rng(4).normal(0, 1e3, (n + 3, n + 3, 2))
    chem = W["chem"]
    tab = O.ordering([dict(col_info=r.col_info(), Eth=r.Eth, E=r.E, S=r.S) for r in chem.reactions],
                     chem.masses(), chem.ntype_all)
    setup = dict(ntype=2, ntype_all=3, ns_coll=10, plt_src=0, n_cell=n * n, n_rg=1, ixl_rg=[1],
                 ixr_rg=[n + 1], iyl_rg=1, iyr_rg=n + 1, flag_ionrep=0,
                 charge={1: -L.QE, 2: L.QE, 3: 0.0}, mass=chem.masses(),
                 Nm={1: W["Nm"], 2: W["Nm"], 3: W["Nm"]},
                 vt0={1: W["vte"], 2: W["vti"], 3: W["vtg"]},
                 np_mx={1: 0.0, 2: 0.0, 3: W["ngas"]}, nu_uplim={1: 6e7, 2: 5e6, 3: 5e6})
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(harness):
        return cpu_baseline_reference(harness, W, dom, Ei, pos, ve, vi, N, O, setup, tab, budget_s)
    t_push = t_mcc = 0.0
    reps = 0
    start = time.perf_counter()
    while time.perf_counter() - start < budget_s or reps < 1:
        e = np.column_stack([pos, ve])
        i = np.column_stack([pos, vi])
        t_push += O.time_push_deposit(dom, -L.QE, L.ME, W["Nm"], Ei, e)
        t_push += O.time_push_deposit(dom, L.QE, 39.95 * L.AMU, W["Nm"], Ei, i)
        cap = N + N // 10
        pe = np.zeros((cap, 6)); pe[:N] = e
        pi_ = np.zeros((cap, 6)); pi_[:N] = i
        parts = {1: dict(vxp=pe, np_tot=N, flag_dead=np.zeros(cap, np.int8)),
                 2: dict(vxp=pi_, np_tot=N, flag_dead=np.zeros(cap, np.int8))}
        np_red = [np.full((n + 3, n + 3), 1e16), np.full((n + 3, n + 3), 1e16)]
        t = time.perf_counter()
        O.collisions(dom, setup, tab, parts, np_red, 123456)
        t_mcc += time.perf_counter() - t
        reps += 1
    per_step = (t_push + t_mcc / 10.0) / reps     # MCC runs every 10 steps
    return dict(value=2 * N / per_step, unit="particles/s", cores=1, kind="port",
                sample=f"oracle C restatement, 1 core: {n}x{n} cells x 100 ppc x 2 species "
                       f"({2 * N} particles), push+deposit each step + MCC/10, {reps} reps")


def _ref_push_deposit(harness, W, dom, Ei, vxp, charge, mass, reps, O):
    """Seconds of reps x (part_mover + charge_deposition) of the reference's
    own Fortran (part_expmover.f90, compiled by oracle/Makefile), one core."""
    import subprocess
    import tempfile
    n, h = dom.n1, dom.h1
    N = vxp.shape[0]
    ints = [n, n, dom.nB[0], dom.nB[1], N, dom.ngrid, 1, dom.flag_gridB, dom.flag_B_pos, 0, reps]
    dbls = [float(np.float32(1.60217646e-19)), 1.66053886e-27, h, h, dom.hB[0], dom.hB[1], dom.dt,
            charge, mass, W["Nm"], dom.zmax, n * h, n * h]
    kq = O.kq(dom)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int64(len(ints)).tobytes() + np.asarray(ints, np.int64).tobytes())
            f.write(np.int64(len(dbls)).tobytes() + np.asarray(dbls, np.float64).tobytes())
            for a in (np.asarray(Ei, np.float64), np.asarray(dom.Bi, np.float64),
                      np.asarray(dom.bcnd, np.int32), np.ascontiguousarray(vxp, np.float64),
                      np.zeros(N, np.int8), np.zeros((n + 3, n + 3)), np.asarray(kq, np.float64)):
                f.write(np.ascontiguousarray(a).tobytes())
        subprocess.run([harness, "bench", fin, fout], check=True)
        sec = float(np.frombuffer(open(fout, "rb").read(8), np.float64)[0])
    return sec


def cpu_baseline_reference(harness, W, dom, Ei, pos, ve, vi, N, O, setup, tab, budget_s):
    """The reference's own push + deposition (Fortran, one core) on the
    sample; MCC every 10 steps from the C restatement (collisions.f90 needs
    MPI, so it cannot be compiled here)."""
    e = np.column_stack([pos, ve])
    i = np.column_stack([pos, vi])
    probe = _ref_push_deposit(harness, W, dom, Ei, e, -L.QE, L.ME, 4, O) + \
        _ref_push_deposit(harness, W, dom, Ei, i, L.QE, 39.95 * L.AMU, 4, O)
    reps = max(4, int(budget_s / max(probe / 4, 1e-6)))
    t_push = _ref_push_deposit(harness, W, dom, Ei, e, -L.QE, L.ME, reps, O) + \
        _ref_push_deposit(harness, W, dom, Ei, i, L.QE, 39.95 * L.AMU, reps, O)
    n = dom.n1
    cap = N + N // 10
    pe = np.zeros((cap, 6)); pe[:N] = e
    pi_ = np.zeros((cap, 6)); pi_[:N] = i
    parts = {1: dict(vxp=pe, np_tot=N, flag_dead=np.zeros(cap, np.int8)),
             2: dict(vxp=pi_, np_tot=N, flag_dead=np.zeros(cap, np.int8))}
    np_red = [np.full((n + 3, n + 3), 1e16), np.full((n + 3, n + 3), 1e16)]
    t = time.perf_counter()
    O.collisions(dom, setup, tab, parts, np_red, 123456)
    t_mcc = time.perf_counter() - t
    per_step = t_push / reps + t_mcc / 10.0
    return dict(value=2 * N / per_step, unit="particles/s", cores=1, kind="reference",
                sample=f"reference part_mover + charge_deposition (part_expmover.f90 built by amdflang -O2, "
                       f"1 core) on {n}x{n} cells x 100 ppc x 2 species ({2 * N} particles), {reps} steps; "
                       f"plus MCC/10 from the C restatement")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=("e-Ar", "ion-source"), default="e-Ar",
                    help="e-/Ar+ box (--n, --ppc) or the 1024x512 four-species negative-ion-source domain")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=100, help="macroparticles per cell of each species")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-plain-mixed", action="store_true")
    ap.add_argument("--skip-learned", action="store_true")
    ap.add_argument("--learn-steps", type=int, default=40)
    ap.add_argument("--train-iters", type=int, default=300)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--skip-accuracy", action="store_true",
                    help="skip the untimed mixed-vs-FP64 comparison and the ng=1 SOR Poisson run")
    ap.add_argument("--sor-solves", type=int, default=3,
                    help="Poisson solves of the ng=1 SOR run beside MG (0: none)")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="after the runs, one device-to-device copy of GIB GiB (a plain copy kernel "
                         "whose HBM counters calibrate those of the mover under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the multi-GPU "
              "split is a design only)", file=sys.stderr)
        sys.exit(2)
    hip = Hip()
    W = workload(args.n, args.ppc) if args.workload == "e-Ar" else workload_ion_source()
    nsp = len(W["species"])
    N = sum(k * W["n2"] * int(round((xr - xl) / W["h"])) for sp in W["species"] for k, xl, xr in sp["loads"])
    K = args.steps

    def summary(a, wall, cnt):
        pdm = (a["push"] + a["mcc"]) / K            # push + deposit + MCC per step (ms)
        step = (a["field"] + a["sort"] + a["mcc"] + a["push"]) / K
        out = dict(ms_per_step=step, ms_push_deposit=a["push"] / K, ms_mcc=a["mcc"] / K,
                   ms_sort=a["sort"] / K, ms_field=a["field"] / K, vcycles_per_step=a["vcycles"] / K,
                   particles_per_s=sum(cnt) / (pdm * 1e-3), wall_s=wall, particles=sum(cnt),
                   electron_mover_ms=float(np.mean(a["push_kernel_ms"][0])),
                   ion_mover_ms=float(np.mean([np.mean(x) for x in a["push_kernel_ms"][1:]])))
        if nsp > 2:
            out["species_mover_ms"] = {sp["name"]: float(np.mean(x)) for sp, x in zip(W["species"], a["push_kernel_ms"])}
        if a["loss"] is not None:
            out.update(train_s=a["train_s"], train_loss_first=a["loss"][0], train_loss_last=a["loss"][1],
                       net_probed=a["net_probed"] / K, net_start_used=a["net_used"] / K,
                       res0_gain_median=float(np.median(a["res0_gain"])) if a["res0_gain"] else None)
        return out

    res = {}
    if not args.skip_learned:
        print(f"[bench] mixed precision + learned warm start, {N} particles, "
              f"{args.learn_steps} recorded + {args.warmup}+{K} steps", file=sys.stderr, flush=True)
        r = run_mode(W, L.MIXED, K, args.warmup, hip, learned=True, learn_steps=args.learn_steps,
                     train_iters=args.train_iters)
        res["mixed_learned"] = (r, summary(*r))
    if not args.skip_plain_mixed:
        print(f"[bench] mixed precision, {N} particles, {args.learn_steps}+{args.warmup}+{K} steps", file=sys.stderr,
              flush=True)
        r = run_mode(W, L.MIXED, K, args.warmup, hip, learn_steps=args.learn_steps)
        res["mixed"] = (r, summary(*r))
    if not args.skip_fp64:
        print("[bench] FP64", file=sys.stderr, flush=True)
        r = run_mode(W, L.FP64, K, args.warmup, hip, learn_steps=args.learn_steps)
        res["fp64"] = (r, summary(*r))
    head = "mixed_learned" if "mixed_learned" in res else "mixed"
    (acc_h, _, cnt_h), s_h = res[head]
    # roofline of the dominant kernel, the mixed-precision mover with fused
    # deposition (one instantiation serves electrons and ions): algorithmic
    # bytes = 2 x 24 B (int32 cell + 2 FP32 offsets + 3 FP32 velocities, read
    # and written) per particle and launch; launch durations come from HIP
    # events around that kernel alone, on the library's stream
    bpp = 48.0
    kms = [float(np.mean(x)) for x in acc_h["push_kernel_ms"]]
    alg = [c * bpp for c in cnt_h]
    achieved = sum(alg) / (sum(kms) * 1e-3) / 1e9
    # HBM traffic per particle was measured with PMC counters on the default
    # workload only
    measured = W["kind"] == "e-Ar" and args.n == 512 and args.ppc == 100
    cpu = None if (args.skip_cpu or W["kind"] != "e-Ar") else cpu_baseline(W, args.cpu_budget)
    out = {
        "metric": "ms/step & particles/s (push+deposit+MCC), mixed+learned vs FP64 GPU",
        "value": s_h["particles_per_s"],
        "unit": "particles/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": s_h["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": W["desc"] + ", " + ("mixed precision + learned warm start" if head == "mixed_learned"
                                                   else "mixed precision"),
                   "global_batch": N, "seq_len": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": TRAFFIC_BYTES_PER_PARTICLE * sum(cnt_h) / nsp if measured else None,
                     "kernel": "k_mover_mixed<256> (one launch per species)",
                     "alg_bytes_per_launch": sum(alg) / nsp,
                     "avg_launch_ms": sum(kms) / nsp,
                     "avg_launch_ms_electrons": kms[0], "avg_launch_ms_ions": float(np.mean(kms[1:]))},
        "cpu_baseline": cpu,
    }
    for k, (_, sm) in res.items():
        out[k] = sm
    if "fp64" in res:
        (acc_f, _, cnt_f), s_f = res["fp64"]
        kf = np.array(acc_f["push_kernel_ms"][0])
        out["fp64"]["roofline_push_GBs"] = cnt_f[0] * 96.0 / (kf.mean() * 1e-3) / 1e9
        out["speedup_vs_fp64_step"] = s_f["ms_per_step"] / s_h["ms_per_step"]
        out["speedup_vs_fp64_pdm"] = s_h["particles_per_s"] / s_f["particles_per_s"]
    if not args.skip_accuracy:
        print("[bench] mixed against FP64 (untimed), ng=1 SOR beside MG", file=sys.stderr, flush=True)
        out["accuracy_mixed_vs_fp64"], out["poisson_sor_vs_mg"] = accuracy_and_sor(W, sor_solves=args.sor_solves)
    if args.copy_calibration > 0:
        out["copy_calibration_bytes"] = hip.copy_once(int(args.copy_calibration * (1 << 30)))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
