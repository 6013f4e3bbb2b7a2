"""jd_lepic_amd -- MI355X (gfx950) implementation of LePIC's per-timestep
particle/field hot path behind a C ABI (include/lepic.h).

``Context`` mirrors the reference's hot-path routine list (part_mover,
charge_deposition, calc_rhs, iterative_solver, calc_Efield, part_sorting,
dens_red, collisions) on GPU-resident state.  Every call goes to the HIP
library liblepic_hip.so; when it is missing, creating a Context raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import (FP64, GRID_E, GRID_NP, GRID_NPRED, GRID_PHI, GRID_RHS, GRID_SINK, GRID_SRC, MCC_CHECK,
                   MCC_PROD, MIXED, START_COLD, START_GIVEN, START_NET, START_PREV)
from .reactions import AMU, ME, QE

EPS0 = 8.854187817e-12   # main.f90:122

__all__ = ["Context", "LepicError", "Species", "FP64", "MIXED", "MCC_CHECK", "MCC_PROD",
           "START_PREV", "START_COLD", "START_GIVEN", "START_NET", "GRID_PHI", "GRID_RHS", "GRID_E", "GRID_NP",
           "GRID_NPRED", "GRID_SINK", "GRID_SRC", "QE", "AMU", "ME", "EPS0", "backup_write", "backup_header",
           "backup_read_species"]


class LepicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"lepic error {code}: {msg}")
        self.code = code


@dataclass
class Species:
    charge: float
    mass: float
    Nm: float
    capacity: int


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _bp(a):
    return a.ctypes.data_as(C.POINTER(C.c_int8))


class Context:
    """GPU-resident state of one simulation (grid, fields, particles)."""

    def __init__(self, n1, n2, h1, h2, dt, species, zmax=1.0, eps0=EPS0, ng=1, ngrid=1,
                 precision=FP64, flag_B_pos=0, tag_neg=0, track_z=1, qe=QE, mg_tail_nodes=0):
        self.lib = capi.load()
        self.n1, self.n2, self.h = n1, n2, (h1, h2)
        self.nsp = len(species)
        self.ngrid = ngrid
        self.species = list(species)
        self.precision = precision
        g = capi.GridDesc(n1, n2, h1, h2, zmax, dt, eps0, qe, ng, ngrid, len(species), precision,
                          flag_B_pos, tag_neg, track_z, mg_tail_nodes)
        sp = (capi.SpeciesDesc * len(species))(*[capi.SpeciesDesc(s.charge, s.mass, s.Nm,
                                                                   int(s.capacity)) for s in species])
        self._ctx = C.c_void_p()
        rc = self.lib.lepic_create(C.byref(g), sp, C.byref(self._ctx))
        if rc:
            msg = self.lib.lepic_last_error(self._ctx).decode() if self._ctx.value else ""
            if self._ctx.value:
                self.lib.lepic_destroy(self._ctx)
            self._ctx = None
            raise LepicError(rc, msg)
        self.nodes = (n1 + 3) * (n2 + 3)

    # ---- plumbing ----
    def _check(self, rc):
        if rc:
            raise LepicError(rc, self.lib.lepic_last_error(self._ctx).decode())

    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.lepic_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def sync(self):
        self._check(self.lib.lepic_sync(self._ctx))

    def kernel_timing(self, enable=True):
        """Bracket each mover kernel with HIP events (measurement only)."""
        self._check(self.lib.lepic_kernel_timing(self._ctx, 1 if enable else 0))

    def kernel_ms(self, ptype):
        """Duration (ms) of the last timed mover kernel of ptype."""
        ms = C.c_float()
        self._check(self.lib.lepic_get_kernel_ms(self._ctx, ptype, C.byref(ms)))
        return float(ms.value)

    # ---- static inputs ----
    def set_boundary(self, bcnd, phi):
        b = np.ascontiguousarray(bcnd, dtype=np.int32)
        p = np.ascontiguousarray(phi, dtype=np.float64)
        assert b.size == self.nodes and p.size == self.nodes
        self._check(self.lib.lepic_set_boundary(self._ctx, _ip(b), _dp(p)))

    def set_bfield(self, Bi, nB, hB):
        B = np.ascontiguousarray(Bi, dtype=np.float64)
        assert B.size == 4 * (nB[0] + 3) * (nB[1] + 3)
        self._check(self.lib.lepic_set_bfield(self._ctx, nB[0], nB[1], hB[0], hB[1], _dp(B)))

    # ---- particles ----
    def upload(self, ptype, vxp):
        v = np.ascontiguousarray(vxp, dtype=np.float64).reshape(-1, 6)
        self._check(self.lib.lepic_upload_particles(self._ctx, ptype, _dp(v), v.shape[0]))

    def count(self, ptype):
        n = C.c_int64()
        self._check(self.lib.lepic_get_count(self._ctx, ptype, C.byref(n)))
        return n.value

    def download(self, ptype):
        n = self.count(ptype)
        out = np.zeros((max(n, 1), 6))
        got = C.c_int64()
        self._check(self.lib.lepic_download_particles(self._ctx, ptype, _dp(out), out.shape[0],
                                                      C.byref(got)))
        return out[:got.value]

    def download_range(self, ptype, first, count):
        """Particles first .. first+count-1 of the species, (count, 6)."""
        out = np.zeros((max(count, 1), 6))
        self._check(self.lib.lepic_download_particles_range(self._ctx, ptype, first, count, _dp(out)))
        return out[:count]

    # ---- averaged moments (part_moments / calc_avg) ----
    def moments_accumulate(self, size_na=1):
        self._check(self.lib.lepic_moments_accumulate(self._ctx, size_na))
        self._size_na = size_na

    def moments(self):
        """(data_pavg (ntype, na2+1, na1+1, 7), phi_avg (n2+3, n1+3, 3), count)."""
        na = getattr(self, "_size_na", 1)
        d = np.zeros((self.nsp, self.n2 // na + 1, self.n1 // na + 1, 7))
        ph = np.zeros((self.n2 + 3, self.n1 + 3, 3))
        c = C.c_int32()
        self._check(self.lib.lepic_moments_get(self._ctx, _dp(d), _dp(ph), C.byref(c)))
        return d, ph, c.value

    def moments_reset(self):
        self._check(self.lib.lepic_moments_reset(self._ctx))

    # ---- charge and energy totals (stocks for drift checks) ----
    def species_totals(self, ptype):
        """dict(count, kinetic (J), charge (C)) of one species."""
        out = np.zeros(3)
        self._check(self.lib.lepic_species_totals(self._ctx, ptype, _dp(out)))
        return dict(count=int(out[0]), kinetic=float(out[1]), charge=float(out[2]))

    def field_energy(self):
        """0.5 eps0 sum |E|^2 h1 h2 zmax over the interior nodes (J)."""
        w = C.c_double()
        self._check(self.lib.lepic_field_energy(self._ctx, C.byref(w)))
        return w.value

    # ---- backup files in the reference's particlesNNN.bak format ----
    def save_particles(self, path, time=0.0, gtype=0, pabs_cor=1.0):
        self._check(self.lib.lepic_save_particles(self._ctx, str(path).encode(), time, gtype, pabs_cor))

    def load_particles(self, path):
        t, g, p = C.c_double(), C.c_int32(), C.c_double()
        self._check(self.lib.lepic_load_particles(self._ctx, str(path).encode(), C.byref(t), C.byref(g),
                                                  C.byref(p)))
        return t.value, g.value, p.value

    def set_flag_dead(self, ptype, flags):
        f = np.ascontiguousarray(flags, dtype=np.int8)
        self._check(self.lib.lepic_set_flag_dead(self._ctx, ptype, _bp(f), f.size))

    def get_flag_dead(self, ptype, n=None):
        n = self.count(ptype) if n is None else n
        f = np.zeros(max(n, 1), np.int8)
        self._check(self.lib.lepic_get_flag_dead(self._ctx, ptype, _bp(f), n))
        return f[:n]

    def load_uniform(self, ptype, ppc, vt, xl, xr, seed):
        self._check(self.lib.lepic_load_uniform(self._ctx, ptype, ppc, vt, xl, xr, seed))

    # ---- grids ----
    def _shape(self, which):
        n1, n2 = self.n1, self.n2
        if which == GRID_RHS:
            return (n2 + 1, n1 + 1)
        if which == GRID_E:
            return (n2 + 3, n1 + 3, 2)
        return (n2 + 3, n1 + 3)

    def get_grid(self, which, ptype=0):
        out = np.zeros(self._shape(which))
        self._check(self.lib.lepic_get_grid(self._ctx, which, ptype, _dp(out)))
        return out

    def set_grid(self, which, arr, ptype=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self._shape(which)
        self._check(self.lib.lepic_set_grid(self._ctx, which, ptype, _dp(a)))

    # ---- phases (the reference's routines) ----
    def part_mover(self, ptype, deposit=True):
        self._check(self.lib.lepic_part_mover(self._ctx, ptype, 1 if deposit else 0))

    def charge_deposition(self, ptype):
        self._check(self.lib.lepic_charge_deposition(self._ctx, ptype))

    def mover_tally(self, ptype, ngrid=None):
        """P_loss(1:4), p_mac (2, ngrid+1) and cnt_dead of one species; p_mac
        has one column per wall label of the context (ngrid as created)."""
        if ngrid is not None and ngrid != self.ngrid:
            raise ValueError(f"the context has ngrid={self.ngrid}, not {ngrid}")
        ngrid = self.ngrid
        P = np.zeros(4)
        pm = np.zeros(2 * (ngrid + 1))
        cd = np.zeros(1)
        self._check(self.lib.lepic_get_mover_tally(self._ctx, ptype, _dp(P), _dp(pm), _dp(cd)))
        return dict(P_loss=P, p_mac=pm.reshape(2, ngrid + 1), cnt_dead=float(cd[0]))

    def reset_tallies(self):
        self._check(self.lib.lepic_reset_tallies(self._ctx))

    def calc_rhs(self):
        self._check(self.lib.lepic_calc_rhs(self._ctx))

    def calc_efield(self):
        self._check(self.lib.lepic_calc_efield(self._ctx))

    def iterative_solver(self, eps=1e-10, ncycl=20000, start=START_PREV, log_cap=4096):
        info = capi.SolveInfo()
        log = np.zeros(log_cap, np.int32)
        self._check(self.lib.lepic_iterative_solver(self._ctx, eps, ncycl, start, C.byref(info),
                                                    _ip(log), log_cap))
        return dict(vcycles=info.vcycles, status=info.status, res=info.res, ktot=info.ktot,
                    res0=info.res0, res0_alt=info.res0_alt, start_used=info.start_used,
                    sweeps=log[:min(info.nlog, log_cap)].copy())

    # ---- learned warm start (no counterpart in the reference) ----
    def net_create(self, capacity=64, seed=1):
        self._check(self.lib.lepic_net_create(self._ctx, capacity, seed))

    def net_recording(self, on=True):
        self._check(self.lib.lepic_net_recording(self._ctx, 1 if on else 0))

    def net_samples(self):
        n = C.c_int32()
        self._check(self.lib.lepic_net_samples(self._ctx, C.byref(n)))
        return n.value

    def net_copy_samples(self, src):
        """Append the recorded samples of context `src` (same grid) to this
        context's sample store; returns how many were copied."""
        n = C.c_int32()
        self._check(self.lib.lepic_net_copy_samples(self._ctx, src._ctx, C.byref(n)))
        return n.value

    def net_copy(self, src):
        """This context's network becomes a copy of `src`'s trained one."""
        self._check(self.lib.lepic_net_copy(self._ctx, src._ctx))

    def net_policy(self, min_gain=2.0, max_backoff=256):
        self._check(self.lib.lepic_net_policy(self._ctx, min_gain, max_backoff))

    def net_train(self, iters, batch=4, lr=1e-3):
        a, b = C.c_float(), C.c_float()
        self._check(self.lib.lepic_net_train(self._ctx, iters, batch, lr, C.byref(a), C.byref(b)))
        return float(a.value), float(b.value)

    def net_set_params(self, flat):
        f = np.ascontiguousarray(flat, np.float32)
        self._check(self.lib.lepic_net_set_params(self._ctx, f.ctypes.data_as(C.POINTER(C.c_float)), f.size))

    def net_get_params(self, n):
        f = np.zeros(n, np.float32)
        self._check(self.lib.lepic_net_get_params(self._ctx, f.ctypes.data_as(C.POINTER(C.c_float)), n))
        return f

    def net_forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        B, H, W, _ = x.shape
        y = np.zeros((B, H, W), np.float32)
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lepic_net_forward(self._ctx, x.ctypes.data_as(fp), B, H, W, y.ctypes.data_as(fp)))
        return y

    def net_loss_grad(self, x, target, mask, nparam, lam=0.0, cw=0.0, cn=0.0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        m = np.ascontiguousarray(mask, np.float32)
        B, H, W, _ = x.shape
        g = np.zeros(nparam, np.float32)
        loss = C.c_double()
        fp = C.POINTER(C.c_float)
        self._check(self.lib.lepic_net_loss_grad(self._ctx, x.ctypes.data_as(fp), t.ctypes.data_as(fp),
                                                 m.ctypes.data_as(fp), B, H, W, lam, cw, cn, C.byref(loss),
                                                 g.ctypes.data_as(fp), nparam))
        return float(loss.value), g

    def part_sorting(self, ptype, want_plist=False):
        if want_plist:
            pl = np.zeros(self.n1 + self.n1 * self.n2 + 2, np.int32)
            self._check(self.lib.lepic_part_sorting(self._ctx, ptype, _ip(pl)))
            return pl
        self._check(self.lib.lepic_part_sorting(self._ctx, ptype, None))
        return None

    def dens_red(self):
        mx = np.zeros(8)
        self._check(self.lib.lepic_dens_red(self._ctx, _dp(mx)))
        return mx[:self.nsp]

    # ---- collisions ----
    def set_reactions(self, chem, table, ns_coll=10, regions=None, vt0=None, nu_uplim=None,
                      flag_ionrep=0):
        """chem: reactions.Chemistry; table: tables.ordering(chem)."""
        nt = chem.ntype_all
        n1, n2 = self.n1, self.n2
        regions = regions or [(1, n1 + 1)]
        iy = (1, n2 + 1)
        arr = lambda vals: np.ascontiguousarray(vals, dtype=np.float64)
        charge = arr([0.0] + [s.charge for s in chem.species])
        mass = arr([0.0] + [s.mass for s in chem.species])
        Nm = arr([0.0] + [s.Nm if hasattr(s, "Nm") else 0.0 for s in chem.species])
        for k in range(self.nsp):
            Nm[k + 1] = self.species[k].Nm
        vt = arr([0.0] + list(vt0)) if vt0 is not None else arr(
            [0.0] + [np.sqrt(2.0 * QE * s.T / abs(s.mass)) for s in chem.species])
        np_mx = arr([0.0] + [s.ni0 if s.charge == 0.0 else 0.0 for s in chem.species])
        upl = arr([0.0] + (list(nu_uplim) if nu_uplim is not None else
                           [6e7] + [5e6] * (nt - 1)))        # main.f90:494-495
        self._rx_keep = keep = dict(
            sig_Er=arr(table["sig_Er"]), sig=arr(table["sig"]).ravel(), Eth=arr(table["Eth"]),
            sigv=arr(table["sigv_mx"]).ravel(),
            col=np.ascontiguousarray(table["col_info"], np.int32).ravel(),
            pn=np.ascontiguousarray(table["p_ncol"], np.int32),
            sl=np.ascontiguousarray(table["sig_list"], np.int32).ravel(),
            charge=charge, mass=mass, Nm=Nm, vt=vt, np_mx=np_mx, upl=upl)
        r = capi.Reactions()
        r.ntype_all, r.ncol, r.npt, r.ns_coll = nt, table["ncol"], table["npt"], ns_coll
        r.n_cell, r.n_rg, r.iyl_rg, r.iyr_rg = n1 * n2, len(regions), iy[0], iy[1]
        for k, (a, b) in enumerate(regions):
            r.ixl_rg[k], r.ixr_rg[k] = a, b
        r.flag_ionrep = flag_ionrep
        r.sig_Er, r.sig, r.Eth, r.sigv_mx = (_dp(keep["sig_Er"]), _dp(keep["sig"]), _dp(keep["Eth"]),
                                             _dp(keep["sigv"]))
        r.col_info, r.p_ncol, r.sig_list = _ip(keep["col"]), _ip(keep["pn"]), _ip(keep["sl"])
        r.charge, r.mass, r.Nm, r.vt0 = _dp(charge), _dp(mass), _dp(Nm), _dp(vt)
        r.np_mx, r.nu_uplim = _dp(np_mx), _dp(upl)
        self._ncol, self._nrg = table["ncol"], len(regions)
        self._check(self.lib.lepic_set_reactions(self._ctx, C.byref(r)))

    def collisions(self, it, mode=MCC_PROD, seed=0, iseed=123456):
        info = capi.MccInfo()
        s = C.c_int32(iseed)
        self._check(self.lib.lepic_collisions(self._ctx, it, mode, seed, C.byref(s), C.byref(info)))
        return dict(seed=s.value, Nc=np.array(info.Nc[:self.nsp]),
                    nu_max=np.array(info.nu_max[:self.nsp]), events=np.array(info.events[:self._ncol]),
                    created=np.array(info.created[:self.nsp]), missed=info.missed)

    def cnt_col(self):
        out = np.zeros(3 * self._ncol * self._nrg)
        self._check(self.lib.lepic_get_cnt_col(self._ctx, _dp(out)))
        return out.reshape(3, self._ncol, self._nrg)

    def mcc_ploss(self):
        out = np.zeros(9)
        self._check(self.lib.lepic_get_mcc_ploss(self._ctx, _dp(out)))
        return out

    # ---- electron heating (e_heating.f90:1-112) ----
    @staticmethod
    def _heating(heat):
        """heat: dict with flag_c, flag_gpp, eheat_type, xa, dr, yl_pow, yr_pow,
        nudt, Pabs_cor and vt (n_icp thermal speeds).  Returns the struct and
        the array it points at (kept alive by the caller)."""
        vt = np.ascontiguousarray(heat["vt"], dtype=np.float64).ravel()
        h = capi.Heating()
        h.flag_c, h.flag_gpp, h.eheat_type = heat["flag_c"], heat["flag_gpp"], heat["eheat_type"]
        h.n_icp = vt.size
        h.xa, h.dr, h.yl_pow, h.yr_pow = heat["xa"], heat["dr"], heat.get("yl_pow", 0.0), heat.get("yr_pow", 0.0)
        h.nudt, h.Pabs_cor, h.vt = heat["nudt"], heat.get("Pabs_cor", 1.0), _dp(vt)
        return h, vt

    def eheating(self, heat, it=0, mode=MCC_PROD, seed=0, iseed=123456):
        """One call of eheating on species 1; returns (dP, iseed): the energy
        change P_loss(2,1) gets and, in check mode, the ran2 state after it."""
        h, keep = self._heating(heat)
        s, dP = C.c_int32(iseed), C.c_double(0.0)
        self._check(self.lib.lepic_eheating(self._ctx, C.byref(h), mode, seed, it, C.byref(s), C.byref(dP)))
        del keep
        return dP.value, s.value

    # ---- surface emitter (part_flux_emitter.f90:1-217) ----
    @staticmethod
    def _emitter(em):
        """em: dict with jne, Lg, Lh, xg1, THm, n_holes, n_icp, ptype_fx,
        tag_b, flag_chmfrd, ns_Hm and optionally chmfrd_array (x1, x2, y1, y2)."""
        e = capi.Emitter()
        for k in ("jne", "Lg", "Lh", "xg1", "THm", "n_holes", "n_icp", "ptype_fx", "tag_b", "flag_chmfrd", "ns_Hm"):
            setattr(e, k, em[k])
        for j, v in enumerate(em.get("chmfrd_array", (0.0, 0.0, 0.0, 0.0))):
            e.chmfrd_array[j] = v
        return e

    def flux_emitter(self, em, it=0, mode=MCC_PROD, seed=0, iseed=123456):
        """One call of part_flux_emitter; returns dict(Nh, created, cnt_b,
        P_loss4, Lgrd, dNh, seed)."""
        e, info, s = self._emitter(em), capi.EmitterInfo(), C.c_int32(iseed)
        self._check(self.lib.lepic_flux_emitter(self._ctx, C.byref(e), it, mode, seed, C.byref(s), C.byref(info)))
        return dict(Nh=info.Nh, created=np.array(info.created[:self.nsp]), cnt_b=np.array(info.cnt_b[:]),
                    P_loss4=np.array(info.P_loss4[:self.nsp]), Lgrd=info.Lgrd, dNh=info.dNh, seed=s.value)

    # ---- volume injection (part_injection.f90:1-237) ----
    def _injection(self, inj):
        """inj: dict with opt_inj, flag_c, flag_nmn, flag_Tp, flag_diffsrc,
        ns_inj, I_inj, xl_pow, xr_pow, yl_pow, yr_pow, xa, dr, THm, th_B and
        ni0, vt0 (sequences over the species).  Returns the struct and the
        arrays it points at."""
        e = capi.Injection()
        for k in ("opt_inj", "flag_c", "flag_nmn", "flag_Tp", "flag_diffsrc", "ns_inj"):
            setattr(e, k, int(inj.get(k, 0 if k != "ns_inj" else 1)))
        for k in ("I_inj", "xl_pow", "xr_pow", "yl_pow", "yr_pow", "xa", "dr", "THm", "th_B"):
            setattr(e, k, float(inj.get(k, 0.0)))
        keep = [np.ascontiguousarray(inj["ni0"], dtype=np.float64), np.ascontiguousarray(inj["vt0"], dtype=np.float64)]
        assert keep[0].size >= self.nsp and keep[1].size >= self.nsp
        e.ni0, e.vt0 = _dp(keep[0]), _dp(keep[1])
        return e, keep

    def part_injection(self, inj, N_inj=None, it=0, mode=MCC_PROD, seed=0, iseed=123456, deposit=False):
        """One call of the injection species loop; N_inj: particles of each
        species lost in this step's push (opt_inj = +-2).  Returns
        dict(injected, P_loss4, P_loss1, seed)."""
        e, keep = self._injection(inj)
        n = np.zeros(self.nsp, np.int64) if N_inj is None else np.ascontiguousarray(N_inj, dtype=np.int64)
        info, s = capi.InjectionInfo(), C.c_int32(iseed)
        self._check(self.lib.lepic_part_injection(self._ctx, C.byref(e), n.ctypes.data_as(C.POINTER(C.c_int64)),
                                                  it, mode, seed, C.byref(s), 1 if deposit else 0, C.byref(info)))
        del keep
        return dict(injected=np.array(info.injected[:self.nsp]), P_loss4=np.array(info.P_loss4[:self.nsp]),
                    P_loss1=np.array(info.P_loss1[:self.nsp]), seed=s.value)

    # ---- whole step ----
    def step(self, it, eps=1e-10, ncycl=20000, nsort=10, ns_coll=10, mcc_mode=MCC_PROD, seed=0,
             iseed=123456, start=START_PREV, heat=None, ns_heat=1, heat_mode=MCC_PROD, emit=None, inj=None):
        a = capi.StepArgs()
        a.eps, a.ncycl, a.nsort, a.ns_coll, a.mcc_mode = eps, ncycl, nsort, ns_coll, mcc_mode
        a.seed, a.iseed, a.start = seed, iseed, start
        keep = None
        a.heat_mode = heat_mode
        if heat is not None:
            h, keep = self._heating(heat)
            a.heat, a.ns_heat = C.pointer(h), ns_heat
        if emit is not None:
            e = self._emitter(emit)
            a.emit = C.pointer(e)
        if inj is not None:
            ij, keep_inj = self._injection(inj)
            a.inj = C.pointer(ij)
        self._check(self.lib.lepic_step(self._ctx, it, C.byref(a)))
        del keep
        return dict(iseed=a.iseed, vcycles=a.solve.vcycles, res=a.solve.res, P_heat=a.P_heat, emitted=a.emitted,
                    injected=a.injected,
                    ms=dict(field=a.ms_rhs_poisson_e, sort=a.ms_sort, mcc=a.ms_mcc, push=a.ms_push))


# ---- backup files without a device (restart.f90:56-113, main.f90:1125-1131) ----
def backup_write(path, time, gtype, pabs_cor, parts):
    """parts[iproc][ptype-1]: (np, 6) arrays in vxp order."""
    lib = capi.load()
    nproc, ntype = len(parts), len(parts[0])
    arrs = [np.ascontiguousarray(parts[ip][p], dtype=np.float64).reshape(-1, 6)
            for ip in range(nproc) for p in range(ntype)]
    np_tot = np.array([a.shape[0] for a in arrs], np.int32)
    ptrs = (C.POINTER(C.c_double) * len(arrs))(*[_dp(a) for a in arrs])
    rc = lib.lepic_backup_write(str(path).encode(), time, gtype, pabs_cor, ntype, nproc, _ip(np_tot), ptrs)
    if rc:
        raise LepicError(rc, f"cannot write {path}")


def backup_header(path, ntype):
    lib = capi.load()
    t, g, pc, nproc = C.c_double(), C.c_int32(), C.c_double(), C.c_int32()
    cnt = np.zeros(4096, np.int32)
    rc = lib.lepic_backup_header(str(path).encode(), ntype, C.byref(t), C.byref(g), C.byref(pc), C.byref(nproc),
                                 _ip(cnt), cnt.size)
    if rc:
        raise LepicError(rc, f"cannot read {path}")
    return t.value, g.value, pc.value, cnt[:ntype * nproc.value].reshape(nproc.value, ntype)


def backup_read_species(path, ntype, ptype):
    lib = capi.load()
    counts = backup_header(path, ntype)[3]
    total = int(counts[:, ptype - 1].sum())
    out = np.zeros((max(total, 1), 6))
    n = C.c_int64()
    rc = lib.lepic_backup_read_species(str(path).encode(), ntype, ptype, _dp(out), out.shape[0], C.byref(n))
    if rc:
        raise LepicError(rc, f"cannot read {path}")
    return out[:n.value]
