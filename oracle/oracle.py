"""ctypes front end of the CPU oracle (oracle/liblepic_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module; the product package
``jd_lepic_amd`` never does.

Array conventions (numpy, C order) follow the reference's Fortran arrays:
  node arrays  A(0:n1+2,0:n2+2)      -> shape (n2+3, n1+3), A[iy, ix]
  Ei(2,0:n1+2,0:n2+2)                -> shape (n2+3, n1+3, 2)
  Bi(4,0:nB1+2,0:nB2+2)              -> shape (nB2+3, nB1+3, 4)
  rhs(n1+1,n2+1)                     -> shape (n2+1, n1+1), rhs[iy-1, ix-1]
  vxp(6,nmax)                        -> shape (N, 6)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblepic_oracle.so")
_lib = None

# constants of the reference (main.f90:120-124)
QE = float(np.float32(1.60217646e-19))   # single-precision literal widened
EPS0 = 8.854187817e-12
AMU = 1.66053886e-27
ME = 9.10938188e-31                      # read_reactions.f90:105
PI = 4.0 * np.arctan(1.0)

MAX_LEVELS, MAX_REAC, MAX_TYPES, MAX_RG = 12, 64, 32, 5


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


class LoDomain(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("h1", C.c_double), ("h2", C.c_double),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("dt", C.c_double), ("qe", C.c_double), ("amu", C.c_double),
                ("eps0", C.c_double), ("bcnd", _ip), ("ngrid", C.c_int),
                ("flag_B", C.c_int), ("nB1", C.c_int), ("nB2", C.c_int),
                ("hB1", C.c_double), ("hB2", C.c_double), ("flag_gridB", C.c_int),
                ("flag_B_pos", C.c_int), ("Bi", _dp)]


class LoSpecies(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("Nm", C.c_double),
                ("tag_neg", C.c_int)]


class LoMoverTally(C.Structure):
    _fields_ = [("P_loss", C.c_double * 4), ("cnt_dead", C.c_double), ("np_lost", C.c_longlong)]


class LoEheat(C.Structure):
    _fields_ = [("flag_c", C.c_int), ("flag_gpp", C.c_int), ("eheat_type", C.c_int), ("n_icp", C.c_int),
                ("xa", C.c_double), ("dr", C.c_double), ("yl_pow", C.c_double), ("yr_pow", C.c_double),
                ("ymax", C.c_double), ("nudt", C.c_double), ("dt", C.c_double), ("Pabs_cor", C.c_double),
                ("Nm", C.c_double), ("mass", C.c_double), ("pi", C.c_double), ("vt", _dp)]


class LoEmitter(C.Structure):
    _fields_ = [("jne", C.c_double), ("Lg", C.c_double), ("Lh", C.c_double), ("xg1", C.c_double),
                ("THm", C.c_double), ("chmfrd_array", C.c_double * 4), ("n_holes", C.c_int), ("n_icp", C.c_int),
                ("ptype_fx", C.c_int), ("tag_neg", C.c_int), ("tag_b", C.c_int), ("flag_chmfrd", C.c_int),
                ("ns_Hm", C.c_int), ("Nm", C.c_double * (MAX_TYPES + 1)), ("mass", C.c_double * (MAX_TYPES + 1)),
                ("pi", C.c_double)]


class LoEmitterOut(C.Structure):
    _fields_ = [("cnt_b", C.c_longlong * 2), ("P_loss4", C.c_double * (MAX_TYPES + 1)), ("Lgrd", C.c_double),
                ("dNh", C.c_double), ("Nh", C.c_longlong)]


class LoInjection(C.Structure):
    _fields_ = [(k, C.c_int) for k in ("opt_inj", "flag_c", "flag_nmn", "flag_Tp", "flag_diffsrc", "ns_inj",
                                       "tag_neg", "ntype")] + \
               [(k, C.c_double) for k in ("I_inj", "xl_pow", "xr_pow", "yl_pow", "yr_pow", "xa", "dr", "THm",
                                          "th_B", "pi")] + \
               [(k, C.c_double * (MAX_TYPES + 1)) for k in ("ni0", "vt0", "Nm", "mass", "charge")]


class LoInjectionOut(C.Structure):
    _fields_ = [("injected", C.c_longlong * (MAX_TYPES + 1)), ("P_loss4", C.c_double * (MAX_TYPES + 1)),
                ("P_loss1", C.c_double * (MAX_TYPES + 1))]


class LoPoisson(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("ng", C.c_int), ("h1", C.c_double),
                ("h2", C.c_double), ("eps0", C.c_double), ("bcnd", _ip)]


class LoSweepLog(C.Structure):
    _fields_ = [("ksor", _ip), ("cap", C.c_int), ("count", C.c_int)]


class LoReactionInput(C.Structure):
    _fields_ = [("ncol", C.c_int), ("col_info", (C.c_int * 10) * MAX_REAC),
                ("Eth", C.c_double * MAX_REAC), ("npt", C.c_int * MAX_REAC),
                ("E", _dp * MAX_REAC), ("S", _dp * MAX_REAC)]


class LoReactionTable(C.Structure):
    _fields_ = [("ncol", C.c_int), ("npt_mx", C.c_int), ("sig_Er", _dp), ("sig", _dp),
                ("col_info", (C.c_int * 10) * MAX_REAC), ("Eth", C.c_double * MAX_REAC),
                ("p_ncol", C.c_int * (MAX_TYPES + 1)),
                ("sig_list", (C.c_int * MAX_REAC) * (MAX_TYPES + 1)),
                ("sigv_mx", (C.c_double * MAX_REAC) * (MAX_TYPES + 1))]


class LoMccSetup(C.Structure):
    _fields_ = [("ntype", C.c_int), ("ntype_all", C.c_int),
                ("charge", C.c_double * (MAX_TYPES + 1)), ("mass", C.c_double * (MAX_TYPES + 1)),
                ("Nm", C.c_double * (MAX_TYPES + 1)), ("vt0", C.c_double * (MAX_TYPES + 1)),
                ("np_mx", C.c_double * (MAX_TYPES + 1)), ("nu_uplim", C.c_double * (MAX_TYPES + 1)),
                ("ns_coll", C.c_int), ("plt_src", C.c_int), ("n_cell", C.c_int),
                ("n_rg", C.c_int), ("ixl_rg", C.c_int * MAX_RG), ("ixr_rg", C.c_int * MAX_RG),
                ("iyl_rg", C.c_int), ("iyr_rg", C.c_int), ("flag_ionrep", C.c_int)]


class LoParticles(C.Structure):
    _fields_ = [("vxp", _dp * (MAX_TYPES + 1)), ("flag_dead", _bp * (MAX_TYPES + 1)),
                ("np_tot", C.c_longlong * (MAX_TYPES + 1)), ("cap", C.c_longlong * (MAX_TYPES + 1)),
                ("src", _dp * (MAX_TYPES + 1)), ("sink", _dp * (MAX_TYPES + 1))]


class LoMccOut(C.Structure):
    _fields_ = [("cnt_col", _dp), ("P_loss3", C.c_double * (MAX_TYPES + 1)),
                ("Nc", C.c_longlong * (MAX_TYPES + 1)), ("nu_max", C.c_double * (MAX_TYPES + 1)),
                ("events", C.c_longlong * MAX_REAC), ("event_log", _ip),
                ("log_cap", C.c_longlong), ("log_len", C.c_longlong)]


def _declare(L):
    L.lo_ran2.restype = C.c_double
    L.lo_ran2.argtypes = [_ip]
    L.lo_load_gauss.argtypes = [_dp, _dp, C.c_double, _dp, C.c_double]
    L.lo_eheating.restype = C.c_double
    L.lo_eheating.argtypes = [C.POINTER(LoEheat), _dp, C.c_longlong, _ip]
    L.lo_part_injection.restype = C.c_int
    L.lo_part_injection.argtypes = [C.POINTER(LoDomain), C.POINTER(LoInjection), C.POINTER(C.c_longlong), _dp,
                                    C.POINTER(LoParticles), _ip, C.c_longlong, C.POINTER(LoInjectionOut)]
    L.lo_flux_emitter.restype = C.c_int
    L.lo_flux_emitter.argtypes = [C.POINTER(LoDomain), C.POINTER(LoEmitter), C.POINTER(LoParticles), _ip,
                                  C.c_longlong, C.POINTER(LoEmitterOut)]
    L.lo_part_moments.argtypes = [C.POINTER(LoDomain), C.POINTER(LoSpecies), _dp, C.c_longlong, _dp, C.c_int, _dp]
    L.lo_calc_avg.argtypes = [C.POINTER(LoDomain), C.c_int, C.POINTER(LoSpecies), C.POINTER(_dp), C.POINTER(_dp),
                              C.POINTER(_dp), C.POINTER(_dp), C.c_int, _dp]
    L.lo_part_mover.restype = C.c_int
    L.lo_part_mover.argtypes = [C.POINTER(LoDomain), C.POINTER(LoSpecies), _dp, _dp, _bp,
                                C.POINTER(C.c_longlong), C.POINTER(LoMoverTally), _dp, _dp]
    L.lo_charge_deposition.argtypes = [C.POINTER(LoDomain), C.POINTER(LoSpecies), _dp,
                                       C.c_longlong, _dp, _dp]
    L.lo_kq.argtypes = [C.POINTER(LoDomain), _dp]
    L.lo_calc_rhs.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.POINTER(_dp), _dp]
    L.lo_dens_red.argtypes = [C.c_int, C.c_int, C.c_int, _ip, C.POINTER(_dp), C.POINTER(_dp), _dp]
    L.lo_calc_efield.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _dp]
    L.lo_part_sorting.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_longlong, _ip]
    L.lo_sor_rb.argtypes = [_dp, _dp, _dp, C.c_double, _ip, C.c_int, C.c_int, C.c_double,
                            C.c_double, C.c_int, C.c_int, _dp, _ip]
    L.lo_restriction.argtypes = [_dp, _dp, _dp, C.c_double, _ip, _ip, _dp, _dp, C.c_int, C.c_int]
    L.lo_prolongation.argtypes = [_dp, _dp, _ip, C.c_int, C.c_int]
    L.lo_mg.restype = C.c_int
    L.lo_mg.argtypes = [C.POINTER(LoPoisson), _dp, _dp, C.c_double, C.c_double, C.c_int, _dp,
                        _dp, C.POINTER(LoSweepLog)]
    L.lo_iterative_solver.restype = C.c_int
    L.lo_iterative_solver.argtypes = [C.POINTER(LoPoisson), _dp, _dp, C.c_int, C.c_double, _ip,
                                      _dp, _dp, C.POINTER(LoSweepLog)]
    L.lo_ordering.restype = C.c_int
    L.lo_ordering.argtypes = [C.POINTER(LoReactionInput), _dp, C.c_double, C.c_int, _dp, _dp,
                              C.POINTER(LoReactionTable)]
    L.lo_time_push_deposit_omp.restype = C.c_double
    L.lo_time_push_deposit_omp.argtypes = [C.POINTER(LoDomain), C.POINTER(LoSpecies), _dp, _dp, _dp,
                                           C.c_longlong, C.c_int, C.c_int]
    L.lo_time_collisions_omp.restype = C.c_double
    L.lo_time_collisions_omp.argtypes = [C.POINTER(LoDomain), C.POINTER(LoMccSetup), C.POINTER(LoReactionTable),
                                         C.POINTER(LoParticles), C.POINTER(_dp), C.c_int]
    L.lo_collisions.restype = C.c_int
    L.lo_collisions.argtypes = [C.POINTER(LoDomain), C.POINTER(LoMccSetup),
                                C.POINTER(LoReactionTable), C.POINTER(LoParticles),
                                C.POINTER(_dp), _ip, C.POINTER(LoMccOut)]


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
@dataclass
class Domain:
    """Grid, walls and static B field (the scalars of part_expmover.f90)."""
    n1: int
    n2: int
    h1: float
    h2: float
    dt: float
    bcnd: np.ndarray
    zmax: float = 1.0
    eps0: float = EPS0
    ngrid: int = 1
    Bi: np.ndarray | None = None      # (nB2+3, nB1+3, 4)
    nB: tuple = (1, 1)
    hB: tuple = (0.0, 0.0)
    flag_gridB: int = 0
    flag_B_pos: int = 0
    _keep: list = field(default_factory=list, repr=False)

    def struct(self) -> LoDomain:
        bc = np.ascontiguousarray(self.bcnd, dtype=np.int32)
        Bi = self.Bi if self.Bi is not None else np.zeros((4, 4, 4))
        Bi = np.ascontiguousarray(Bi, dtype=np.float64)
        self._keep = [bc, Bi]
        return LoDomain(self.n1, self.n2, self.h1, self.h2, self.n1 * self.h1, self.n2 * self.h2,
                        self.zmax, self.dt, QE, AMU, self.eps0, _p(bc, _ip), self.ngrid,
                        1 if self.Bi is not None else 0, self.nB[0], self.nB[1], self.hB[0],
                        self.hB[1], self.flag_gridB, self.flag_B_pos, _p(Bi))


def box_bcnd(n1: int, n2: int, label: int = 1) -> np.ndarray:
    """Grounded box of generate_boundary.f90:193-230 drawn by the four walls
    of a rectangle: rows/columns 0, 1, n+1, n+2 carry the wall label."""
    b = -np.ones((n2 + 3, n1 + 3), dtype=np.int32)
    b[:2, :] = label
    b[n2 + 1:, :] = label
    b[:, :2] = label
    b[:, n1 + 1:] = label
    return b


def ran2(seed: int, count: int):
    s = C.c_int32(seed)
    out = np.array([lib().lo_ran2(C.byref(s)) for _ in range(count)])
    return out, s.value


def load_gauss(vt: float, rnd: np.ndarray):
    rnd = np.ascontiguousarray(rnd, dtype=np.float64).reshape(-1, 2)
    out = np.empty_like(rnd)
    vx, vy = C.c_double(), C.c_double()
    for k in range(rnd.shape[0]):
        lib().lo_load_gauss(C.byref(vx), C.byref(vy), vt, _p(rnd[k]), PI)
        out[k] = (vx.value, vy.value)
    return out


def kq(dom: Domain) -> np.ndarray:
    out = np.zeros((dom.n2 + 3, dom.n1 + 3))
    d = dom.struct()
    lib().lo_kq(C.byref(d), _p(out))
    return out


def part_mover(dom: Domain, charge, mass, Nm, Ei, vxp, flag_dead=None, tag_neg=0):
    vxp = np.array(vxp, dtype=np.float64, order="C").reshape(-1, 6)
    N = vxp.shape[0]
    fd = np.zeros(N, np.int8) if flag_dead is None else np.array(flag_dead, np.int8)
    Ei = np.ascontiguousarray(Ei, dtype=np.float64)
    d = dom.struct()
    sp = LoSpecies(charge, mass, Nm, tag_neg)
    t = LoMoverTally()
    npt = C.c_longlong(N)
    p_mac = np.zeros(2 * (dom.ngrid + 1))
    ss2D = np.zeros((dom.n2 + 3, dom.n1 + 3))
    rc = lib().lo_part_mover(C.byref(d), C.byref(sp), _p(Ei), _p(vxp), _p(fd, _bp), C.byref(npt),
                             C.byref(t), _p(p_mac), _p(ss2D))
    if rc:
        raise RuntimeError(f"lo_part_mover rc={rc}")
    n = npt.value
    return dict(vxp=vxp[:n].copy(), np_tot=n, P_loss=np.array(t.P_loss[:]), cnt_dead=t.cnt_dead,
                p_mac=p_mac.reshape(2, dom.ngrid + 1), ss2D=ss2D, flag_dead=fd)


def charge_deposition(dom: Domain, Nm, vxp, kq_arr=None):
    vxp = np.ascontiguousarray(vxp, dtype=np.float64).reshape(-1, 6)
    k = kq(dom) if kq_arr is None else np.ascontiguousarray(kq_arr, dtype=np.float64)
    out = np.zeros((dom.n2 + 3, dom.n1 + 3))
    d = dom.struct()
    sp = LoSpecies(0.0, 1.0, Nm, 0)
    lib().lo_charge_deposition(C.byref(d), C.byref(sp), _p(vxp), vxp.shape[0], _p(k), _p(out))
    return out


def eheating(vxp, heat, iseed):
    """e_heating.f90:1-112 (lo_eheating).  heat: dict of flag_c, flag_gpp,
    eheat_type, xa, dr, yl_pow, yr_pow, ymax, nudt, dt, Pabs_cor, Nm, mass,
    vt (array of n_icp).  Returns (vxp, dP_loss2, iseed)."""
    v = np.array(vxp, dtype=np.float64).reshape(-1, 6)
    vt = np.ascontiguousarray(heat["vt"], dtype=np.float64)
    e = LoEheat(heat["flag_c"], heat["flag_gpp"], heat["eheat_type"], vt.size, heat["xa"], heat["dr"],
                heat["yl_pow"], heat["yr_pow"], heat["ymax"], heat["nudt"], heat["dt"], heat["Pabs_cor"],
                heat["Nm"], heat["mass"], PI, _p(vt))
    s = C.c_int32(iseed)
    dP = lib().lo_eheating(C.byref(e), _p(v), v.shape[0], C.byref(s))
    return v, dP, s.value


EMITTER_KEYS = ("jne", "Lg", "Lh", "xg1", "THm", "n_holes", "n_icp", "ptype_fx", "tag_neg", "tag_b",
                "flag_chmfrd", "ns_Hm")


def flux_emitter(dom: Domain, em: dict, parts: dict, iseed: int, max_tries: int = 1 << 20):
    """part_flux_emitter.f90:1-217 (lo_flux_emitter).  em: the keys of
    EMITTER_KEYS, chmfrd_array (4), and Nm / mass as {ptype: value}.  parts:
    {ptype: dict(vxp=(cap,6), np_tot, src=(n2+3,n1+3) optional)}, appended
    to in place.  Returns dict(rc, seed, cnt_b, P_loss4, Lgrd, dNh, Nh)."""
    e = LoEmitter()
    for k in EMITTER_KEYS:
        setattr(e, k, em[k])
    for j, v in enumerate(em.get("chmfrd_array", (0.0, 0.0, 0.0, 0.0))):
        e.chmfrd_array[j] = v
    for p, v in em["Nm"].items():
        e.Nm[p] = v
    for p, v in em["mass"].items():
        e.mass[p] = v
    e.pi = PI
    P = LoParticles()
    for p, d in parts.items():
        P.vxp[p] = _p(d["vxp"])
        P.np_tot[p] = d["np_tot"]
        P.cap[p] = d["vxp"].shape[0]
        if "src" in d:
            P.src[p] = _p(d["src"])
    o = LoEmitterOut()
    s = C.c_int32(iseed)
    d = dom.struct()
    rc = lib().lo_flux_emitter(C.byref(d), C.byref(e), C.byref(P), C.byref(s), max_tries, C.byref(o))
    for p, dd in parts.items():
        dd["np_tot"] = P.np_tot[p]
    return dict(rc=rc, seed=s.value, cnt_b=np.array(o.cnt_b[:]), P_loss4=np.array(o.P_loss4[:]), Lgrd=o.Lgrd,
                dNh=o.dNh, Nh=o.Nh)


INJECTION_INTS = ("opt_inj", "flag_c", "flag_nmn", "flag_Tp", "flag_diffsrc", "ns_inj", "tag_neg")
INJECTION_DBLS = ("I_inj", "xl_pow", "xr_pow", "yl_pow", "yr_pow", "xa", "dr", "THm", "th_B")


def part_injection(dom: Domain, inj: dict, parts: dict, iseed: int, N_inj=None, phi=None,
                   max_tries: int = 1 << 20):
    """main.f90:1063-1071 with part_injection.f90:1-237 (lo_part_injection).
    inj: INJECTION_INTS / INJECTION_DBLS keys and ni0, vt0, Nm, mass, charge
    as {ptype: value}.  parts as for flux_emitter (species 1..ntype), appended
    to in place; N_inj {ptype: lost count} for opt_inj = +-2.  Returns
    dict(rc, seed, injected, P_loss4, P_loss1) indexed by ptype."""
    e = LoInjection()
    for k in INJECTION_INTS:
        setattr(e, k, int(inj.get(k, 0)))
    for k in INJECTION_DBLS:
        setattr(e, k, float(inj.get(k, 0.0)))
    e.ntype = max(parts)
    e.pi = PI
    for k in ("ni0", "vt0", "Nm", "mass", "charge"):
        arr = getattr(e, k)
        for p, v in inj[k].items():
            arr[p] = v
    P = LoParticles()
    for p, d in parts.items():
        P.vxp[p] = _p(d["vxp"])
        P.np_tot[p] = d["np_tot"]
        P.cap[p] = d["vxp"].shape[0]
        if "src" in d:
            P.src[p] = _p(d["src"])
    n = (C.c_longlong * (MAX_TYPES + 1))()
    for p, v in (N_inj or {}).items():
        n[p] = v
    ph = np.ascontiguousarray(phi, np.float64) if phi is not None else None
    o = LoInjectionOut()
    s = C.c_int32(iseed)
    d = dom.struct()
    rc = lib().lo_part_injection(C.byref(d), C.byref(e), n, _p(ph) if ph is not None else None, C.byref(P),
                                 C.byref(s), max_tries, C.byref(o))
    for p, dd in parts.items():
        dd["np_tot"] = P.np_tot[p]
    return dict(rc=rc, seed=s.value, injected=np.array(o.injected[:]), P_loss4=np.array(o.P_loss4[:]),
                P_loss1=np.array(o.P_loss1[:]))


def part_moments(dom: Domain, Nm, vxp, size_na=1, kq_arr=None):
    """utils.f90:1-136 (lo_part_moments): p_mts as (na2+1, na1+1, 4) with
    [..., 0] = sum k v^2 and [..., 1:4] = sum k v."""
    vxp = np.ascontiguousarray(vxp, dtype=np.float64).reshape(-1, 6)
    k = kq(dom) if kq_arr is None else np.ascontiguousarray(kq_arr, dtype=np.float64)
    na1, na2 = dom.n1 // size_na, dom.n2 // size_na
    out = np.zeros((na2 + 1, na1 + 1, 4))
    d = dom.struct()
    sp = LoSpecies(0.0, 1.0, Nm, 0)
    lib().lo_part_moments(C.byref(d), C.byref(sp), _p(vxp), vxp.shape[0], _p(k), size_na, _p(out))
    return out


def calc_avg(dom: Domain, masses, Nms, nps, p_mts, sinks, size_na=1, data_pavg=None, srcs=None):
    """utils.f90:138-265 (lo_calc_avg): adds into data_pavg, shaped
    (ntype, na2+1, na1+1, 7) = [n, Tp, jx, jy, jz, source, sink]; srcs
    (the ss2D(1) maps) default to zero."""
    ntype = len(nps)
    na1, na2 = dom.n1 // size_na, dom.n2 // size_na
    out = np.zeros((ntype, na2 + 1, na1 + 1, 7)) if data_pavg is None else data_pavg
    sps = (LoSpecies * ntype)(*[LoSpecies(0.0, m, w, 0) for m, w in zip(masses, Nms)])
    arr = lambda xs: [np.ascontiguousarray(a, dtype=np.float64) for a in xs]
    a_np, a_m, a_s = arr(nps), arr(p_mts), arr(sinks)
    a_src = arr(srcs) if srcs is not None else None
    ptr = lambda xs: (_dp * ntype)(*[_p(a) for a in xs])
    d = dom.struct()
    lib().lo_calc_avg(C.byref(d), ntype, sps, ptr(a_np), ptr(a_m), ptr(a_src) if a_src is not None else None,
                      ptr(a_s), size_na, _p(out))
    return out


def calc_rhs(n1, n2, charges, nps):
    ch = np.ascontiguousarray(charges, dtype=np.float64)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in nps]
    ptrs = (_dp * len(arrs))(*[_p(a) for a in arrs])
    rhs = np.zeros((n2 + 1, n1 + 1))
    lib().lo_calc_rhs(n1, n2, len(arrs), _p(ch), ptrs, _p(rhs))
    return rhs


def dens_red(n1, n2, bcnd, nps):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in nps]
    outs = [np.zeros_like(a) for a in arrs]
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    mx = np.zeros(len(arrs))
    lib().lo_dens_red(n1, n2, len(arrs), _p(bc, _ip), (_dp * len(arrs))(*[_p(a) for a in arrs]),
                      (_dp * len(outs))(*[_p(a) for a in outs]), _p(mx))
    return outs, mx


def calc_efield(n1, n2, h, phi, bcnd, Ei0=None):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    Ei = np.zeros((n2 + 3, n1 + 3, 2)) if Ei0 is None else np.array(Ei0, dtype=np.float64)
    hh = np.array(h, dtype=np.float64)
    lib().lo_calc_efield(n1, n2, _p(hh), _p(phi), _p(bc, _ip), _p(Ei))
    return Ei


def part_sorting(n1, n2, h, vxp):
    v = np.array(vxp, dtype=np.float64, order="C").reshape(-1, 6)
    pl_max = n1 + n1 * n2 + 1
    Plist = np.zeros(pl_max + 1, np.int32)
    hh = np.array(h, dtype=np.float64)
    lib().lo_part_sorting(n1, n2, _p(hh), _p(v), v.shape[0], _p(Plist, _ip))
    return v, Plist


def sor_rb(u, b, h, eps0, bcnd, n1, n2, omega=1.4, eps=0.0, kmg=2, dig=-1):
    u = np.array(u, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    hh = np.array(h, dtype=np.float64)
    res, ks = C.c_double(), C.c_int32()
    lib().lo_sor_rb(_p(u), _p(b), _p(hh), eps0, _p(bc, _ip), n1, n2, omega, eps, kmg, dig,
                    C.byref(res), C.byref(ks))
    return u, res.value, ks.value


def restriction(u, b, h, eps0, bcnd, n1, n2):
    u = np.ascontiguousarray(u, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    m1, m2 = n1 // 2, n2 // 2
    bc2 = np.zeros((m2 + 3, m1 + 3), np.int32)
    r2 = np.zeros((m2 + 1, m1 + 1))
    e2 = np.ones((m2 + 3, m1 + 3))
    hh = np.array(h, dtype=np.float64)
    lib().lo_restriction(_p(u), _p(b), _p(hh), eps0, _p(bc, _ip), _p(bc2, _ip), _p(r2), _p(e2),
                         n1, n2)
    return r2, e2, bc2


def prolongation(u, e2h, bcnd, n1, n2):
    u = np.array(u, dtype=np.float64)
    e = np.ascontiguousarray(e2h, dtype=np.float64)
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    lib().lo_prolongation(_p(u), _p(e), _p(bc, _ip), n1, n2)
    return u


def iterative_solver(phi, rhs, bcnd, h, n1, n2, ng, eps=1e-10, eps0=EPS0, ncycl=20000,
                     log_cap=1 << 16):
    u = np.array(phi, dtype=np.float64)
    b = np.ascontiguousarray(rhs, dtype=np.float64)
    bc = np.ascontiguousarray(bcnd, dtype=np.int32)
    P = LoPoisson(n1, n2, ng, h[0], h[1], eps0, _p(bc, _ip))
    ks = np.zeros(log_cap, np.int32)
    log = LoSweepLog(_p(ks, _ip), log_cap, 0)
    k, ktot, res = C.c_int32(), C.c_double(), C.c_double()
    rc = lib().lo_iterative_solver(C.byref(P), _p(u), _p(b), ncycl, eps, C.byref(k), C.byref(ktot),
                                   C.byref(res), C.byref(log))
    if rc:
        raise RuntimeError(f"lo_iterative_solver rc={rc}")
    return dict(phi=u, vcycles=k.value, ktot=ktot.value, res=res.value,
                sweeps=ks[:min(log.count, log_cap)].copy())


def mg_alpha(n1, n2, h, eps0=EPS0):
    """alpha of iterative_solver.f90:22-23."""
    aw = eps0 / (h[0] * h[0])
    an = eps0 / (h[1] * h[1])
    ac = -(an + an + aw + aw)
    return abs((n1 + 1) * (n2 + 1) * ac)


# ---------------------------------------------------------------------------
def ordering(reactions, masses, ntype_all):
    """reactions: list of dicts with keys col_info (list of 10 ints), Eth, E, S."""
    ri = LoReactionInput()
    ri.ncol = len(reactions)
    keep = []
    for c, r in enumerate(reactions):
        for j, v in enumerate(r["col_info"]):
            ri.col_info[c][j] = int(v)
        ri.Eth[c] = r["Eth"]
        E = np.ascontiguousarray(r["E"], dtype=np.float64)
        S = np.ascontiguousarray(r["S"], dtype=np.float64)
        keep += [E, S]
        ri.npt[c] = len(E)
        ri.E[c] = _p(E)
        ri.S[c] = _p(S)
    total = sum(len(r["E"]) for r in reactions)
    sig_Er = np.zeros(max(total, 1))
    sig = np.zeros(max(total, 1) * max(len(reactions), 1))
    m = np.zeros(MAX_TYPES + 1)
    for k, v in masses.items():
        m[k] = v
    t = LoReactionTable()
    rc = lib().lo_ordering(C.byref(ri), _p(m), QE, ntype_all, _p(sig_Er), _p(sig), C.byref(t))
    if rc:
        raise RuntimeError(f"lo_ordering rc={rc}")
    n = t.npt_mx
    return dict(struct=t, sig_Er=sig_Er[:n].copy(), sig=sig[:n * len(reactions)].reshape(len(reactions), n).copy(),
                p_ncol=np.array(t.p_ncol[:]), sigv_mx=np.array([list(r) for r in t.sigv_mx]),
                sig_list=np.array([list(r) for r in t.sig_list]), _keep=(sig_Er, sig, keep))


def collisions(dom: Domain, setup: dict, table: dict, parts: dict, np_red, iseed: int,
               log_cap: int = 0):
    """parts: {ptype: {'vxp': (cap,6) array, 'np_tot': int, 'flag_dead': (cap,) int8}}"""
    s = LoMccSetup()
    for k in ("ntype", "ntype_all", "ns_coll", "plt_src", "n_cell", "n_rg", "iyl_rg", "iyr_rg",
              "flag_ionrep"):
        setattr(s, k, int(setup[k]))
    for k in ("charge", "mass", "Nm", "vt0", "np_mx", "nu_uplim"):
        arr = getattr(s, k)
        for p, v in setup[k].items():
            arr[p] = v
    for j in range(setup["n_rg"]):
        s.ixl_rg[j] = setup["ixl_rg"][j]
        s.ixr_rg[j] = setup["ixr_rg"][j]
    P = LoParticles()
    for p, d in parts.items():
        P.vxp[p] = _p(d["vxp"])
        P.flag_dead[p] = _p(d["flag_dead"], _bp)
        P.np_tot[p] = d["np_tot"]
        P.cap[p] = d["vxp"].shape[0]
        if "src" in d:
            P.src[p] = _p(d["src"])
        if "sink" in d:
            P.sink[p] = _p(d["sink"])
    ncol = table["struct"].ncol
    cnt = np.zeros(3 * ncol * max(setup["n_rg"], 1))
    log = np.zeros(2 * max(log_cap, 1), np.int32)
    o = LoMccOut()
    o.cnt_col = _p(cnt)
    o.event_log = _p(log, _ip) if log_cap else _ip()
    o.log_cap = log_cap
    reds = [np.ascontiguousarray(a, dtype=np.float64) for a in np_red]
    seed = C.c_int32(iseed)
    d = dom.struct()
    rc = lib().lo_collisions(C.byref(d), C.byref(s), C.byref(table["struct"]), C.byref(P),
                             (_dp * len(reds))(*[_p(a) for a in reds]), C.byref(seed), C.byref(o))
    if rc:
        raise RuntimeError(f"lo_collisions rc={rc}")
    for p, dd in parts.items():
        dd["np_tot"] = P.np_tot[p]
    return dict(seed=seed.value, cnt_col=cnt.reshape(3, ncol, max(setup["n_rg"], 1)),
                P_loss3=np.array(o.P_loss3[:]), Nc=np.array(o.Nc[:]), nu_max=np.array(o.nu_max[:]),
                events=np.array(o.events[:ncol]), log=log[:2 * o.log_len].reshape(-1, 2).copy())


def time_push_deposit(dom: Domain, charge, mass, Nm, Ei, vxp, reps=1):
    """Seconds spent in lo_part_mover + lo_charge_deposition alone (no
    wrapper copies), for the CPU baseline; vxp is left unchanged."""
    import time
    Ei = np.ascontiguousarray(Ei, dtype=np.float64)
    src = np.ascontiguousarray(vxp, dtype=np.float64).reshape(-1, 6)
    work = src.copy()
    N = src.shape[0]
    fd = np.zeros(N, np.int8)
    d = dom.struct()
    sp = LoSpecies(charge, mass, Nm, 0)
    k = kq(dom)
    rho = np.zeros((dom.n2 + 3, dom.n1 + 3))
    p_mac = np.zeros(2 * (dom.ngrid + 1))
    ss2D = np.zeros((dom.n2 + 3, dom.n1 + 3))
    total = 0.0
    for _ in range(reps):
        work[:] = src
        t = LoMoverTally()
        npt = C.c_longlong(N)
        t0 = time.perf_counter()
        lib().lo_part_mover(C.byref(d), C.byref(sp), _p(Ei), _p(work), _p(fd, _bp), C.byref(npt),
                            C.byref(t), _p(p_mac), _p(ss2D))
        lib().lo_charge_deposition(C.byref(d), C.byref(sp), _p(work), npt.value, _p(k), _p(rho))
        total += time.perf_counter() - t0
    return total / reps


def time_push_deposit_omp(dom: Domain, charge, mass, Nm, Ei, vxp, nthreads, reps=1):
    """Seconds of `reps` steps of push + deposit of one species with the
    particles split into nthreads partitions, one OpenMP thread each with its
    own density array (the reference's main.f90:1024-1089); vxp unchanged."""
    Ei = np.ascontiguousarray(Ei, dtype=np.float64)
    src = np.ascontiguousarray(vxp, dtype=np.float64).reshape(-1, 6)
    d = dom.struct()
    sp = LoSpecies(charge, mass, Nm, 0)
    k = kq(dom)
    return lib().lo_time_push_deposit_omp(C.byref(d), C.byref(sp), _p(Ei), _p(k), _p(src), src.shape[0],
                                          int(nthreads), int(reps))


def time_collisions_omp(dom: Domain, setup: dict, table: dict, parts: dict, np_red, nthreads):
    """Seconds of one collisions call with every species' particles split
    into nthreads partitions, each thread colliding its own partition with
    its own seed (collision_OMP per thread, collisions.f90:160-181)."""
    s = LoMccSetup()
    for k in ("ntype", "ntype_all", "ns_coll", "plt_src", "n_cell", "n_rg", "iyl_rg", "iyr_rg",
              "flag_ionrep"):
        setattr(s, k, int(setup[k]))
    for k in ("charge", "mass", "Nm", "vt0", "np_mx", "nu_uplim"):
        arr = getattr(s, k)
        for p, v in setup[k].items():
            arr[p] = v
    for j in range(setup["n_rg"]):
        s.ixl_rg[j] = setup["ixl_rg"][j]
        s.ixr_rg[j] = setup["ixr_rg"][j]
    P = LoParticles()
    keep = []
    for p, dd in parts.items():
        v = np.ascontiguousarray(dd["vxp"], dtype=np.float64)
        keep.append(v)
        P.vxp[p] = _p(v)
        P.np_tot[p] = dd["np_tot"]
        P.cap[p] = v.shape[0]
    reds = [np.ascontiguousarray(a, dtype=np.float64) for a in np_red]
    d = dom.struct()
    return lib().lo_time_collisions_omp(C.byref(d), C.byref(s), C.byref(table["struct"]), C.byref(P),
                                        (_dp * len(reds))(*[_p(a) for a in reds]), int(nthreads))
