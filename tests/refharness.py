"""Runs oracle/_ref/ref_harness: the reference's own Fortran routines compiled
from the reference checkout (test infrastructure; see oracle/Makefile)."""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
QE = float(np.float32(1.60217646e-19))
AMU = 1.66053886e-27


def available() -> bool:
    return os.path.exists(HARNESS)


def _run(cmd, ints, dbls, arrays):
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int64(len(ints)).tobytes())
            f.write(np.asarray(ints, np.int64).tobytes())
            f.write(np.int64(len(dbls)).tobytes())
            f.write(np.asarray(dbls, np.float64).tobytes())
            for a in arrays:
                f.write(np.ascontiguousarray(a).tobytes())
        subprocess.run([HARNESS, cmd, fin, fout], check=True)
        return open(fout, "rb").read()


class _Reader:
    def __init__(self, buf):
        self.buf, self.off = buf, 0

    def take(self, dtype, count):
        dt = np.dtype(dtype)
        a = np.frombuffer(self.buf, dtype=dt, count=count, offset=self.off)
        self.off += dt.itemsize * count
        return a.copy()


def ran2(seed, count):
    r = _Reader(_run("ran2", [seed, count], [], []))
    return r.take(np.float64, count), int(r.take(np.int32, 1)[0])


def load_gauss(vt, rnd):
    rnd = np.ascontiguousarray(rnd, np.float64).reshape(-1, 2)
    r = _Reader(_run("gauss", [rnd.shape[0]], [QE, AMU, vt], [rnd]))
    return r.take(np.float64, 2 * rnd.shape[0]).reshape(-1, 2)


def part_mover(dom, charge, mass, Nm, Ei, vxp, flag_dead=None, tag_neg=0, phi=None):
    n1, n2 = dom.n1, dom.n2
    vxp = np.ascontiguousarray(vxp, np.float64).reshape(-1, 6)
    N = vxp.shape[0]
    fd = np.zeros(N, np.int8) if flag_dead is None else np.asarray(flag_dead, np.int8)
    Bi = dom.Bi if dom.Bi is not None else np.zeros((dom.nB[1] + 3, dom.nB[0] + 3, 4))
    phi = np.zeros((n2 + 3, n1 + 3)) if phi is None else phi
    ints = [n1, n2, dom.nB[0], dom.nB[1], N, dom.ngrid, 1 if dom.Bi is not None else 0,
            dom.flag_gridB, dom.flag_B_pos, tag_neg]
    dbls = [QE, AMU, dom.h1, dom.h2, dom.hB[0], dom.hB[1], dom.dt, charge, mass, Nm, dom.zmax,
            n1 * dom.h1, n2 * dom.h2]
    r = _Reader(_run("mover", ints, dbls, [np.asarray(Ei, np.float64), np.asarray(Bi, np.float64),
                                           np.asarray(dom.bcnd, np.int32), vxp, fd,
                                           np.asarray(phi, np.float64)]))
    npt = int(r.take(np.int64, 1)[0])
    out = dict(np_tot=npt, vxp=r.take(np.float64, 6 * npt).reshape(-1, 6),
               P_loss=r.take(np.float64, 4))
    # p_mac(ntype=1, 2, 0:ngrid) -> [k, igrid]
    out["p_mac"] = r.take(np.float64, 2 * (dom.ngrid + 1)).reshape(dom.ngrid + 1, 2).T.copy()
    out["cnt_dead"] = float(r.take(np.float64, 1)[0])
    out["ss2D"] = r.take(np.float64, 2 * (n1 + 3) * (n2 + 3)).reshape(n2 + 3, n1 + 3, 2)[..., 1].copy()
    out["flag_dead"] = r.take(np.int8, max(N, 1))[:N]
    return out


def charge_deposition(dom, Nm, vxp, kq):
    vxp = np.ascontiguousarray(vxp, np.float64).reshape(-1, 6)
    r = _Reader(_run("deposit", [dom.n1, dom.n2, vxp.shape[0]],
                     [QE, AMU, dom.h1, dom.h2, Nm, dom.zmax], [vxp, np.asarray(kq, np.float64)]))
    return r.take(np.float64, (dom.n1 + 3) * (dom.n2 + 3)).reshape(dom.n2 + 3, dom.n1 + 3)


def calc_efield(n1, n2, h, phi, bcnd, Ei0=None):
    Ei0 = np.zeros((n2 + 3, n1 + 3, 2)) if Ei0 is None else Ei0
    r = _Reader(_run("efield", [n1, n2], [QE, AMU, h[0], h[1]],
                     [np.asarray(phi, np.float64), np.asarray(bcnd, np.int32),
                      np.asarray(Ei0, np.float64)]))
    return r.take(np.float64, 2 * (n1 + 3) * (n2 + 3)).reshape(n2 + 3, n1 + 3, 2)


def part_sorting(n1, n2, h, vxp):
    vxp = np.ascontiguousarray(vxp, np.float64).reshape(-1, 6)
    r = _Reader(_run("sort", [n1, n2, vxp.shape[0]], [QE, AMU, h[0], h[1]], [vxp]))
    npt = int(r.take(np.int64, 1)[0])
    v = r.take(np.float64, 6 * npt).reshape(-1, 6)
    return v, r.take(np.int32, n1 + n1 * n2 + 2)


def flux_emitter(dom, em, parts, iseed):
    """The reference's part_flux_emitter (part_flux_emitter.f90) on one
    partition; parts as for oracle.flux_emitter with species 1..ntype.
    Returns (parts copy, src maps, cnt_b, P_loss4, iseed, Lgrd)."""
    ntype = max(parts)
    nmax = max(d["vxp"].shape[0] for d in parts.values())
    n1, n2 = dom.n1, dom.n2
    vxp = np.zeros((ntype, nmax, 6))
    np_tot = np.zeros(ntype, np.int32)
    for p, d in parts.items():
        vxp[p - 1, :d["vxp"].shape[0]] = d["vxp"]
        np_tot[p - 1] = d["np_tot"]
    ints = [n1, n2, ntype, nmax, em["ptype_fx"], em["tag_neg"], em["tag_b"], em["flag_chmfrd"], em["n_holes"],
            em["n_icp"], em["ns_Hm"], iseed]
    dbls = [QE, AMU, dom.h1, dom.h2, em["jne"], em["Lg"], em["Lh"], em["xg1"], em["THm"], dom.n2 * dom.h2, dom.zmax,
            dom.dt, *em.get("chmfrd_array", (0.0, 0.0, 0.0, 0.0)),
            *[em["Nm"].get(p, 0.0) for p in range(1, ntype + 1)],
            *[em["mass"].get(p, 0.0) for p in range(1, ntype + 1)]]
    bc = np.ascontiguousarray(dom.bcnd, np.int32)
    r = _Reader(_run("emit", ints, dbls, [bc, np_tot, vxp]))
    v = r.take(np.float64, 6 * nmax * ntype).reshape(ntype, nmax, 6)
    counts = r.take(np.int32, ntype)
    ss = r.take(np.float64, 2 * (n1 + 3) * (n2 + 3) * ntype).reshape(ntype, n2 + 3, n1 + 3, 2)
    cnt_b = r.take(np.int32, 2)
    ploss = r.take(np.float64, 4 * ntype).reshape(ntype, 4)
    seed = int(r.take(np.int32, 1)[0])
    Lgrd = float(r.take(np.float64, 1)[0])
    out = {p: dict(vxp=v[p - 1], np_tot=int(counts[p - 1])) for p in range(1, ntype + 1)}
    return out, ss[..., 0], cnt_b, ploss[:, 3], seed, Lgrd


def eheating(vxp, heat, iseed):
    """The reference's eheating (e_heating.f90) on one partition."""
    vxp = np.ascontiguousarray(vxp, np.float64).reshape(-1, 6)
    vt = np.ascontiguousarray(heat["vt"], np.float64)
    ints = [vxp.shape[0], vt.size, heat["flag_c"], heat["flag_gpp"], heat["eheat_type"], iseed]
    dbls = [QE, AMU, heat["xa"], heat["dr"], heat["yl_pow"], heat["yr_pow"], heat["ymax"], heat["nudt"],
            heat["dt"], heat["Pabs_cor"], heat["Nm"], heat["mass"]]
    r = _Reader(_run("heat", ints, dbls, [vxp, vt]))
    v = r.take(np.float64, 6 * vxp.shape[0]).reshape(-1, 6)
    dP = float(r.take(np.float64, 1)[0])
    return v, dP, int(r.take(np.int32, 1)[0])
