"""Times the reference's own part_mover + charge_deposition (part_expmover.f90,
compiled from the reference checkout by oracle/Makefile into
oracle/_ref/ref_harness) beside the oracle's C restatement of the same
routines, on the CPU baseline's sample: a 128 x 128-cell box of the e-/Ar+
plasma at 100 particles per cell for both species, with the particles of a
species split into one partition per OpenMP thread as the reference's main
loop runs them (main.f90:1024-1089).  Runs only where the reference checkout
and its harness exist (the build container); prints one JSON line with
particles/s of both and their ratio.

    python oracle/time_reference.py [--threads T] [--reps R]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from jd_lepic_amd import workloads as WL  # noqa: E402


def ref_ompbench(harness, dom, Ei, vxp, charge, mass, Nm, reps, threads):
    n, h = dom.n1, dom.h1
    N = vxp.shape[0]
    ints = [n, n, dom.nB[0], dom.nB[1], N, dom.ngrid, 1, dom.flag_gridB, dom.flag_B_pos, 0, reps, threads]
    dbls = [float(np.float32(1.60217646e-19)), 1.66053886e-27, h, h, dom.hB[0], dom.hB[1], dom.dt, charge, mass,
            Nm, dom.zmax, n * h, n * h]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(np.int64(len(ints)).tobytes() + np.asarray(ints, np.int64).tobytes())
            f.write(np.int64(len(dbls)).tobytes() + np.asarray(dbls, np.float64).tobytes())
            for a in (np.asarray(Ei, np.float64), np.asarray(dom.Bi, np.float64), np.asarray(dom.bcnd, np.int32),
                      np.ascontiguousarray(vxp, np.float64), np.zeros(N, np.int8), np.zeros((n + 3, n + 3)),
                      np.asarray(O.kq(dom), np.float64)):
                f.write(np.ascontiguousarray(a).tobytes())
        subprocess.run([harness, "ompbench", fin, fout], check=True)
        return float(np.frombuffer(open(fout, "rb").read(8), np.float64)[0])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    harness = os.path.join(HERE, "_ref", "ref_harness")
    if not os.path.exists(harness):
        sys.exit("oracle/_ref/ref_harness is not built (needs the reference checkout)")
    W = WL.argon_box(128, 100)
    n, h = W["n1"], W["h"]
    dom = O.Domain(n1=n, n2=n, h1=h, h2=h, dt=W["dt"], bcnd=W["bcnd"], Bi=W["Bi"], nB=(n, n), hB=(h, h))
    rng = np.random.default_rng(3)
    N = n * n * W["ppc"]
    pos = np.column_stack([rng.uniform(0, n * h, N), rng.uniform(0, n * h, N), np.zeros(N)])
    e = np.column_stack([pos, rng.normal(0, W["vte"] / np.sqrt(2), (N, 3))])
    i = np.column_stack([pos, rng.normal(0, W["vti"] / np.sqrt(2), (N, 3))])
    Ei = np.random.default_rng(4).normal(0, 1e3, (n + 3, n + 3, 2))
    M = 39.95 * O.AMU
    out = {}
    for T in sorted({1, args.threads}):
        t_ref = ref_ompbench(harness, dom, Ei, e, -O.QE, O.ME, W["Nm"], args.reps, T) + \
            ref_ompbench(harness, dom, Ei, i, O.QE, M, W["Nm"], args.reps, T)
        t_res = O.time_push_deposit_omp(dom, -O.QE, O.ME, W["Nm"], Ei, e, T, args.reps) + \
            O.time_push_deposit_omp(dom, O.QE, M, W["Nm"], Ei, i, T, args.reps)
        out[f"threads_{T}"] = dict(reference_particles_per_s=2 * N * args.reps / t_ref,
                                   restatement_particles_per_s=2 * N * args.reps / t_res,
                                   reference_over_restatement=(2 * N * args.reps / t_ref) / (2 * N * args.reps / t_res))
    out["sample"] = f"{n}x{n} cells x {W['ppc']} ppc x 2 species ({2 * N} particles), {args.reps} steps"
    print(json.dumps(out))


if __name__ == "__main__":
    main()
