"""Offline training of the learned Poisson warm start (no counterpart in the
reference).

The samples are the (change of rhs, change of phi) pairs of consecutive
converged multigrid solves of short e-/Ar+ runs on an n x n grid (the
512 x 512 plasma by default), each run with its own B-filter peak drawn
uniformly in [0, 20] mT and its own density drawn log-uniformly in
[1e15, 1e17] m^-3.  Every run records into its own context; its samples are
appended to one sample store (lepic_net_copy_samples), where the U-Net is
trained on the GPU by the library's BF16-MFMA forward and backward kernels
(lepic_net_train).  The trained network is then copied into the context that
uses it (lepic_net_copy), on any grid: the network is convolutional and its
inputs and outputs are normalised per solve.

    python -m jd_lepic_amd.warmstart --samples 20000 --runs 40
trains on 20000 samples and evaluates the network on the 1024 x 1024
workload, printing one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from . import workloads as WL


def run_steps(ctx, W, steps, start_cold=True, seed=7, on_solve=None):
    """`steps` steps of the hot path in main.f90:689-1090 order (rhs ->
    Poisson -> E -> [sort] -> [dens_red + MCC] -> push + deposit)."""
    import jd_lepic_amd as L
    nsp = len(W["species"])
    infos = []
    for it in range(1, steps + 1):
        ctx.calc_rhs()
        start = L.START_COLD if (it == 1 and start_cold) else L.START_PREV
        info = ctx.iterative_solver(eps=1e-10, ncycl=20000, start=start)
        infos.append(info)
        if on_solve:
            on_solve(it, info)
        ctx.calc_efield()
        if it % 10 == 1:
            for p in range(1, nsp + 1):
                ctx.part_sorting(p)
        if it > 1 and it % 10 == 1:
            ctx.dens_red()
            ctx.collisions(it, mode=L.MCC_PROD, seed=seed)
        for p in range(1, nsp + 1):
            ctx.part_mover(p, deposit=True)
    return infos


def build_dataset(samples, runs, n=512, ppc=50, skip=5, seed=1, log=None):
    """A sample store holding `samples` (drhs, dphi) pairs from `runs` runs
    on the n x n argon box with randomised B peak and density.  Returns the
    store context and one record per run."""
    import jd_lepic_amd as L
    rng = np.random.default_rng(seed)
    per_run = int(math.ceil(samples / runs))
    store, meta = None, []
    for r in range(runs):
        bpeak = float(rng.uniform(0.0, 0.02))
        n0 = float(10.0 ** rng.uniform(15.0, 17.0))
        W = WL.argon_box(n, ppc, n0=n0, Bpeak=bpeak)
        if store is None:
            store = L.Context(n, n, W["h"], W["h"], W["dt"], [L.Species(-L.QE, L.ME, 1.0, 16)], ng=W["ng"],
                              precision=L.MIXED)
            store.set_boundary(W["bcnd"], np.zeros((n + 3, n + 3)))
            store.net_create(capacity=samples, seed=seed)
        t = time.perf_counter()
        with WL.build_context(W, L.MIXED, seed=int(rng.integers(1, 2**31))) as ctx:
            ctx.net_create(capacity=per_run, seed=seed)
            # recording starts after `skip` steps; a sample needs the solve before it
            run_steps(ctx, W, skip, seed=seed)
            ctx.net_recording(True)
            infos = run_steps(ctx, W, per_run + 1, start_cold=False, seed=seed)
            copied = store.net_copy_samples(ctx)
        meta.append(dict(bpeak_T=bpeak, n0=n0, samples=copied, vcycles=float(np.mean([i["vcycles"] for i in infos])),
                         seconds=time.perf_counter() - t))
        if log:
            log(f"run {r + 1}/{runs}: B {bpeak * 1e3:.1f} mT, n0 {n0:.2e}, {copied} samples, "
                f"{meta[-1]['vcycles']:.1f} V-cycles/solve, {meta[-1]['seconds']:.1f} s")
    return store, meta


def train(store, iters, batch=4, lr=2e-3):
    """Train the store's network; returns (first loss, last loss, seconds)."""
    t = time.perf_counter()
    first, last = store.net_train(iters, batch=batch, lr=lr)
    return first, last, time.perf_counter() - t


def evaluate(W, store, steps=20, warmup=10, seed=20261202):
    """V-cycles, field time and the network's residual gain on workload W,
    with the store's network (START_NET) and without it (START_PREV), from the
    same load over the same steps."""
    import jd_lepic_amd as L
    out = {}
    for mode in ("prev", "net"):
        with WL.build_context(W, L.MIXED, seed=seed) as ctx:
            if mode == "net":
                ctx.net_create(capacity=0)
                ctx.net_copy(store)
                ctx.net_policy(min_gain=1.0, max_backoff=0)   # evaluate the network on every solve
            run_steps(ctx, W, warmup)
            vc, gains, used, ms = [], [], 0, []
            nsp = len(W["species"])
            for it in range(warmup + 1, warmup + steps + 1):
                ctx.calc_rhs()
                ctx.sync()
                t = time.perf_counter()
                info = ctx.iterative_solver(eps=1e-10, ncycl=20000,
                                            start=L.START_NET if mode == "net" else L.START_PREV)
                ctx.sync()
                ms.append(1e3 * (time.perf_counter() - t))
                vc.append(info["vcycles"])
                if mode == "net" and info["res0_alt"] >= 0.0:
                    ok = info["start_used"] == L.START_NET
                    used += ok
                    gains.append(info["res0_alt"] / info["res0"] if ok else info["res0"] / info["res0_alt"])
                ctx.calc_efield()
                if it % 10 == 1:
                    for p in range(1, nsp + 1):
                        ctx.part_sorting(p)
                for p in range(1, nsp + 1):
                    ctx.part_mover(p, deposit=True)
            out[mode] = dict(vcycles_per_solve=float(np.mean(vc)), ms_per_solve=float(np.median(ms)))
            if mode == "net":
                out[mode].update(net_start_used=used / steps,
                                 res0_gain_median=float(np.median(gains)) if gains else None,
                                 res0_gain_max=float(np.max(gains)) if gains else None)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--runs", type=int, default=40)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=50)
    ap.add_argument("--iters", type=int, default=0, help="training iterations (0: one pass over the samples)")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", default="argon-1024", help="workload the trained network is evaluated on")
    ap.add_argument("--eval-steps", type=int, default=20)
    args = ap.parse_args()

    def log(msg):
        print(f"[warmstart] {msg}", flush=True)

    t = time.perf_counter()
    store, meta = build_dataset(args.samples, args.runs, n=args.n, ppc=args.ppc, log=log)
    t_data = time.perf_counter() - t
    iters = args.iters or max(1, store.net_samples() // args.batch)
    log(f"{store.net_samples()} samples in {t_data:.1f} s; training {iters} iterations")
    first, last, t_train = train(store, iters, batch=args.batch, lr=args.lr)
    log(f"loss {first:.4g} -> {last:.4g} in {t_train:.1f} s; evaluating on {args.eval}")
    ev = evaluate(WL.CONFIGS[args.eval](), store, steps=args.eval_steps)
    store.close()
    print(json.dumps(dict(samples=int(sum(m["samples"] for m in meta)), runs=args.runs, grid=args.n, ppc=args.ppc,
                          dataset_s=t_data, train_iters=iters, train_s=t_train, loss_first=first, loss_last=last,
                          evaluation=dict(workload=args.eval, **ev), runs_detail=meta)))


if __name__ == "__main__":
    main()
