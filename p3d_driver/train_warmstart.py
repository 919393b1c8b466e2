"""Trains the warm-start network of one problem family on the GPU and writes
its weights (python -m pot3d_amd.train_warmstart --domain pfss|open).

1. Seeded synthetic Br maps on the coarse grid (51x91x181 by default),
   seeds first_seed .. first_seed+pairs-1 (disjoint from the evaluation
   seeds 0..999 of the workloads).
2. Each map solved by this library's PCG (p3d_solve, ILU0, FP32 recurrence
   with the FP64 solution, ||b-Ax||2/||b||2 <= 1e-9): the training pairs.
3. Adam on the device by the library's forward/backward kernels, one pair
   per step, with the exact symmetries of the discrete problem as
   augmentation (phi rotation, theta reflection, sign), the learning rate
   stepping down to lr/4 and lr/16 for the last 40% and 15% of the steps.
4. The last `val` pairs are held out: their guess error and the PCG
   iterations from zero and from the guess are stored with the weights.

With --monopole 1 (the default for 'open', whose |Br| maps have a large
mean) the network learns the maps' non-monopole part only; the monopole's
response comes from one PCG solve for br0 = 1 (pot3d_amd.warmstart).
"""
from __future__ import annotations

import argparse
import math
import os
import sys
import time

import numpy as np


def _map(job):
    n_bipoles, seed = job
    from p3d_driver import inputs
    return inputs.synthetic_br_map(seed, n_bipoles)


def build_maps(domain, seeds, workers):
    """The 1-degree Br map of every seed, on the host before the GPU is
    touched (a fork-based pool)."""
    from p3d_driver import workloads
    jobs = [(workloads.DOMAINS[domain]["n_bipoles"], s) for s in seeds]
    if workers <= 1:
        return [_map(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_map, jobs, chunksize=8)


def symmetric_theta(mesh) -> bool:
    th = np.asarray(mesh.th)
    return bool(np.allclose(th + th[::-1], math.pi, rtol=0, atol=1e-12))


def uniform_phi(mesh) -> bool:
    dph = np.asarray(mesh.dph)
    return bool(np.ptp(dph) <= 1e-12 * np.max(dph))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--domain", choices=("pfss", "open"), required=True)
    ap.add_argument("--grid", default="51x91x181")
    ap.add_argument("--pairs", type=int, default=2000, help="PCG-solved pairs (incl. held out)")
    ap.add_argument("--val", type=int, default=48, help="pairs held out for validation")
    ap.add_argument("--first-seed", type=int, default=1000)
    ap.add_argument("--epochs", type=float, default=8.0)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--augment", type=int, default=1)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--monopole", type=int, default=-1,
                    help="split br0's area-weighted mean off (1), or not (0); default 1 for "
                         "open, 0 for pfss")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    from p3d_driver import workloads
    grid = tuple(int(v) for v in args.grid.split("x"))
    seeds = list(range(args.first_seed, args.first_seed + args.pairs))
    t0 = time.perf_counter()
    maps = build_maps(args.domain, seeds, args.workers)
    print(f"[train] {len(maps)} {args.domain} maps in {time.perf_counter() - t0:.1f}s",
          flush=True)

    from pot3d_amd import _lib, warmstart as W
    from pot3d_amd.solver import Pot3dSolver
    w0 = workloads.domain_workload(args.domain, grid, 0)
    # br0 on the coarse mesh: the driver's readbr + set_flux (p3d_readbr)
    opt = workloads.DOMAINS[args.domain]["option"]
    brs = [workloads.readbr(w0.mesh, tn, pn, bn, opt) for tn, pn, bn in maps]
    del maps
    s = Pot3dSolver(w0.mesh, w0.pm_r1, 2)
    s.set_option(_lib.OPT_STOP, 1)
    s.set_option(_lib.OPT_MIXED, 2)
    xs, iters = [], []
    t0 = time.perf_counter()
    for q, b in enumerate(brs):
        s.set_br0(b)
        x = np.zeros(s.N)
        ierr, ncg, _ = s.solve(x, None, 1e-9, ncgmax=100000)
        if ierr != 0:
            raise RuntimeError(f"pair {seeds[q]}: the solve did not converge")
        xs.append(x)
        iters.append(ncg)
        if (q + 1) % 200 == 0:
            print(f"[train] {q + 1} pairs solved ({time.perf_counter() - t0:.1f}s, "
                  f"{np.mean(iters):.0f} iterations each)", flush=True)
    t_pairs = time.perf_counter() - t0
    ntr = len(brs) - args.val
    tr_b, tr_x = brs[:ntr], xs[:ntr]
    va_b, va_x = brs[ntr:], xs[ntr:]

    net = W.WarmStartNet(w0.mesh, seed=args.init_seed)
    mono = args.monopole if args.monopole >= 0 else int(args.domain == "open")
    if mono:
        net.set_monopole(W.monopole_response(w0.mesh, w0.pm_r1), W.area_weights(w0.mesh))
    net.set_scales(*W.fit_scales((net.split(b)[0] for b in tr_b),
                                 (net.split(b, x)[1] for b, x in zip(tr_b, tr_x))))
    for b, x in zip(tr_b, tr_x):
        net.add_sample(b, x)
    aug = (uniform_phi(w0.mesh), symmetric_theta(w0.mesh), True) if args.augment else (0, 0, 0)
    net.set_augment(*aug)
    total = int(args.epochs * ntr)
    n1, n2 = int(0.6 * total), int(0.25 * total)
    phases = [(n1, args.lr), (n2, args.lr / 4), (total - n1 - n2, args.lr / 16)]
    losses = []
    t0 = time.perf_counter()
    chunk = 2000
    for ph, (n, lr) in enumerate(phases):
        done = 0
        while done < n:
            m = min(chunk, n - done)
            losses.append(net.train(m, lr, seed=1000 * ph + done))
            done += m
            print(f"[train] phase {ph} lr {lr:.1e}: {done}/{n} steps, loss "
                  f"{float(np.mean(losses[-1][-200:])):.3e} ({time.perf_counter() - t0:.1f}s)",
                  flush=True)
    t_train = time.perf_counter() - t0
    losses = np.concatenate(losses)

    # held-out pairs: guess error, and PCG iterations from zero / from the guess
    errs, it0, itw = [], [], []
    s.set_option(_lib.OPT_GUESS_CHECK, 2)
    for q, (b, x) in enumerate(zip(va_b, va_x)):
        g = net.predict(b)
        errs.append(float(np.linalg.norm(g - x) / np.linalg.norm(x)))
        if q < 16:
            s.set_br0(b)
            it0.append(iters[ntr + q])
            xw = g.copy()
            assert s.solve(xw, None, 1e-9, ncgmax=100000)[0] == 0
            itw.append(s.stats()["ncg"])
    print(f"[train] held out: guess rel err {np.mean(errs):.4f} (min {np.min(errs):.4f}, "
          f"max {np.max(errs):.4f}); iterations zero {np.mean(it0):.1f} -> guess "
          f"{np.mean(itw):.1f}", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    d = workloads.DOMAINS[args.domain]
    net.save(args.out, domain=args.domain, option=d["option"], r0=d["r0"], r1=d["r1"],
             n_bipoles=d["n_bipoles"], pairs=ntr, held_out=args.val, first_seed=args.first_seed,
             epochs=args.epochs, lr=args.lr, augment=np.array(aug, dtype=np.int32),
             steps=total, loss_first=float(np.mean(losses[:200])),
             loss_last=float(np.mean(losses[-200:])), val_guess_rel_err=float(np.mean(errs)),
             val_iters_zero=float(np.mean(it0)), val_iters_guess=float(np.mean(itw)),
             t_pair_solves_s=t_pairs, t_train_s=t_train)
    print(f"[train] wrote {args.out}: {total} steps in {t_train:.1f}s, pairs in "
          f"{t_pairs:.1f}s", flush=True)
    net.close()
    s.close()


if __name__ == "__main__":
    sys.exit(main())
