"""Benchmark: PCG time-to-solution of the POT3D Laplace solve on one MI355X.

Workload: 151x301x601 PFSS (26.7 M unknowns), r in [1,2.5] with a geometric
radial mesh, uniform theta/phi, seeded synthetic Br (8 bipoles + l<=15
background), phi=0 at r=2.5, ILU0 preconditioner, stopping at
||b - Ax||2/||b||2 <= 1e-9.  A step is one complete solve from the zero
guess with the vectors already resident in HBM.

Reported:
  value        time-to-solution of the fastest path (mixed precision: FP32
               inner ILU0-PCG in FP64 defect correction), seconds
  fp64_plain   the same solve with the plain FP64 ILU0 PCG (the drop-in
               baseline), seconds, and the speed-up
  roofline     the dominant kernel's achieved HBM bandwidth (algorithmic
               bytes / HIP-event time on the solver's stream) vs 8 TB/s
  cpu_baseline the CPU oracle's ILU0 PCG (the reference's CPU algorithm:
               Saad ilu0 + lusol) timed on a bounded sample on the host
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

HBM_PEAK_GBS = 8000.0


class Hip:
    """Minimal HIP runtime access for device buffers (no torch needed)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = C.c_void_p()
        if self.L.hipMalloc(C.byref(p), C.c_size_t(n)) != 0:
            raise MemoryError("hipMalloc failed")
        return p.value

    def h2d(self, d, h: np.ndarray):
        assert self.L.hipMemcpy(C.c_void_p(d), h.ctypes.data, h.nbytes, 1) == 0

    def d2h(self, h: np.ndarray, d):
        assert self.L.hipMemcpy(h.ctypes.data, C.c_void_p(d), h.nbytes, 2) == 0

    def free(self, d):
        self.L.hipFree(C.c_void_p(d))

    def sync(self):
        self.L.hipDeviceSynchronize()


def cpu_baseline(workload, gpu_iters: int, sample_iters: int, threads: int):
    """The oracle's ILU0 PCG (ifprec=2: diacsr + Saad ilu0 + lusol, the
    reference's CPU path) on the same problem: set-up plus `sample_iters`
    iterations timed, extrapolated to the GPU's iteration count."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    o = oracle.Oracle(workload.mesh, workload.pm_r1)
    t0 = time.perf_counter()
    b = o.rhs(workload.br0)
    o.coefs()
    t_coef = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.cgsolve(b, ifprec=2, epscg=1e-30, ncgmax=0)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.cgsolve(b, ifprec=2, epscg=1e-30, ncgmax=sample_iters)
    t_run = time.perf_counter() - t0 - t_setup
    per_iter = t_run / sample_iters
    est = t_coef + t_setup + per_iter * gpu_iters
    return {"value": round(est, 3), "unit": "s (estimated time-to-solution)", "cores": threads,
            "kind": "port",
            "sample": (f"oracle ILU0 PCG on the same 151x301x601 problem: matrix+ILU0 set-up "
                       f"({t_coef + t_setup:.2f} s) and {sample_iters} iterations "
                       f"({per_iter:.3f} s/iteration, {threads} OpenMP threads for the "
                       f"operator, sequential trisolves), extrapolated to the GPU's "
                       f"{gpu_iters} iterations")}


# (name, time_kernel id, algorithmic bytes per unknown, padded-array bytes per point)
KERNELS = [
    ("stencil_fp64 (k_stencil_march<double>)", 0, 8, 8),
    ("stencil_fp32 (k_stencil_march<float>)", 1, 4, 4),
    ("ilu_apply_fp64 (k_tile_fwd+k_tile_bwd<double>)", 2, 32, 0),
    ("ilu_apply_fp32 (k_tile_fwd+k_tile_bwd<float>)", 7, 24, 0),
    ("cg_update_fp64 (k_cg_update<double>)", 4, 48, 0),
    ("p_update_fp64 (k_p_update<double>)", 5, 24, 0),
    ("copy_fp64 (k_copy<double>)", 6, 16, 0),
]


def measure_kernels(s, N: int, padded: int, reps: int = 20):
    """Average HIP-event time per launch on the solver's stream, algorithmic
    bytes per launch and the achieved bandwidth of each kernel."""
    table = {}
    for name, kid, per_n, per_pad in KERNELS:
        ms = s.time_kernel(kid, reps if kid not in (2, 7) else 5)
        nbytes = per_n * N + per_pad * padded
        table[name] = {"ms": round(ms, 5), "bytes": nbytes,
                       "GBs": round(nbytes / (ms * 1e-3) / 1e9, 1)}
    # the dominant kernel of the solve: the ILU apply or the operator, by time
    cand = {k: v for k, v in table.items() if k.startswith(("ilu_apply_fp64", "stencil_fp64"))}
    name = max(cand, key=lambda k: cand[k]["ms"])
    return {"table": table, "dominant": dict(name=name, **table[name])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=str, default="151x301x601")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=12)
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the kernels (for profiler runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(--gpus {args.gpus} requested); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from pot3d_amd import _lib, solver, workloads
    nr, nt, np_ = (int(v) for v in args.grid.split("x"))
    t0 = time.perf_counter()
    w = workloads.pfss(nr, nt, np_, seed=0, ifprec=2, name=f"pfss_{args.grid}")
    print(f"[bench] workload {w.name} N={w.N} built in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    hip = Hip()
    N = w.N
    s = solver.Pot3dSolver(w.mesh, w.pm_r1, 2)
    s.set_br0(w.br0)
    s.set_option(_lib.OPT_STOP, 1)
    rhs = s.rhs()
    if args.kernels_only:
        kern = measure_kernels(s, N, nr * nt * np_, reps=5)
        print(json.dumps({"workload": w.name, "unknowns": N, "kernels": kern["table"]}))
        s.close()
        return
    d_x = hip.malloc(8 * N)
    d_b = hip.malloc(8 * N)
    zero = np.zeros(N)

    def run(mixed: bool):
        s.set_option(_lib.OPT_MIXED, 1 if mixed else 0)
        hip.h2d(d_x, zero)
        hip.h2d(d_b, rhs)
        hip.sync()
        t = time.perf_counter()
        ierr, ncg, epsn = s.solve_dev(d_x, d_b, 1e-9, ncgmax=20000)
        hip.sync()
        dt = time.perf_counter() - t
        st = s.stats()
        assert ierr == 0, f"solve did not converge ({st})"
        return dt, st

    # FP64 plain ILU0 PCG (drop-in baseline)
    for _ in range(args.warmup):
        run(False)
    t64 = []
    for _ in range(args.steps):
        dt, st64 = run(False)
        t64.append(dt)
        print(f"[bench] fp64 ILU0 PCG: {dt:.3f}s ncg={st64['ncg']} "
              f"relres={st64['true_relres']:.2e}", file=sys.stderr, flush=True)
    x64 = np.empty(N)
    hip.d2h(x64, d_x)
    # mixed precision
    for _ in range(args.warmup):
        run(True)
    tmx = []
    for _ in range(args.steps):
        dt, stmx = run(True)
        tmx.append(dt)
        print(f"[bench] mixed ILU0 PCG: {dt:.3f}s ncg={stmx['ncg']} outer={stmx['n_outer']} "
              f"relres={stmx['true_relres']:.2e}", file=sys.stderr, flush=True)
    xmx = np.empty(N)
    hip.d2h(xmx, d_x)
    rel_x = float(np.max(np.abs(xmx - x64)) / np.max(np.abs(x64)))

    kern = measure_kernels(s, N, nr * nt * np_)
    dom = kern["dominant"]
    roof = {"kernel": dom["name"], "bound": "hbm", "achieved": dom["GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(dom["GBs"] / HBM_PEAK_GBS, 4), "traffic": None,
            "avg_ms": dom["ms"], "bytes_per_launch": dom["bytes"]}
    per_iter64 = min(t64) / max(st64["ncg"], 1)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(w, st64["ncg"], args.cpu_iters, min(16, os.cpu_count() or 1))
        except Exception as e:  # reported, not fatal
            cpu = {"error": str(e)}
    best = min(tmx)
    out = {
        "metric": "PCG time-to-||r||/||b||<=1e-9 (s): mixed+learned vs FP64 plain; "
                  "stencil HBM GB/s",
        "value": round(best, 4), "unit": "s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * float(np.mean(tmx)), 2),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "f64 outer / f32 inner", "data": "synthetic (seeded bipole+harmonic Br)",
        "config": {"workload": f"PFSS {args.grid} ILU0 PCG to ||r||2/||b||2<=1e-9",
                   "unknowns": N, "precond": "ILU0 (D-ILU)", "epscg": 1e-9},
        "fp64_plain_s": round(min(t64), 4), "fp64_plain_ncg": st64["ncg"],
        "mixed_ncg": stmx["ncg"], "mixed_outer": stmx["n_outer"],
        "speedup_vs_fp64": round(min(t64) / best, 3),
        "mixed_vs_fp64_max_rel_dx": rel_x,
        "fp64_ms_per_iter": round(1e3 * per_iter64, 4),
        "kernels": kern["table"],
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    s.close()
    hip.free(d_x)
    hip.free(d_b)


if __name__ == "__main__":
    main()
