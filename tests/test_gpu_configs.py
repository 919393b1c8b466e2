"""GPU solves at the benchmark configurations' own workloads, through the C
ABI, against the CPU oracle (where it finishes in seconds) and against the
FP64 GPU solve (full sizes, size-independent properties).

  51x91x181 PFSS ('ss', phi=0 at r1), FP64 PCG with the diagonal
        preconditioner (ifprec=1): against the oracle's cgsolve.
  51x91x181 with ILU0, FP32 inner PCG inside FP64 defect
        correction (both mixed modes), and the Neumann variant ('potential',
        dphi/dr=0 at r1, src/pot3d.F90:1250-1255): against the oracle's FP64
        solve; the Neumann system is singular (constants), so phi is compared
        after removing its mean.
  151x301x601 PFSS, ILU0: the FP64 solve's residual recomputed by the
        oracle's operator; mixed precision with the block-Jacobi ILU0 against
        the FP64 solve.
  301x601x1201 'open' (r in [2.5, 10], 16 bipoles): mixed precision
        with the block-Jacobi ILU0 against the FP64 solve.

Bars (the accuracy bars of the solve at epscg = 1e-9): potential and B-field (getb,
src/pot3d.F90:6187-6233) within 1e-8 relative to their largest magnitude,
magnetic energy (:6235-6321) within 1e-7 relative, recomputed FP64
||b - A x||2/||b||2 <= 1e-9 where the stop is on it (the mixed solves) and
<= 1.05e-9 where the stop is on the recurrence residual (FP64 PCG).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PHI_TOL = 1e-8
B_TOL = 1e-8
E_TOL = 1e-7


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _maxrel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _check_fields(s, phi, phi_ref, B_ref, w_ref, remove_mean=False, label=""):
    """phi, B and energy of the GPU solver's current phi against a reference."""
    if remove_mean:
        d = _maxrel(phi - phi.mean(), phi_ref - phi_ref.mean())
    else:
        d = _maxrel(phi, phi_ref)
    B = s.getb(phi)
    dB = max(_maxrel(g, r) for g, r in zip(B, B_ref))
    w = s.magnetic_energy(phi)
    dE = float(np.max(np.abs(w - w_ref) / np.abs(w_ref)))
    say(f"    {label}: phi {d:.2e}  B {dB:.2e}  energy {dE:.2e}")
    assert d <= PHI_TOL and dB <= B_TOL and dE <= E_TOL, (d, dB, dE)
    return d, dB, dE


def _set(s, **opts):
    from pot3d_amd import _lib
    for k, v in opts.items():
        s.set_option(getattr(_lib, "OPT_" + k.upper()), v)


def test_pfss_51x91x181_fp64_diag_vs_oracle(oracle_mod):
    """51x91x181 PFSS: FP64 PCG, diagonal preconditioner, the reference's epsn stop."""
    from pot3d_amd import solver
    from p3d_driver import workloads
    w = workloads.config(1)
    say(f"[gpu] {w.name} N={w.N}")
    o = oracle_mod.Oracle(w.mesh, w.pm_r1)
    b = o.rhs(w.br0)
    ref = o.cgsolve(b, ifprec=1, epscg=1e-9)
    phir = o.fill_phi(ref["x"], w.br0)
    with solver.Pot3dSolver(w.mesh, w.pm_r1, 1) as s:
        s.set_br0(w.br0)
        x = np.zeros(w.N)
        ierr, ncg, epsn = s.solve(x, None, 1e-9, ncgmax=20000)
        say(f"    ncg {ncg} (oracle {ref['ncg']}), epsn {epsn:.3e}")
        assert ierr == 0 and epsn < 1e-9 and abs(ncg - ref["ncg"]) <= 2
        assert _maxrel(x, ref["x"]) <= PHI_TOL
        phi = s.fill_phi()
        _check_fields(s, phi, phir, o.getb(phir), o.energy(phir), label="fp64 diag")
        assert s.stats()["true_relres"] < 1e-6


@pytest.mark.parametrize("option", ["ss", "potential"])
@pytest.mark.parametrize("mixed", [1, 2])
def test_51x91x181_mixed_ilu_vs_oracle(oracle_mod, option, mixed):
    """51x91x181: FP32 inner ILU0-PCG (inner tolerance 1e-5 in mode 1) inside
    FP64 defect correction to ||b - A x||2/||b||2 <= 1e-9, on the 51x91x181 grid,
    with phi=0 at r1 ('ss') and the Neumann variant ('potential')."""
    from pot3d_amd import solver
    from p3d_driver import workloads
    w = workloads.pfss(51, 91, 181, seed=0, option=option, ifprec=2)
    say(f"[gpu] {w.name} option={option} mixed={mixed}")
    o = oracle_mod.Oracle(w.mesh, w.pm_r1)
    b = o.rhs(w.br0)
    ref = o.cgsolve(b, ifprec=3, epscg=1e-12)
    assert ref["ierr"] == 0
    phir = o.fill_phi(ref["x"], w.br0)
    Br = o.getb(phir)
    wr = o.energy(phir)
    singular = option == "potential"
    with solver.Pot3dSolver(w.mesh, w.pm_r1, 2) as s:
        s.set_br0(w.br0)
        _set(s, stop=1, mixed=mixed, inner_tol=1e-5)
        x = np.zeros(w.N)
        ierr, ncg, epsn = s.solve(x, None, 1e-9, ncgmax=20000)
        st = s.stats()
        say(f"    ncg {ncg} outer {st['n_outer']} true relres {st['true_relres']:.3e} "
            f"(oracle FP64 ncg {ref['ncg']} to epsn 1e-12)")
        assert ierr == 0 and st["true_relres"] <= 1e-9 and st["n_outer"] >= 1
        # the oracle's operator agrees on the residual
        rr = np.linalg.norm(b - o.ax(x)) / np.linalg.norm(b)
        assert rr <= 1.05e-9, rr
        phi = s.fill_phi()
        _check_fields(s, phi, phir, Br, wr, remove_mean=singular,
                      label=f"mixed {mixed} {option}")
        # the FP64 GPU solve of the same problem, for the iteration count
        _set(s, mixed=0)
        x64 = np.zeros(w.N)
        assert s.solve(x64, None, 1e-9, ncgmax=20000)[0] == 0
        n64 = s.stats()["ncg"]
        say(f"    fp64 GPU ncg {n64}")
        assert ncg <= 1.3 * n64 + 20


def _fp64_then_mixed(w, blocks, label, oracle_check=None):
    from pot3d_amd import solver
    with solver.Pot3dSolver(w.mesh, w.pm_r1, 2) as s:
        s.set_br0(w.br0)
        _set(s, stop=1)
        x64 = np.zeros(w.N)
        ierr, ncg64, _ = s.solve(x64, None, 1e-9, ncgmax=20000)
        st = s.stats()
        say(f"    {label} fp64 ILU0: ncg {ncg64} true relres {st['true_relres']:.3e} "
            f"t {st['t_solve']:.2f}s")
        # the FP64 PCG stops on its recurrence residual, as the reference's
        # does; the recomputed b - Ax drifts from it by the recurrence's
        # rounding (a few parts in 1e3 of epscg over ~2000 iterations)
        assert ierr == 0 and st["true_relres"] <= 1.05e-9
        if oracle_check:
            oracle_check(s, x64)
        phi64 = s.fill_phi()
        B64 = s.getb(phi64)
        w64 = s.magnetic_energy(phi64)
        _set(s, ilu_blocks_r=blocks[0], ilu_blocks_t=blocks[1], ilu_blocks_p=blocks[2],
             tile_k=16, tile_stage=2, mixed=2)
        x = np.zeros(w.N)
        ierr, ncg, _ = s.solve(x, None, 1e-9, ncgmax=20000)
        st = s.stats()
        say(f"    {label} mixed block-Jacobi {blocks}: ncg {ncg} outer {st['n_outer']} "
            f"true relres {st['true_relres']:.3e} t {st['t_solve']:.2f}s")
        assert ierr == 0 and st["true_relres"] <= 1e-9
        phi = s.fill_phi()
        del x
        _check_fields(s, phi, phi64, B64, w64, label=f"{label} mixed vs fp64")


def test_151x301x601_fp64_residual_and_mixed(oracle_mod):
    """151x301x601: the FP64 ILU0 solve's residual recomputed by the
    oracle's operator; mixed block-Jacobi solve against it."""
    from p3d_driver import workloads
    w = workloads.config(2)
    say(f"[gpu] {w.name} N={w.N}")

    def oracle_residual(s, x):
        o = oracle_mod.Oracle(w.mesh, w.pm_r1)
        b = o.rhs(w.br0)
        rr = np.linalg.norm(b - o.ax(x)) / np.linalg.norm(b)
        say(f"    oracle ||b - A x||/||b|| = {rr:.3e}")
        assert rr <= 1.05e-9

    _fp64_then_mixed(w, (8, 16, 16), "151x301x601", oracle_residual)


@pytest.mark.timeout(600)
def test_301x601x1201_open_mixed_vs_fp64():
    """301x601x1201 'open': mixed block-Jacobi solve against the FP64
    ILU0 solve (phi, B, energy; recomputed residuals of both)."""
    from p3d_driver import workloads
    w = workloads.config(5)
    say(f"[gpu] {w.name} N={w.N}")
    _fp64_then_mixed(w, (8, 16, 16), "301x601x1201")


def test_301x601x1201_ilu_launch_schedules_same_bits():
    """301x601x1201, block-Jacobi ILU0 on register tiles (8x16x16 blocks,
    16-plane tiles, 20 tile diagonals per sweep): one launch per diagonal,
    one block-persistent launch per sweep, and the small end diagonals
    block-persistent with one launch per large diagonal give the same
    apply bit for bit, and the same mixed solve."""
    from pot3d_amd import solver
    from p3d_driver import workloads
    w = workloads.config(5)
    say(f"[gpu] {w.name} N={w.N} ILU launch schedules")
    with solver.Pot3dSolver(w.mesh, w.pm_r1, 2) as s:
        _set(s, ilu_blocks_r=8, ilu_blocks_t=16, ilu_blocks_p=16, tile_k=16, tile_stage=2)
        x = np.random.default_rng(5).standard_normal(w.N)
        z = {}
        for mode in (0, 1, 2):
            _set(s, block_sweep=mode)
            z[mode] = s.prec_inv(x)
        assert np.array_equal(z[1], z[0]) and np.array_equal(z[2], z[0])
        del z, x
        s.set_br0(w.br0)
        _set(s, stop=1, mixed=2)
        xs = {}
        for mode in (0, 2):
            _set(s, block_sweep=mode)
            xs[mode] = np.zeros(w.N)
            assert s.solve(xs[mode], None, 1e-9, ncgmax=20000)[0] == 0
            say(f"    block_sweep {mode}: ncg {s.stats()['ncg']} t {s.stats()['t_solve']:.2f}s")
        assert np.array_equal(xs[2], xs[0])
        _set(s, block_sweep=-1)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("wl,family", [(4, "pfss"), (5, "open")])
def test_learned_warm_start_real_grids(wl, family):
    """151x301x601 PFSS and 301x601x1201 'open': the family's
    network trained on 51x91x181 PCG solutions (committed weights, held-out
    seed 0) predicts on 51x91x181, is prolonged to the configuration's grid
    and starts the value path's solve (mixed precision, block-Jacobi ILU0):
    the guess is kept, the answer is the zero guess's (phi, B, energy) and
    the solve takes fewer iterations."""
    import os
    from p3d_driver import workloads
    from pot3d_amd import solver, warmstart as W
    wpath = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                         "pot3d_amd", "weights", f"warmstart_{family}_51x91x181.npz")
    if not os.path.exists(wpath):
        pytest.skip(f"no trained weights at pot3d_amd/weights/ (python -m "
                    f"p3d_driver.train_warmstart --domain {family} writes them)")
    w = workloads.config(wl)
    wc = workloads.domain_workload(family, workloads.COARSE_GRID, 0)
    say(f"[gpu] {w.name} N={w.N} warm start from {os.path.basename(wpath)}")
    net, meta = W.WarmStartNet.load(wpath, wc.mesh, wc.pm_r1)
    with solver.Pot3dSolver(wc.mesh, wc.pm_r1, 2) as sc:
        sc.set_br0(wc.br0)
        phic = sc.fill_phi(net.predict(wc.br0))
    net.close()
    xg = W.prolong(W.prolong_tables(wc.mesh, w.mesh), phic)
    with solver.Pot3dSolver(w.mesh, w.pm_r1, 2) as s:
        s.set_br0(w.br0)
        _set(s, stop=1, mixed=2, ilu_blocks_r=8, ilu_blocks_t=16, ilu_blocks_p=16, tile_k=16,
             tile_stage=2)
        x0 = np.zeros(w.N)
        assert s.solve(x0, None, 1e-9, ncgmax=20000)[0] == 0
        n0 = s.stats()["ncg"]
        phi0 = s.fill_phi(x0)
        B0 = s.getb(phi0)
        w0 = s.magnetic_energy(phi0)
        _set(s, guess_check=2)
        x = xg.copy()
        assert s.solve(x, None, 1e-9, ncgmax=20000)[0] == 0
        st = s.stats()
        e = float(np.linalg.norm(xg - x0) / np.linalg.norm(x0))
        say(f"    zero guess {n0} iterations; learned guess (rel err {e:.3f}) {st['ncg']} "
            f"iterations, guess {st['guess']}, true relres {st['true_relres']:.2e}")
        assert st["guess"] == 1 and st["ncg"] < n0 and st["true_relres"] <= 1e-9
        _check_fields(s, s.fill_phi(x), phi0, B0, w0, label=f"{w.name} learned vs zero guess")
