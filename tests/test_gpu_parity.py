"""GPU parity: every operation of libpot3d_hip.so against the CPU oracle on
the same seeded inputs, through the C ABI.

Tolerances: FP64 operator/rhs/phi/energy to 1e-13 relative to the natural
scale (the GPU computes the coefficients from separable factor tables and
may contract products into FMAs, so last-bit differences are expected);
FP32 preconditioners to 2e-5 relative; solves to the solve bar
(potential 1e-8 relative, energy 1e-7).
"""
import sys

import numpy as np
import pytest

from conftest import ROOT, random_br0, small_mesh

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solver_mod():
    from pot3d_amd import solver
    return solver


MESHES = {
    "tiny": dict(nr=9, nt=8, np_=12),
    "odd": dict(nr=13, nt=11, np_=18, drratio=2.0),
    "mid": dict(nr=33, nt=45, np_=70),
}


def _mesh(name):
    return small_mesh(**MESHES[name])


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("pm_r1", [-1.0, 1.0])
def test_operator(oracle_mod, solver_mod, name, pm_r1):
    say(f"[gpu] operator mesh={name} pm_r1={pm_r1}")
    mesh = _mesh(name)
    o = oracle_mod.Oracle(mesh, pm_r1)
    with solver_mod.Pot3dSolver(mesh, pm_r1, 1) as s:
        rng = np.random.default_rng(7)
        for _ in range(2):
            x = rng.standard_normal(o.N)
            y = s.ax(x)
            ref = o.ax(x)
            scale = np.abs(o.coefs()).sum(axis=0) * np.abs(x).max()
            assert np.max(np.abs(y - ref) / scale) < 1e-14


# grids for the operator kernels' layouts: a < 62 and > 62 (several waves
# along i), b not a multiple of 8 (a partial row block), few phi planes
# (lookahead past the segment wraps more than once)
KERNEL_GRIDS = [(9, 8, 12), (13, 11, 7), (33, 45, 70), (70, 27, 40), (140, 19, 23),
                (131, 67, 9)]


@pytest.mark.parametrize("dims", KERNEL_GRIDS)
@pytest.mark.parametrize("pm_r1", [-1.0, 1.0])
def test_operator_kernels_agree(oracle_mod, solver_mod, dims, pm_r1):
    """The row-blocked and the LDS-staged operator kernels (P3D_OPT_STENCIL=1,
    and 3 and 4: tall and short patches) give the column march's values bit
    for bit, in FP64 and in the FP32 operator of the mixed recurrence; the
    FP64 values against the oracle as test_operator."""
    from pot3d_amd import _lib
    say(f"[gpu] operator kernels {dims} pm_r1={pm_r1}")
    nr, nt, np_ = dims
    mesh = small_mesh(nr=nr, nt=nt, np_=np_, drratio=1.7)
    o = oracle_mod.Oracle(mesh, pm_r1)
    with solver_mod.Pot3dSolver(mesh, pm_r1, 1) as s:
        x = np.random.default_rng(11).standard_normal(o.N)
        for fp32 in (0, 1):
            s.set_option(_lib.OPT_APPLY_FP32, fp32)
            s.set_option(_lib.OPT_STENCIL, 0)
            y0 = s.ax(x)
            ref = o.ax(x)
            scale = np.abs(o.coefs()).sum(axis=0) * np.abs(x).max()
            tol = 1e-14 if not fp32 else 1e-6
            for kernel in (1, 3, 4):
                s.set_option(_lib.OPT_STENCIL, kernel)
                y1 = s.ax(x)
                assert np.array_equal(y0.view(np.uint64), y1.view(np.uint64)), (fp32, kernel)
                assert np.max(np.abs(y1 - ref) / scale) < tol


@pytest.mark.parametrize("name", ["tiny", "mid"])
def test_rhs_phi_energy(oracle_mod, solver_mod, name):
    say(f"[gpu] rhs/phi/energy mesh={name}")
    mesh = _mesh(name)
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 3)
    with solver_mod.Pot3dSolver(mesh, -1.0, 1) as s:
        s.set_br0(br0)
        b = s.rhs()
        bref = o.rhs(br0)
        assert np.max(np.abs(b - bref)) <= 1e-14 * np.max(np.abs(bref))
        x = np.random.default_rng(1).standard_normal(o.N)
        phi = s.fill_phi(x)
        pref = o.fill_phi(x, br0)
        assert np.max(np.abs(phi - pref)) <= 1e-14 * np.max(np.abs(pref))
        w = s.magnetic_energy()
        wref = o.energy(pref)
        assert np.allclose(w, wref, rtol=1e-12, atol=0)
        br, bt, bp = s.getb()
        rbr, rbt, rbp = o.getb(pref)
        for g_, r_ in ((br, rbr), (bt, rbt), (bp, rbp)):
            assert np.max(np.abs(g_ - r_)) <= 1e-13 * np.max(np.abs(r_))


@pytest.mark.parametrize("name", list(MESHES))
def test_diag_preconditioner(oracle_mod, solver_mod, name):
    say(f"[gpu] diag prec mesh={name}")
    mesh = _mesh(name)
    o = oracle_mod.Oracle(mesh, -1.0)
    with solver_mod.Pot3dSolver(mesh, -1.0, 1) as s:
        x = np.random.default_rng(2).standard_normal(o.N)
        assert np.max(np.abs(s.prec_inv(x) - o.prec_apply(x, 1))) <= 1e-6 * np.max(
            np.abs(o.prec_apply(x, 1)))


@pytest.mark.parametrize("name", list(MESHES))
def test_ilu_pivots_and_apply(oracle_mod, solver_mod, name):
    say(f"[gpu] ilu mesh={name}")
    mesh = _mesh(name)
    o = oracle_mod.Oracle(mesh, -1.0)
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        dinv = s.ilu_dinv()
        dref = o.dilu_dinv()
        assert np.max(np.abs(dinv - dref) / np.abs(dref)) < 1e-6
        x = np.random.default_rng(4).standard_normal(o.N)
        z = s.prec_inv(x)
        zref = o.dilu_apply(x.astype(np.float32)).astype(np.float64)
        assert np.max(np.abs(z - zref)) <= 2e-5 * np.max(np.abs(zref))
        # the reference's own CPU ILU0 path (Saad ilu0 + lusol)
        zs = o.prec_apply(x, 2)
        assert np.max(np.abs(z - zs)) <= 1e-4 * np.max(np.abs(zs))


@pytest.mark.parametrize("dims", [(9, 8, 12), (13, 11, 18), (21, 31, 48), (40, 19, 30),
                                  (33, 45, 70)])
def test_ilu_tile_sweeps_match_hyperplanes(solver_mod, dims):
    """The triangular-solve schedules (tile diagonals of every tile shape,
    one launch per hyperplane) do the same arithmetic per point: results
    must be bit-identical, on every repeat."""
    say(f"[gpu] ilu tile sweeps vs hyperplanes dims={dims}")
    from pot3d_amd import _lib
    mesh = small_mesh(*dims)
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        rng = np.random.default_rng(8)
        for rep in range(3):
            x = rng.standard_normal(s.N)
            s.set_option(_lib.OPT_SWEEP, 1)
            z_lv = s.prec_inv(x)
            s.set_option(_lib.OPT_SWEEP, 0)
            z_tl = s.prec_inv(x)
            z_tl2 = s.prec_inv(x)
            assert np.array_equal(z_tl, z_lv), (rep, np.max(np.abs(z_tl - z_lv)))
            assert np.array_equal(z_tl, z_tl2)
            # other tile shapes: waves per tile x planes per tile, and
            # single-wave tiles staged by the marching wave alone
            for waves, tk, stage in ((1, 16, 0), (1, 64, 0), (4, 16, 0), (4, 32, 0),
                                     (4, 64, 0), (8, 16, 0), (8, 32, 0), (1, 16, 1),
                                     (1, 32, 1), (1, 64, 1), (1, 16, 2), (1, 32, 2)):
                s.set_option(_lib.OPT_TILE_WAVES, waves)
                s.set_option(_lib.OPT_TILE_K, tk)
                s.set_option(_lib.OPT_TILE_STAGE, stage)
                z_tk = s.prec_inv(x)
                assert np.array_equal(z_tk, z_lv), (rep, waves, tk, stage,
                                                    np.max(np.abs(z_tk - z_lv)))
            s.set_option(_lib.OPT_TILE_WAVES, 1)
            s.set_option(_lib.OPT_TILE_K, 32)
            s.set_option(_lib.OPT_TILE_STAGE, 0)


_DATAFLOW_REMOVED = ("the persistent dataflow ILU sweep (P3D_OPT_SWEEP=2: one launch per "
                     "sweep, tiles waiting on each other's progress counters) did not "
                     "terminate inside the CG iteration loop on the MI355X (three runs); "
                     "its cause was not shown, and the schedule has been removed from the "
                     "library")


@pytest.mark.skip(reason=_DATAFLOW_REMOVED)
def test_ilu_dataflow_sweep_single_apply(solver_mod):
    """Single applications of the persistent dataflow sweep against the
    hyperplane schedule (bit for bit)."""
    from pot3d_amd import _lib
    mesh = small_mesh(21, 31, 48)
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        x = np.random.default_rng(8).standard_normal(s.N)
        s.set_option(_lib.OPT_SWEEP, 2)
        z_df = s.prec_inv(x)
        s.set_option(_lib.OPT_SWEEP, 1)
        assert np.array_equal(z_df, s.prec_inv(x))


@pytest.mark.skip(reason=_DATAFLOW_REMOVED)
def test_ilu_solve_same_with_both_sweeps(solver_mod):
    say("[gpu] ilu solve dataflow vs hyperplanes")
    from pot3d_amd import _lib
    mesh = small_mesh(21, 31, 48, drratio=2.5)
    br0 = random_br0(mesh, 9)
    out = []
    for mode in (0, 1):
        with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
            s.set_br0(br0)
            s.set_option(_lib.OPT_SWEEP, mode)
            x = np.zeros(s.N)
            ierr, ncg, epsn = s.solve(x, None, 1e-10, ncgmax=2000)
            assert ierr == 0
            out.append((ncg, x))
    # same preconditioner bit for bit; the dataflow path reduces r.z inside
    # the backward sweep (tile order), so the CG scalars differ in the last
    # bits only
    assert abs(out[0][0] - out[1][0]) <= 1
    assert _rel(out[0][1], out[1][1]) < 1e-10


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_ilu_solve_multiwave_tiles(solver_mod):
    """A whole PCG solve with 16x32-column tiles equals the default tiles."""
    say("[gpu] ilu solve multi-wave tiles")
    from pot3d_amd import _lib
    mesh = small_mesh(33, 45, 70)
    br0 = random_br0(mesh, 5)
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        s.set_br0(br0)
        x1 = np.zeros(s.N)
        assert s.solve(x1, None, 1e-9, ncgmax=2000)[0] == 0
        n1 = s.stats()["ncg"]
        for waves in (4, 8):
            s.set_option(_lib.OPT_TILE_WAVES, waves)
            x2 = np.zeros(s.N)
            assert s.solve(x2, None, 1e-9, ncgmax=2000)[0] == 0
            assert s.stats()["ncg"] == n1 and np.array_equal(x1, x2)


def test_ilu_solve_tiles_vs_hyperplanes(solver_mod):
    """ILU0 PCG with the tile-diagonal schedule and with one launch per
    hyperplane: identical preconditioner bits, so identical iterates."""
    say("[gpu] ilu solve tiles vs hyperplanes")
    from pot3d_amd import _lib
    mesh = small_mesh(21, 31, 48, drratio=2.5)
    br0 = random_br0(mesh, 9)
    out = []
    for mode in (0, 1):
        with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
            s.set_br0(br0)
            s.set_option(_lib.OPT_SWEEP, mode)
            x = np.zeros(s.N)
            ierr, ncg, epsn = s.solve(x, None, 1e-10, ncgmax=2000)
            assert ierr == 0
            out.append((ncg, x))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])


def test_dataflow_sweep_option_rejected(solver_mod):
    """The removed persistent schedule is refused at the option."""
    from pot3d_amd import _lib
    mesh = small_mesh()
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        with pytest.raises(_lib.Pot3dError) as e:
            s.set_option(_lib.OPT_SWEEP, 2)
        assert e.value.code == _lib.P3D_EINVAL


def test_ilu_pivot_guard(solver_mod):
    """A non-finite mesh entry makes the D-ILU pivots non-finite: the
    factorisation refuses them (P3D_ENONFINITE) instead of iterating on
    inf/NaN (the reference's host ilu0 prints a zero pivot and goes on,
    src/pot3d.F90:4695-4697)."""
    say("[gpu] ILU pivot guard")
    from pot3d_amd import _lib
    mesh = small_mesh(13, 11, 18)
    mesh.dph = mesh.dph.copy()
    mesh.dph[5] = np.nan
    with pytest.raises(_lib.Pot3dError) as e:
        solver_mod.Pot3dSolver(mesh, -1.0, 2)
    assert e.value.code == _lib.P3D_ENONFINITE
    assert "pivot" in str(e.value)
    # the same mesh with the diagonal preconditioner is created (no pivots)
    with solver_mod.Pot3dSolver(mesh, -1.0, 1):
        pass


@pytest.mark.parametrize("ifprec,oprec", [(1, 1), (2, 3), (0, 0)])
def test_solve_matches_oracle(oracle_mod, solver_mod, ifprec, oprec):
    say(f"[gpu] solve ifprec={ifprec}")
    mesh = small_mesh(21, 31, 48, drratio=2.5)
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 9)
    b = o.rhs(br0)
    ref = o.cgsolve(b, ifprec=oprec, epscg=1e-9)
    with solver_mod.Pot3dSolver(mesh, -1.0, ifprec) as s:
        s.set_br0(br0)
        x = np.zeros(o.N)
        r = b.copy()
        ierr, ncg, epsn = s.solve(x, r, 1e-9, ncgmax=5000)
        assert ierr == 0
        assert abs(ncg - ref["ncg"]) <= max(2, ref["ncg"] // 50)
        assert epsn < 1e-9
        assert _rel(x, ref["x"]) < 1e-8
        # residual returned in rhs, as cgsolve's r
        assert np.max(np.abs(r - (b - o.ax(x)))) < 1e-8 * np.max(np.abs(b))
        phi = s.fill_phi()
        w = s.magnetic_energy()
        wref = o.energy(o.fill_phi(ref["x"], br0))
        assert np.allclose(w, wref, rtol=1e-7, atol=0)
        st = s.stats()
        assert st["status"] == 1 and st["true_relres"] < 1e-6


@pytest.mark.parametrize("mixed", [1, 2])
def test_two_norm_stop_and_mixed_precision(oracle_mod, solver_mod, mixed):
    say(f"[gpu] mixed precision (mode {mixed})")
    from pot3d_amd import _lib
    mesh = small_mesh(25, 37, 60, drratio=3.0)
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 12)
    b = o.rhs(br0)
    ref = o.cgsolve(b, ifprec=1, epscg=1e-12)
    wref = o.energy(o.fill_phi(ref["x"], br0))
    for ifprec in (1, 2):
        with solver_mod.Pot3dSolver(mesh, -1.0, ifprec) as s:
            s.set_br0(br0)
            s.set_option(_lib.OPT_STOP, 1)
            x = np.zeros(o.N)
            assert s.solve(x, None, 1e-9, ncgmax=5000)[0] == 0
            st = s.stats()
            assert st["true_relres"] <= 1.2e-9
            x64 = x.copy()
            ncg64 = st["ncg"]
            s.set_option(_lib.OPT_MIXED, mixed)
            x = np.zeros(o.N)
            assert s.solve(x, None, 1e-9, ncgmax=5000)[0] == 0
            st = s.stats()
            say(f"  ifprec={ifprec} mixed: outer={st['n_outer']} ncg={st['ncg']} "
                f"(fp64 {ncg64}) relres={st['true_relres']:.3e}")
            assert st["true_relres"] <= 1e-9 and st["n_outer"] >= 1
            assert _rel(x, ref["x"]) < 1e-8 and _rel(x64, ref["x"]) < 1e-8
            s.fill_phi()
            assert np.allclose(s.magnetic_energy(), wref, rtol=1e-7, atol=0)


def test_mixed_reliable_updates_open_boundary(solver_mod):
    """The FP32 recurrence with the FP64 solution (mode 2) on the 'open'
    problem (r in [2.5, 10], |Br| monopole: smooth, badly conditioned):
    its true residual follows the recurrence, so it needs about the FP64
    iteration count and three residual replacements for 1e-9, and lands on
    the FP64 potential and energy."""
    say("[gpu] mixed mode 2 on the open problem")
    from pot3d_amd import _lib
    from p3d_driver import workloads
    w = workloads.pfss(41, 81, 160, seed=3, n_bipoles=16, option="open", r0=2.5, r1=10.0,
                       ifprec=2)
    with solver_mod.Pot3dSolver(w.mesh, w.pm_r1, 2) as s:
        s.set_br0(w.br0)
        s.set_option(_lib.OPT_STOP, 1)
        x64 = np.zeros(s.N)
        assert s.solve(x64, None, 1e-9, ncgmax=5000)[0] == 0
        n64 = s.stats()["ncg"]
        s.fill_phi()
        w64 = s.magnetic_energy()
        s.set_option(_lib.OPT_MIXED, 2)
        # the kernel timings write the FP32 work vectors (the FP32 operator's
        # output in natural order included): the solve must not depend on
        # what they leave there
        for kid in (1, 7, 8, 9, 10):
            s.time_kernel(kid, 1)
        x = np.zeros(s.N)
        assert s.solve(x, None, 1e-9, ncgmax=5000)[0] == 0
        st = s.stats()
        say(f"  fp64 ncg={n64}; mixed ncg={st['ncg']} replacements={st['n_outer']} "
            f"relres={st['true_relres']:.3e}")
        assert st["true_relres"] <= 1e-9
        assert st["ncg"] <= 1.05 * n64 + 10 and st["n_outer"] <= 3
        assert _rel(x, x64) < 1e-8
        s.fill_phi()
        assert np.allclose(s.magnetic_energy(), w64, rtol=1e-7, atol=0)
        # the benchmark's configuration: block-Jacobi cut, 16-plane tiles
        # staged by the marching wave (ragged edges: 39 x 79 x 158 points)
        for key, v in zip((_lib.OPT_ILU_BLOCKS_R, _lib.OPT_ILU_BLOCKS_T, _lib.OPT_ILU_BLOCKS_P),
                          (2, 4, 2)):
            s.set_option(key, v)
        s.set_option(_lib.OPT_TILE_K, 16)
        s.set_option(_lib.OPT_TILE_STAGE, 1)
        x = np.zeros(s.N)
        assert s.solve(x, None, 1e-9, ncgmax=5000)[0] == 0
        st = s.stats()
        say(f"  block-Jacobi 2x4x2, 16-plane tiles: ncg={st['ncg']} "
            f"replacements={st['n_outer']} relres={st['true_relres']:.3e}")
        assert st["true_relres"] <= 1e-9 and st["ncg"] <= 1.3 * n64 + 10
        assert _rel(x, x64) < 1e-8
        s.fill_phi()
        assert np.allclose(s.magnetic_energy(), w64, rtol=1e-7, atol=0)
        # the same FP32 solve with the tiles held in registers: the same
        # arithmetic per point, so the same iterates bit for bit
        s.set_option(_lib.OPT_TILE_STAGE, 2)
        x2 = np.zeros(s.N)
        assert s.solve(x2, None, 1e-9, ncgmax=5000)[0] == 0
        st2 = s.stats()
        say(f"  register tiles: ncg={st2['ncg']} replacements={st2['n_outer']}")
        assert st2["ncg"] == st["ncg"] and st2["n_outer"] == st["n_outer"]
        assert np.array_equal(x2, x)
        # the residual update folded into the register tiles' forward sweep
        # (the default from 2^26 unknowns): the same fmaf, the same iterates
        s.set_option(_lib.OPT_FUSE_RESIDUAL, 1)
        s.time_kernel(11, 1)
        x3 = np.zeros(s.N)
        assert s.solve(x3, None, 1e-9, ncgmax=5000)[0] == 0
        st3 = s.stats()
        say(f"  residual update in the forward sweep: ncg={st3['ncg']}")
        assert st3["ncg"] == st["ncg"] and st3["n_outer"] == st["n_outer"]
        assert np.array_equal(x3, x)
        # the register tiles above ran block-persistent (a workgroup per
        # block marching its diagonals, the default); one launch per
        # diagonal over all blocks gives the same iterates
        s.set_option(_lib.OPT_BLOCK_SWEEP, 0)
        x4 = np.zeros(s.N)
        assert s.solve(x4, None, 1e-9, ncgmax=5000)[0] == 0
        st4 = s.stats()
        say(f"  one launch per tile diagonal: ncg={st4['ncg']}")
        assert st4["ncg"] == st["ncg"] and st4["n_outer"] == st["n_outer"]
        assert np.array_equal(x4, x)
        s.set_option(_lib.OPT_FUSE_RESIDUAL, -1)
        s.set_option(_lib.OPT_BLOCK_SWEEP, -1)


def test_failure_paths(oracle_mod, solver_mod):
    say("[gpu] failure paths")
    mesh = small_mesh()
    with solver_mod.Pot3dSolver(mesh, -1.0, 1) as s:
        s.set_br0(np.zeros((mesh.nt, mesh.np)))
        x = np.ones(s.N)
        ierr, ncg, epsn = s.solve(x, None, 1e-9, ncgmax=100)
        assert ierr == 0 and ncg == 0 and np.all(x == 0)  # zero rhs (:4302-4311)
        br0 = random_br0(mesh, 1)
        s.set_br0(br0)
        x = np.zeros(s.N)
        ierr, ncg, _ = s.solve(x, None, 1e-14, ncgmax=3)
        assert ierr == 1 and ncg == 3
        bad = br0.copy()
        bad[3, 4] = np.nan
        s.set_br0(bad)
        from pot3d_amd._lib import Pot3dError, P3D_ENONFINITE
        with pytest.raises(Pot3dError) as e:
            s.solve(np.zeros(s.N), None, 1e-9, ncgmax=1000)
        assert e.value.code == P3D_ENONFINITE


def test_legacy_lu_abi(oracle_mod, solver_mod):
    """load_external_lu_solver / external_lu_solver / unload with host
    arrays (the -DHIPSPARSE convention without OpenMP mapping): LU factors
    written back over the CSR values, and the trisolve, against the
    oracle's Saad ilu0 + lusol (src/pot3d.F90:4717-4813, 5227-5302)."""
    say("[gpu] legacy ABI")
    import ctypes as C
    from pot3d_amd import _lib
    L = _lib.load()
    mesh = small_mesh(17, 14, 26, drratio=2.0)
    o = oracle_mod.Oracle(mesh, -1.0)
    IA, JA, Av, dptr = o.csr()
    _, _, LUref, _, _ = o.ilu0()
    lu = Av.copy()
    L.load_external_lu_solver(lu.ctypes.data, IA.ctypes.data, JA.ctypes.data, o.N, len(Av))
    nnz = IA[-1] - 1
    assert np.max(np.abs(lu[:nnz] - LUref[:nnz]) / np.abs(LUref[:nnz])) < 1e-4
    x = np.random.default_rng(5).standard_normal(o.N)
    ref = o.lusol(x.astype(np.float32)).astype(np.float64)
    y = x.copy()
    L.external_lu_solver(y.ctypes.data)
    assert np.max(np.abs(y - ref)) <= 1e-4 * np.max(np.abs(ref))
    L.unload_external_lu_solver()


def test_legacy_lu_abi_device_pointers(oracle_mod, solver_mod):
    """The -DCUSPARSE calling convention (src/pot3d.F90:4675-4683,
    use_device_addr): CSR arrays and x are device addresses.  The LU factors
    are written back over the device CSR values, x is solved in place on the
    device, and IA/JA are left as they were; repeated applies give the same
    bits as the host-array convention."""
    say("[gpu] legacy ABI, device pointers")
    from pot3d_amd import _hip, _lib
    H = _hip.Hip()
    L = _lib.load()
    mesh = small_mesh(17, 14, 26, drratio=2.0)
    o = oracle_mod.Oracle(mesh, -1.0)
    IA, JA, Av, dptr = o.csr()
    _, _, LUref, _, _ = o.ilu0()
    nnz = IA[-1] - 1
    x = np.random.default_rng(5).standard_normal(o.N)
    # host convention first, for the bit-for-bit comparison
    lu_h = Av.copy()
    L.load_external_lu_solver(lu_h.ctypes.data, IA.ctypes.data, JA.ctypes.data, o.N, len(Av))
    y_h = x.copy()
    L.external_lu_solver(y_h.ctypes.data)
    L.unload_external_lu_solver()
    bufs = [H.to_device(a) for a in (Av, IA, JA, x)]
    d_av, d_ia, d_ja, d_x = bufs
    try:
        L.load_external_lu_solver(d_av, d_ia, d_ja, o.N, len(Av))
        lu = np.empty_like(Av)
        H.d2h(lu, d_av)
        assert np.max(np.abs(lu[:nnz] - LUref[:nnz]) / np.abs(LUref[:nnz])) < 1e-4
        assert np.array_equal(lu[:nnz], lu_h[:nnz])
        ia2, ja2 = np.empty_like(IA), np.empty_like(JA)
        H.d2h(ia2, d_ia)
        H.d2h(ja2, d_ja)
        assert np.array_equal(ia2, IA) and np.array_equal(ja2, JA)
        ref = o.lusol(x.astype(np.float32)).astype(np.float64)
        y = np.empty_like(x)
        for _ in range(2):
            H.h2d(d_x, x)
            L.external_lu_solver(d_x)
            H.d2h(y, d_x)
            assert np.max(np.abs(y - ref)) <= 1e-4 * np.max(np.abs(ref))
            assert np.array_equal(y, y_h)
        L.unload_external_lu_solver()
    finally:
        for d in bufs:
            H.free(d)


def test_legacy_lu_abi_zero_pivot_exits(oracle_mod, tmp_path):
    """A zero ILU0 pivot ends the program with a message and exit status 1,
    the shim's error convention (the reference's host ilu0 reports zero
    pivots, src/pot3d.F90:4695-4697).  Run in a child process."""
    say("[gpu] legacy ABI zero pivot")
    import subprocess
    mesh = small_mesh(13, 11, 18)
    o = oracle_mod.Oracle(mesh, -1.0)
    IA, JA, Av, dptr = o.csr()
    Av[IA[0] - 1:IA[1] - 1] = 0.0   # row 1: zero diagonal and couplings
    np.savez(tmp_path / "csr.npz", IA=IA, JA=JA, Av=Av)
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {ROOT!r})\n"
        "from pot3d_amd import _lib\n"
        f"z = np.load({repr(str(tmp_path / 'csr.npz'))})\n"
        "IA, JA, Av = z['IA'], z['JA'], z['Av'].copy()\n"
        "L = _lib.load()\n"
        "L.load_external_lu_solver(Av.ctypes.data, IA.ctypes.data, JA.ctypes.data, "
        "len(IA) - 1, len(Av))\n"
        "print('returned', flush=True)\n")
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=120)
    say(res.stdout.strip(), res.stderr.strip()[-300:])
    assert res.returncode == 1
    assert "zero or non-finite pivot" in res.stderr and "row 1" in res.stderr
    assert "returned" not in res.stdout


def test_neumann_r1_boundary_values(oracle_mod, solver_mod):
    """The validation run's Neumann value at r1 (br1, src/pot3d.F90:5420-
    5423; rhs at :3941-3950): rhs, phi with boundary values, energy and B of
    a solve against the oracle.  With Neumann ghosts at both r ends the
    system is singular (constants); br0 and br1 are both flux balanced, so it
    is consistent, and phi is compared after removing its mean."""
    say("[gpu] br1 Neumann r1 values")
    mesh = small_mesh(21, 31, 48, drratio=2.5)
    o = oracle_mod.Oracle(mesh, 1.0)
    br0 = random_br0(mesh, 21)
    br1 = 0.2 * random_br0(mesh, 22)
    b = o.rhs(br0, br1)
    bz = o.rhs(br0)
    assert np.max(np.abs(b - bz)) > 1e-3 * np.max(np.abs(b))  # br1 enters
    ref = o.cgsolve(b, ifprec=3, epscg=1e-12)
    phir = o.fill_phi(ref["x"], br0, br1)
    with solver_mod.Pot3dSolver(mesh, 1.0, 2) as s:
        s.set_br0(br0)
        s.set_br1(br1)
        assert np.max(np.abs(s.rhs() - b)) <= 1e-14 * np.max(np.abs(b))
        xr = np.random.default_rng(3).standard_normal(o.N)
        pf = s.fill_phi(xr)
        pref = o.fill_phi(xr, br0, br1)
        assert np.max(np.abs(pf - pref)) <= 1e-14 * np.max(np.abs(pref))
        x = np.zeros(o.N)
        ierr, ncg, epsn = s.solve(x, None, 1e-12, ncgmax=5000)
        say(f"    ncg {ncg} (oracle {ref['ncg']})")
        assert ierr == 0 and abs(ncg - ref["ncg"]) <= 2
        phi = s.fill_phi()
        dm = (phi - phi.mean()) - (phir - phir.mean())
        assert np.max(np.abs(dm)) <= 1e-8 * np.max(np.abs(phir - phir.mean()))
        for g_, r_ in zip(s.getb(), o.getb(phir)):
            assert np.max(np.abs(g_ - r_)) <= 1e-8 * np.max(np.abs(r_))
        assert np.allclose(s.magnetic_energy(), o.energy(phir), rtol=1e-7, atol=0)
        s.set_br1(None)
        assert np.max(np.abs(s.rhs() - bz)) <= 1e-14 * np.max(np.abs(bz))


def _blk_extents(n, nb, q):
    """The solver's block extents (blocks per axis nb, rounded up to q)."""
    if nb <= 1:
        return n
    e = -(-n // nb)
    e = -(-e // q) * q
    return n if e >= n else e


@pytest.mark.parametrize("nblk", [(2, 2, 2), (1, 2, 1), (2, 1, 3)])
def test_block_jacobi_ilu(oracle_mod, solver_mod, nblk):
    """Block-Jacobi ILU0 (the reference's per-rank ILU on a decomposed
    grid): pivots equal the oracle's, every sweep schedule gives the same
    bits, the apply matches the oracle's, and a solve reaches the same
    answer."""
    say(f"[gpu] block-Jacobi ILU blocks={nblk}")
    from pot3d_amd import _lib
    mesh = small_mesh(33, 45, 70)
    a, b, c = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    ext = (_blk_extents(a, nblk[0], 16), _blk_extents(b, nblk[1], 32),
           _blk_extents(c, nblk[2], 64))
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 6)
    with solver_mod.Pot3dSolver(mesh, -1.0, 2) as s:
        s.set_option(_lib.OPT_ILU_BLOCKS_R, nblk[0])
        s.set_option(_lib.OPT_ILU_BLOCKS_T, nblk[1])
        s.set_option(_lib.OPT_ILU_BLOCKS_P, nblk[2])
        dref = o.dilu_dinv(ext)
        assert np.allclose(s.ilu_dinv(), dref, rtol=2e-7, atol=0)
        x = np.random.default_rng(3).standard_normal(s.N)
        s.set_option(_lib.OPT_SWEEP, 1)
        z_lv = s.prec_inv(x)
        s.set_option(_lib.OPT_SWEEP, 0)
        for waves, tk, stage in ((1, 32, 0), (1, 64, 0), (4, 64, 0), (8, 32, 0), (1, 16, 1),
                                 (1, 32, 1), (1, 16, 2), (1, 32, 2)):
            s.set_option(_lib.OPT_TILE_WAVES, waves)
            s.set_option(_lib.OPT_TILE_K, tk)
            s.set_option(_lib.OPT_TILE_STAGE, stage)
            z = s.prec_inv(x)
            assert np.array_equal(z, z_lv), (waves, tk, stage, np.max(np.abs(z - z_lv)))
            if stage == 2:
                # register tiles, one launch per diagonal instead of a
                # workgroup per block marching its diagonals
                s.set_option(_lib.OPT_BLOCK_SWEEP, 0)
                z = s.prec_inv(x)
                s.set_option(_lib.OPT_BLOCK_SWEEP, -1)
                assert np.array_equal(z, z_lv), ("per diagonal", tk, np.max(np.abs(z - z_lv)))
        zref = o.dilu_apply(x.astype(np.float32), blocks=ext).astype(np.float64)
        assert np.max(np.abs(z_lv - zref)) <= 2e-5 * np.max(np.abs(zref))
        s.set_option(_lib.OPT_TILE_WAVES, 1)
        s.set_option(_lib.OPT_TILE_K, 32)
        s.set_option(_lib.OPT_TILE_STAGE, 0)
        s.set_br0(br0)
        s.set_option(_lib.OPT_STOP, 1)
        xs = np.zeros(s.N)
        assert s.solve(xs, None, 1e-10, ncgmax=5000)[0] == 0
        ncg = s.stats()["ncg"]
        ref = o.cgsolve(o.rhs(br0), ifprec=1, epscg=1e-12)
        say(f"    ncg {ncg}")
        assert _rel(xs, ref["x"]) < 1e-8
