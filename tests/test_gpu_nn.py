"""GPU tests of the warm-start network (BF16 matrix cores) against the CPU
oracle (oracle/nn_oracle.py), and of what the trained network gains.

* the operand layout of the 16x16x16 BF16 MFMA every convolution uses,
  against a host matmul of the same BF16 values;
* forward pass, loss and parameter gradient against the oracle's bf16 mode
  (the oracle's own gradient is pinned by finite differences on the CPU);
* trilinear prolongation: exact on multilinear fields;
* training lowers the loss; the trained network's guess cuts PCG
  iterations on held-out boundary maps, with the same converged answer;
  a bad guess is replaced by zero (P3D_OPT_GUESS_CHECK).
"""
import sys

import numpy as np
import pytest

from conftest import random_br0, small_mesh

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_bits(x):
    """Round-to-nearest-even FP32 -> BF16 bits (numpy)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def test_mfma_layout():
    import ctypes as C

    from pot3d_amd import _lib
    say("[gpu] mfma 16x16x16 bf16 layout")
    L = _lib.load()
    rng = np.random.default_rng(3)
    A = bf16_bits(rng.standard_normal((16, 16)))
    B = bf16_bits(rng.standard_normal((16, 16)))
    D = np.zeros((16, 16), np.float32)
    _lib.check(L.p3d_nn_mfma_check(A.ctypes.data_as(C.POINTER(C.c_ushort)),
                                   B.ctypes.data_as(C.POINTER(C.c_ushort)),
                                   D.ctypes.data_as(C.POINTER(C.c_float))))
    ref = bf16_value(A).astype(np.float64) @ bf16_value(B).astype(np.float64)
    assert np.max(np.abs(D - ref)) <= 1e-5 * np.max(np.abs(ref))


def _oracle_params(n, seed=0):
    import nn_oracle as N
    p = N.init_params(seed)
    rng = np.random.default_rng(seed + 1)
    offs, _ = N.param_offsets()
    for (cin, cout, _), (wo, bo) in zip(N.SPEC, offs):
        p[bo:bo + cout] = 0.05 * rng.standard_normal(cout)
    assert p.size == n
    return p.astype(np.float32)


@pytest.mark.parametrize("dims", [(12, 14, 20), (9, 11, 13)])
def test_forward_and_gradient_match_oracle(dims):
    import nn_oracle as N

    from pot3d_amd.warmstart import WarmStartNet
    nr, nt, np_ = dims
    say(f"[gpu] nn forward/backward vs oracle {dims}")
    mesh = small_mesh(nr=nr, nt=nt, np_=np_)
    net = WarmStartNet(mesh, seed=1)
    p = _oracle_params(net.nparams)
    net.set_params(p)
    bscale, phiscale = 2.0, 0.5
    net.set_scales(bscale, phiscale)
    br0 = random_br0(mesh, seed=4)
    a, b, c = nr - 2, nt - 2, np_ - 2
    rng = np.random.default_rng(7)
    x = rng.standard_normal(a * b * c) * 0.3
    x0 = N.features(br0, np.asarray(mesh.rh), bscale, mode="bf16")
    tgt = np.asarray(x / phiscale, np.float32).astype(np.float64).reshape(c, b, a)
    L_o, g_o, out_o = N.loss_and_grad(p.astype(np.float64), x0, tgt, mode="bf16")
    # forward
    xp = net.predict(br0)
    ref = phiscale * out_o[..., 0].ravel()
    err = np.max(np.abs(xp - ref)) / np.max(np.abs(ref))
    say(f"    forward max rel err {err:.2e}")
    assert err < 2e-2
    assert np.linalg.norm(xp - ref) <= 5e-3 * np.linalg.norm(ref)
    # loss and gradient
    L_g, g_g = net.loss_grad(br0, x)
    say(f"    loss gpu {L_g:.6e} oracle {L_o:.6e}")
    assert abs(L_g - L_o) <= 5e-3 * L_o
    offs, _ = N.param_offsets()
    for i, ((cin, cout, _), (wo, bo)) in enumerate(zip(N.SPEC, offs)):
        gw, ow = g_g[wo:bo], g_o[wo:bo]
        gb, ob = g_g[bo:bo + cout], g_o[bo:bo + cout]
        ew = np.linalg.norm(gw - ow) / np.linalg.norm(ow)
        eb = np.linalg.norm(gb - ob) / max(np.linalg.norm(ob), 1e-30)
        say(f"    conv {i} ({cin}->{cout}): weight grad rel err {ew:.2e}, bias {eb:.2e}")
        assert ew < 3e-2 and eb < 3e-2


def test_prolongation_exact_for_multilinear():
    from pot3d_amd import warmstart as W
    from p3d_driver.inputs import global_mesh
    say("[gpu] trilinear prolongation")
    cm = global_mesh(9, 10, 14, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(2.0,))
    fm = global_mesh(17, 19, 27, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(2.0,))

    def f(r, t, p):
        return 1.0 + 2.0 * r - 0.5 * t + 0.25 * p + 0.3 * r * t - 0.1 * t * p + 0.05 * r * t * p

    R, T, P = np.meshgrid(cm.rh, cm.th, cm.ph, indexing="ij")
    phic = f(R, T, P).ravel(order="F")
    tab = W.prolong_tables(cm, fm)
    xf = W.prolong(tab, phic)
    R, T, P = np.meshgrid(fm.rh[1:-1], fm.th[1:-1], fm.ph[1:-1], indexing="ij")
    ref = f(R, T, P).ravel(order="F")
    assert np.max(np.abs(xf - ref)) <= 1e-12 * np.max(np.abs(ref))


def _pairs(wl_fn, seeds, solver):
    from pot3d_amd import warmstart as W
    brs = [wl_fn(s).br0 for s in seeds]
    xs, iters = W.generate_pairs(solver, brs)
    return brs, xs, iters


def test_training_lowers_loss_and_warm_start_gains():
    from pot3d_amd import _lib
    from pot3d_amd import warmstart as W
    from p3d_driver import workloads
    from pot3d_amd.solver import Pot3dSolver
    say("[gpu] train on PCG pairs, warm start on held-out maps")

    def wl(seed):
        return workloads.pfss(21, 37, 72, seed, ifprec=2)

    w0 = wl(0)
    s = Pot3dSolver(w0.mesh, w0.pm_r1, 2)
    s.set_option(_lib.OPT_STOP, 1.0)
    brs, xs, iters = _pairs(wl, range(100, 148), s)
    say(f"    48 training pairs, {iters} PCG iterations")
    net = W.WarmStartNet(w0.mesh, seed=0)
    losses, t = W.train_network(net, brs, xs, epochs=24, lr=2e-3, seed=1)
    first, last = float(np.mean(losses[:20])), float(np.mean(losses[-20:]))
    say(f"    loss {first:.3e} -> {last:.3e} in {t:.1f} s ({losses.size} steps)")
    assert last < 0.2 * first
    # held-out maps: iterations from zero vs from the network's guess
    s.set_option(_lib.OPT_GUESS_CHECK, 1.0)
    zero_it, warm_it = [], []
    for seed in (7, 8, 9, 10):
        b = wl(seed).br0
        s.set_br0(b)
        xz = np.zeros(s.N)
        assert s.solve(xz, None, 1e-9)[0] == 0
        zero_it.append(s.stats()["ncg"])
        xw = net.predict(b)
        assert s.solve(xw, None, 1e-9)[0] == 0
        st = s.stats()
        warm_it.append(st["ncg"])
        say(f"    seed {seed}: zero {zero_it[-1]} warm {warm_it[-1]} (guess {st['guess']})")
        say(f"      guess rel err {np.linalg.norm(net.predict(b) - xz) / np.linalg.norm(xz):.3e}")
        assert st["guess"] == 1
        assert st["true_relres"] <= 1.5e-9
        assert np.linalg.norm(xw - xz) <= 1e-7 * np.linalg.norm(xz)
    assert all(w <= z for w, z in zip(warm_it, zero_it))
    assert np.mean(warm_it) < 0.95 * np.mean(zero_it)
    # a guess worse than zero is replaced by zero, under both rules; the
    # network's guess is kept under the A-norm rule too
    for mode in (1, 2):
        s.set_option(_lib.OPT_GUESS_CHECK, mode)
        xb = 1e3 * np.random.default_rng(0).standard_normal(s.N)
        assert s.solve(xb, None, 1e-9)[0] == 0
        st = s.stats()
        assert st["guess"] == 2 and st["ncg"] == zero_it[-1]
        xw = net.predict(b)
        assert s.solve(xw, None, 1e-9)[0] == 0
        assert s.stats()["guess"] == 1
    net.close()
    s.close()


def test_monopole_split():
    """The monopole split of the warm start: x_m (the PCG solution for
    br0 = 1) depends on r only; a constant added to br0 moves the prediction
    by exactly that constant times x_m (the network sees the same map); and
    on a grid small enough for the network's receptive field to span the
    sphere (where the monopole is learnable without the split) the split
    costs nothing in guess quality ('open' |Br| maps)."""
    from pot3d_amd import _lib
    from pot3d_amd import warmstart as W
    from p3d_driver import workloads
    from pot3d_amd.solver import Pot3dSolver
    say("[gpu] monopole split of the warm start")

    def wl(seed):
        return workloads.pfss(21, 37, 72, seed, ifprec=2, n_bipoles=8, option="open", r0=2.5,
                              r1=10.0)

    w0 = wl(0)
    mesh = w0.mesh
    xm = W.monopole_response(mesh, w0.pm_r1)
    X = xm.reshape(mesh.np - 2, mesh.nt - 2, mesh.nr - 2)
    spread = float(np.max(np.abs(X - X[:1, :1, :])) / np.max(np.abs(X)))
    say(f"    x_m: max deviation from its first column {spread:.2e}")
    assert spread <= 1e-6
    s = Pot3dSolver(mesh, w0.pm_r1, 2)
    s.set_option(_lib.OPT_STOP, 1.0)
    brs, xs, _ = _pairs(wl, range(100, 148), s)
    means = [float(np.sum(W.area_weights(mesh) * b[1:-1, 1:-1]) / np.sum(W.area_weights(mesh)))
             for b in brs]
    say(f"    mean |Br| of the maps: {np.mean(means):.3e} (rms of the maps "
        f"{np.mean([np.sqrt(np.mean(b ** 2)) for b in brs]):.3e})")
    errs = {}
    for mono in (0, 1):
        net = W.WarmStartNet(mesh, seed=0)
        if mono:
            net.set_monopole(xm, W.area_weights(mesh))
        W.train_network(net, brs, xs, epochs=24, lr=2e-3, seed=1)
        e = []
        for seed in (7, 8, 9, 10):
            w = wl(seed)
            s.set_br0(w.br0)
            xz = np.zeros(s.N)
            assert s.solve(xz, None, 1e-9)[0] == 0
            e.append(float(np.linalg.norm(net.predict(w.br0) - xz) / np.linalg.norm(xz)))
        errs[mono] = float(np.mean(e))
        if mono:
            b = wl(7).br0
            d = net.predict(b + 0.5) - net.predict(b)
            lin = float(np.max(np.abs(d - 0.5 * xm)) / np.max(np.abs(0.5 * xm)))
            say(f"    prediction(br0 + 0.5) - prediction(br0) vs 0.5 x_m: {lin:.2e}")
            assert lin <= 1e-3
        net.close()
    say(f"    held-out guess rel err without / with the split: {errs[0]:.3f} / {errs[1]:.3f}")
    assert errs[1] <= 1.1 * errs[0] + 0.01
    s.close()
