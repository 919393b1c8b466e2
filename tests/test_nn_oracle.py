"""Pins of the network oracle (oracle/nn_oracle.py) on the CPU.

* conv_fwd against a direct loop over voxels and taps (zero padding in r
  and theta, periodic phi), so the shift convention is fixed independently;
* adjoint identities <F x, y> = <x, F^T y> for the convolution, pooling and
  upsampling backward passes;
* the whole-network gradient (mode f64) against central finite differences
  of its own loss, per layer;
* BF16 rounding against torch's round-to-nearest-even conversion.
"""
import numpy as np
import pytest

import nn_oracle as N


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-8, 8, 4000),
                        [1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 1e-40]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(N.bf16(x), ref)
    assert N.bf16(1.0 + 2.0 ** -8) == 1.0           # tie to even
    assert N.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6


def _conv_loop(x, W, b):
    Z, Y, X, _ = x.shape
    out = np.zeros((Z, Y, X, W.shape[1]))
    for z in range(Z):
        for y in range(Y):
            for xx in range(X):
                acc = b.copy()
                for t in range(27):
                    dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
                    xn, yn, zn = xx + dx, y + dy, (z + dz) % Z
                    if 0 <= xn < X and 0 <= yn < Y:
                        acc = acc + W[t] @ x[zn, yn, xn]
                out[z, y, xx] = acc
    return out


@pytest.mark.parametrize("shape", [(3, 4, 5), (1, 2, 3), (2, 1, 1), (5, 3, 2)])
def test_conv_matches_loop(shape):
    rng = np.random.default_rng(1)
    Z, Y, X = shape
    x = rng.standard_normal((Z, Y, X, 4))
    W = rng.standard_normal((27, 3, 4))
    b = rng.standard_normal(3)
    assert np.allclose(N.conv_fwd(x, W, b), _conv_loop(x, W, b), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("fdims", [(7, 5, 6), (4, 4, 4), (1, 3, 5)])
def test_adjoints(fdims):
    rng = np.random.default_rng(2)
    X, Y, Z = fdims
    cdims = ((X + 1) // 2, (Y + 1) // 2, (Z + 1) // 2)
    x = rng.standard_normal((Z, Y, X, 3))
    W = rng.standard_normal((27, 5, 3))
    y = rng.standard_normal((Z, Y, X, 5))
    dx, dW, db = N.conv_bwd(y, x, W)
    assert np.isclose(np.sum(N.conv_fwd(x, W, 0.0) * y), np.sum(x * dx), rtol=1e-12)
    # dW: the convolution is linear in W as well
    assert np.isclose(np.sum(N.conv_fwd(x, W, 0.0) * y), np.sum(W * dW), rtol=1e-12)
    assert np.allclose(db, y.reshape(-1, 5).sum(0))
    yc = rng.standard_normal((cdims[2], cdims[1], cdims[0], 3))
    assert np.isclose(np.sum(N.pool_fwd(x, cdims) * yc), np.sum(x * N.pool_bwd(yc, fdims)),
                      rtol=1e-12)
    assert np.isclose(np.sum(N.up_fwd(yc, fdims) * x), np.sum(yc * N.up_bwd(x, cdims)),
                      rtol=1e-12)
    # pooling of a constant is that constant (child counts at ragged edges)
    assert np.allclose(N.pool_fwd(np.ones_like(x), cdims), 1.0)


def _small_problem(seed=0, dims=(6, 5, 8)):
    a, b, c = dims
    rng = np.random.default_rng(seed)
    p = N.init_params(seed)
    _, n = N.param_offsets()
    p[[bo for _, bo in N.param_offsets()[0]]] = 0.0
    p += 0.01 * rng.standard_normal(n)  # nonzero biases too
    x0 = np.zeros((c, b, a, N.NCH0))
    x0[..., :3] = rng.standard_normal((c, b, a, 3))
    tgt = rng.standard_normal((c, b, a))
    return p, x0, tgt


def test_gradient_matches_finite_differences():
    p, x0, tgt = _small_problem()
    L, g, _ = N.loss_and_grad(p, x0, tgt, mode="f64")
    offs, n = N.param_offsets()
    rng = np.random.default_rng(5)
    h = 1e-6
    for i, ((cin, cout, _), (wo, bo)) in enumerate(zip(N.SPEC, offs)):
        d = np.zeros(n)
        d[wo:bo + cout] = rng.standard_normal(bo + cout - wo)
        lp = N.loss_and_grad(p + h * d, x0, tgt, mode="f64")[0]
        lm = N.loss_and_grad(p - h * d, x0, tgt, mode="f64")[0]
        fd = (lp - lm) / (2 * h)
        an = float(g @ d)
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12) + 1e-10, (i, fd, an)


def test_bf16_mode_close_to_f64():
    # the device's rounding points change the loss and gradient by O(2^-8)
    p, x0, tgt = _small_problem(seed=3)
    L64, g64, _ = N.loss_and_grad(p, x0, tgt, mode="f64")
    Lbf, gbf, _ = N.loss_and_grad(p, x0, N.bf16(tgt), mode="bf16")
    assert abs(Lbf - L64) <= 0.05 * L64
    assert np.linalg.norm(gbf - g64) <= 0.1 * np.linalg.norm(g64)


def test_features_layout():
    rh = np.linspace(0.9, 2.6, 9)
    nt, np_ = 7, 10
    br0 = np.arange(nt * np_, dtype=float).reshape(nt, np_)
    x = N.features(br0, rh, bscale=2.0, mode="f64")
    a, b, c = 7, 5, 8
    assert x.shape == (c, b, a, N.NCH0)
    # voxel (i, j, k) reads br0 at (j+1, k+1)
    assert x[3, 2, 4, 1] == br0[3, 4] / 2.0
    dec, rpos = N.radial_tables(rh)
    assert np.isclose(x[3, 2, 4, 0], br0[3, 4] / 2.0 * dec[4])
    assert np.isclose(rpos[0], (rh[1] - 0.5 * (rh[0] + rh[1])) / (0.5 * (rh[-2] + rh[-1])
                                                                  - 0.5 * (rh[0] + rh[1])))
    assert np.all(x[..., 3:] == 0)
