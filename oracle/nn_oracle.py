"""CPU oracle of the warm-start network (numpy) -- test infrastructure only.

Only tests/ may import this module; the product path never does.  The
reference (predsci/POT3D) has no learned component: its initial guess is
zero (src/pot3d.F90:3922-3925).  This module restates the network that
libpot3d_hip.so trains and evaluates (include/pot3d_hip.h, p3d_nn_*), so
that its forward pass, loss and parameter gradient can be checked.

Two precisions:
  * mode "f64": every operation in float64, no rounding.  Its gradient is
    pinned by finite differences of its own loss (tests/test_nn_oracle.py),
    and its pieces by adjoint identities (<A x, y> = <x, A^T y>).
  * mode "bf16": the same code with the device's rounding points: BF16
    weights, BF16 activations after each convolution / pooling / skip sum,
    BF16 rounding of the FP32 gradient entering each convolution backward,
    FP32 head output and FP32 input features.  The device result is
    compared with this mode.

Layout: tensors are (Z, Y, X, C) arrays -- flattened they are the device's
voxel-major, channel-fastest T[(x + X*(y + Y*z))*C + c]; x = r, y = theta,
z = phi (periodic), zero padding in x and y.
"""
from __future__ import annotations

import numpy as np

# (Cin, Cout, level) of the 8 convolutions, in parameter order
SPEC = [(16, 16, 0), (16, 32, 1), (32, 64, 2), (64, 128, 3),
        (128, 64, 2), (64, 32, 1), (32, 16, 0), (16, 16, 0)]
NCH0 = 16  # channels of the input features (3 used, rest zero)


def param_offsets():
    """[(w_off, b_off)] per convolution and the total parameter count; the
    weights of a convolution are W[tap][c_out][c_in], tap = (dx+1) +
    3(dy+1) + 9(dz+1)."""
    offs, off = [], 0
    for cin, cout, _ in SPEC:
        w_off = off
        off += 27 * cin * cout
        offs.append((w_off, off))
        off += cout
    return offs, off


def unpack_params(p):
    offs, n = param_offsets()
    assert p.size == n
    out = []
    for (cin, cout, _), (wo, bo) in zip(SPEC, offs):
        out.append((p[wo:wo + 27 * cin * cout].reshape(27, cout, cin), p[bo:bo + cout]))
    return out


def level_dims(a, b, c):
    D = [(a, b, c)]
    for _ in range(3):
        X, Y, Z = D[-1]
        D.append(((X + 1) // 2, (Y + 1) // 2, (Z + 1) // 2))
    return D


# -- BF16 rounding (round to nearest even, as the device's f2bf) -------------
def bf16(x):
    x32 = np.asarray(x, dtype=np.float32)
    u = x32.view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(special, (u >> 16) << 16, r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


class _Round:
    def __init__(self, mode):
        assert mode in ("f64", "bf16")
        self.on = mode == "bf16"

    def act(self, x):  # BF16 activation storage
        return bf16(x) if self.on else x

    def f32(self, x):
        return np.asarray(x, np.float32).astype(np.float64) if self.on else x


# -- building blocks ----------------------------------------------------------
def tap_offsets():
    return [((t % 3) - 1, ((t // 3) % 3) - 1, (t // 9) - 1) for t in range(27)]


def shift(x, dx, dy, dz):
    """out[z,y,x] = x[z+dz, y+dy, x+dx]; zero outside in x and y, periodic z."""
    Z, Y, X, C = x.shape
    out = np.zeros_like(x)
    xs = np.roll(x, -dz, axis=0)
    ys0, ys1 = max(0, -dy), min(Y, Y - dy)
    xs0, xs1 = max(0, -dx), min(X, X - dx)
    out[:, ys0:ys1, xs0:xs1] = xs[:, ys0 + dy:ys1 + dy, xs0 + dx:xs1 + dx]
    return out


def conv_fwd(x, W, b):
    """y[v][co] = sum_t sum_ci x[v + o_t][ci] W[t][co][ci] + b[co]."""
    Z, Y, X, _ = x.shape
    y = np.zeros((Z, Y, X, W.shape[1]))
    for t, (dx, dy, dz) in enumerate(tap_offsets()):
        y += shift(x, dx, dy, dz) @ W[t].T
    return y + b


def conv_bwd(dy, x, W):
    """(dx, dW, db) of conv_fwd for the upstream gradient dy."""
    dx = np.zeros_like(x)
    dW = np.zeros_like(W)
    dyf = dy.reshape(-1, dy.shape[-1])
    for t, (ox, oy, oz) in enumerate(tap_offsets()):
        dx += shift(dy @ W[t], -ox, -oy, -oz)
        dW[t] = dyf.T @ shift(x, ox, oy, oz).reshape(-1, x.shape[-1])
    db = dyf.sum(axis=0)
    return dx, dW, db


def _counts(fdims, cdims):
    """children per coarse voxel (Z, Y, X, 1) of 2x2x2 pooling with ceil."""
    X, Y, Z = fdims
    cX, cY, cZ = cdims
    nx = np.minimum(2, X - 2 * np.arange(cX))
    ny = np.minimum(2, Y - 2 * np.arange(cY))
    nz = np.minimum(2, Z - 2 * np.arange(cZ))
    return (nz[:, None, None] * ny[None, :, None] * nx[None, None, :])[..., None].astype(float)


def _up_index(fdims):
    X, Y, Z = fdims
    return np.arange(Z) // 2, np.arange(Y) // 2, np.arange(X) // 2


def up_fwd(x, fdims):
    iz, iy, ix = _up_index(fdims)
    return x[iz][:, iy][:, :, ix]


def up_bwd(dy, cdims):
    """adjoint of up_fwd: sum over the children of each coarse voxel."""
    cX, cY, cZ = cdims
    Z, Y, X, C = dy.shape
    out = np.zeros((cZ, cY, cX, C))
    iz, iy, ix = np.arange(Z) // 2, np.arange(Y) // 2, np.arange(X) // 2
    np.add.at(out, (iz[:, None, None], iy[None, :, None], ix[None, None, :]), dy)
    return out


def pool_fwd(x, cdims):
    fdims = (x.shape[2], x.shape[1], x.shape[0])
    return up_bwd(x, cdims) / _counts(fdims, cdims)


def pool_bwd(dy, fdims):
    cdims = (dy.shape[2], dy.shape[1], dy.shape[0])
    return up_fwd(dy / _counts(fdims, cdims), fdims)


def relu(x):
    return np.maximum(x, 0.0)


# -- features and the network -------------------------------------------------
def radial_tables(rh):
    """decay (r0/r)^2 and rpos (r-r0)/(r1-r0) on the interior shells
    (float32, as the device tables)."""
    nr = rh.size
    r0 = 0.5 * (rh[0] + rh[1])
    r1 = 0.5 * (rh[nr - 2] + rh[nr - 1])
    r = rh[1:nr - 1]
    return (((r0 / r) * (r0 / r)).astype(np.float32), ((r - r0) / (r1 - r0)).astype(np.float32))


def features(br0, rh, bscale, mode="bf16"):
    """Input (Z, Y, X, 16) from br0(nt,np) (C array of the Fortran (j,k)):
    ch0 = br0 (r0/r)^2 / bscale, ch1 = br0 / bscale, ch2 = rpos."""
    rd = _Round(mode)
    dec, rpos = radial_tables(rh)
    a = rh.size - 2
    bint = np.asarray(br0)[1:-1, 1:-1]  # (b, c): j interior, k interior
    if rd.on:
        bz = (bint.astype(np.float32) * np.float32(1.0 / bscale)).T  # (c, b) = (Z, Y)
        ch0 = bz[:, :, None] * dec[None, None, :]
        ch1 = np.broadcast_to(bz[:, :, None], ch0.shape)
        ch2 = np.broadcast_to(rpos[None, None, :], ch0.shape)
    else:
        bz = (bint / bscale).T
        ch0 = bz[:, :, None] * dec.astype(float)[None, None, :]
        ch1 = np.broadcast_to(bz[:, :, None], ch0.shape)
        ch2 = np.broadcast_to(rpos.astype(float)[None, None, :], ch0.shape)
    x = np.zeros(ch0.shape + (NCH0,))
    x[..., 0] = rd.act(ch0.astype(np.float64))
    x[..., 1] = rd.act(np.asarray(ch1, np.float64))
    x[..., 2] = rd.act(np.asarray(ch2, np.float64))
    assert x.shape[2] == a
    return x


def forward(params, x0, mode="bf16"):
    """Forward pass; returns (out (Z,Y,X,16), cache)."""
    rd = _Round(mode)
    P = [(rd.act(W), rd.f32(b)) for W, b in unpack_params(params)]
    D = level_dims(x0.shape[2], x0.shape[1], x0.shape[0])
    c = {"x0": x0}

    def conv(i, x, act=True):
        y = conv_fwd(x, *P[i])
        if act:
            return rd.act(relu(y))
        return rd.f32(y)

    c["e0"] = conv(0, x0)
    c["p1"] = rd.act(pool_fwd(c["e0"], D[1]))
    c["e1"] = conv(1, c["p1"])
    c["p2"] = rd.act(pool_fwd(c["e1"], D[2]))
    c["e2"] = conv(2, c["p2"])
    c["p3"] = rd.act(pool_fwd(c["e2"], D[3]))
    c["e3"] = conv(3, c["p3"])
    c["u2"] = up_fwd(c["e3"], D[2])
    c["d2"] = conv(4, c["u2"])
    c["s2"] = rd.act(c["d2"] + c["e2"])
    c["u1"] = up_fwd(c["s2"], D[1])
    c["d1"] = conv(5, c["u1"])
    c["s1"] = rd.act(c["d1"] + c["e1"])
    c["u0"] = up_fwd(c["s1"], D[0])
    c["d0"] = conv(6, c["u0"])
    c["s0"] = rd.act(c["d0"] + c["e0"])
    out = conv(7, c["s0"], act=False)
    c["P"] = P
    c["D"] = D
    return out, c


def loss_and_grad(params, x0, target, mode="bf16"):
    """L = mean_v (out[v][0] - target[v])^2 and dL/dparams (flat, as the
    device's parameter buffer).  target is (Z, Y, X)."""
    rd = _Round(mode)
    out, c = forward(params, x0, mode)
    P, D = c["P"], c["D"]
    V = target.size
    dlt = out[..., 0] - target
    L = float(np.sum(dlt * dlt) / V)
    g_out = np.zeros_like(out)
    g_out[..., 0] = rd.f32(2.0 * dlt / V)
    grads = [None] * 8

    def cbwd(i, dy, x, need_dx=True):
        dx, dW, db = conv_bwd(rd.act(dy), x, P[i][0])
        # the bias gradient sums the unrounded FP32 gradient
        db = dy.reshape(-1, dy.shape[-1]).sum(axis=0)
        grads[i] = (dW, db)
        return dx if need_dx else None

    def mask(y, g):
        return np.where(y > 0, g, 0.0)

    g_s0 = cbwd(7, g_out, c["s0"])
    g_t0 = mask(c["d0"], g_s0)
    g_u0 = cbwd(6, g_t0, c["u0"])
    g_s1 = up_bwd(g_u0, D[1])
    g_t1 = mask(c["d1"], g_s1)
    g_u1 = cbwd(5, g_t1, c["u1"])
    g_s2 = up_bwd(g_u1, D[2])
    g_t2 = mask(c["d2"], g_s2)
    g_u2 = cbwd(4, g_t2, c["u2"])
    g_e3 = mask(c["e3"], up_bwd(g_u2, D[3]))
    g_p3 = cbwd(3, g_e3, c["p3"])
    g_e2 = mask(c["e2"], pool_bwd(g_p3, D[2]) + g_s2)
    g_p2 = cbwd(2, g_e2, c["p2"])
    g_e1 = mask(c["e1"], pool_bwd(g_p2, D[1]) + g_s1)
    g_p1 = cbwd(1, g_e1, c["p1"])
    g_e0 = mask(c["e0"], pool_bwd(g_p1, D[0]) + g_s0)
    cbwd(0, g_e0, c["x0"], need_dx=False)
    offs, n = param_offsets()
    flat = np.zeros(n)
    for (wo, bo), (dW, db) in zip(offs, grads):
        flat[wo:wo + dW.size] = dW.ravel()
        flat[bo:bo + db.size] = db
    return L, flat, out


def init_params(seed=0):
    """He-normal weights (seeded numpy generator), zero biases; the head
    scaled by 0.1.  (The device draws its own initial weights; tests that
    compare with it set these with p3d_nn_set_params.)"""
    rng = np.random.default_rng(seed)
    offs, n = param_offsets()
    p = np.zeros(n)
    for i, ((cin, cout, _), (wo, bo)) in enumerate(zip(SPEC, offs)):
        sd = np.sqrt(2.0 / (27 * cin)) * (0.1 if i == 7 else 1.0)
        p[wo:wo + 27 * cin * cout] = rng.standard_normal(27 * cin * cout) * sd
    return p


def packed_to_zyx(x, a, b, c):
    """packed x_cg (r fastest) -> (Z=c, Y=b, X=a)."""
    return np.asarray(x).reshape(c, b, a)
