"""Pins the oracle's operator (coefficients, boundary values, seam, delsq,
rhs) against closed forms, invariants and an independent sparse assembly.

Reference: load_matrix_pot3d_solve src/pot3d.F90:4565-4613, ax :5102-5150,
set_boundary_points :5376-5487, seam_hhh :5689-5760, rhs trick :3941-3950.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import random_br0, small_mesh


def _assemble(o, mesh, pm_r1):
    """Independent sparse assembly of the homogeneous operator from the
    coefficient planes: ghost values are linear in the interior unknowns
    (r0: x(1)=x(2); r1: x(nr)=pm_r1 x(nrm1); poles: 2<x(i,2,.)>_phi - x;
    phi: periodic), so each ghost read becomes matrix entries."""
    a = o.coefs()
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    N = A_ * B_ * C_
    idx = np.arange(N).reshape(C_, B_, A_)  # [k, j, i]
    w = mesh.dph[1:-1] / (2.0 * math.pi)
    rows, cols, vals = [], [], []

    def add(r, c, v):
        rows.append(r.ravel())
        cols.append(c.ravel())
        vals.append(v.ravel())

    coef = {m: a[m].reshape(C_, B_, A_) for m in range(7)}
    K, J, I = np.meshgrid(np.arange(C_), np.arange(B_), np.arange(A_), indexing="ij")
    P = idx
    add(P, P, coef[3])
    add(P, idx[(K - 1) % C_, J, I], coef[0])  # phi wrap
    add(P, idx[(K + 1) % C_, J, I], coef[6])
    # r neighbours
    inner = I > 0
    add(P[inner], idx[K, J, I - 1][inner], coef[2][inner])
    add(P[~inner], P[~inner], coef[2][~inner])
    outer = I < A_ - 1
    add(P[outer], idx[K, J, np.minimum(I + 1, A_ - 1)][outer], coef[4][outer])
    add(P[~outer], P[~outer], pm_r1 * coef[4][~outer])
    # theta neighbours; pole ghosts 2*sum_k x(i,pole,k) w_k - x
    jin = J > 0
    add(P[jin], idx[K, J - 1, I][jin], coef[1][jin])
    jout = J < B_ - 1
    add(P[jout], idx[K, np.minimum(J + 1, B_ - 1), I][jout], coef[5][jout])
    for pole, jj, cm in ((0, 0, 1), (1, B_ - 1, 5)):
        for k in range(C_):
            for i in range(A_):
                p = idx[k, jj, i]
                cv = coef[cm][k, jj, i]
                rows.append(np.full(C_, p))
                cols.append(idx[:, jj, i])
                vals.append(2.0 * cv * w)
                rows.append(np.array([p]))
                cols.append(np.array([p]))
                vals.append(np.array([-cv]))
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(N, N))


@pytest.mark.parametrize("pm_r1", [-1.0, 1.0])
def test_ax_matches_sparse_assembly(oracle_mod, pm_r1):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, pm_r1)
    A = _assemble(o, mesh, pm_r1)
    rng = np.random.default_rng(1)
    for _ in range(3):
        x = rng.standard_normal(o.N)
        y = o.ax(x)
        ref = A @ x
        scale = np.abs(o.coefs()).sum(axis=0) * np.abs(x).max()
        assert np.max(np.abs(y - ref) / scale) < 1e-14


def test_coefficients_symmetric_and_row_sums(oracle_mod):
    mesh = small_mesh(11, 9, 14)
    o = oracle_mod.Oracle(mesh, -1.0)
    a = o.coefs()
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    c = {m: a[m].reshape(C_, B_, A_) for m in range(7)}
    # a5(i) == a3(i+1), a6(j) == a2(j+1), a7(k) == a1(k+1) (bitwise)
    assert np.array_equal(c[4][:, :, :-1], c[2][:, :, 1:])
    assert np.array_equal(c[5][:, :-1, :], c[1][:, 1:, :])
    assert np.array_equal(c[6][:-1], c[0][1:])
    # periodic: a7 at the last phi plane couples to the first with a1's value
    assert np.array_equal(c[6][-1], c[0][0])
    # off-diagonals negative, diagonal = -(sum of off-diagonals)
    off = np.stack([c[m] for m in (0, 1, 2, 4, 5, 6)])
    assert np.all(off[[0, 2, 3, 5]] < 0)
    assert np.allclose(c[3], -off.sum(axis=0), rtol=1e-15, atol=0)
    # the north-pole theta face carries sin(0) = 0 exactly
    assert np.all(c[1][:, 0, :] == 0.0)


def test_constant_is_in_neumann_nullspace(oracle_mod):
    """pm_r1=+1 ('potential'): every ghost of a constant field equals the
    constant, so A*1 = 0 up to rounding; with pm_r1=-1 only the outer shell
    feels the Dirichlet ghost: row sum a5*(pm_r1-1) = -2*a5."""
    mesh = small_mesh(10, 9, 13)
    o = oracle_mod.Oracle(mesh, 1.0)
    y = o.ax(np.ones(o.N))
    scale = np.abs(o.coefs()[3]).max()
    assert np.max(np.abs(y)) < 1e-13 * scale
    o2 = oracle_mod.Oracle(mesh, -1.0)
    y2 = o2.ax(np.ones(o2.N))
    A_ = mesh.nr - 2
    a5 = o2.coefs()[4].reshape(-1, A_)
    expect = np.zeros_like(a5)
    expect[:, -1] = -2.0 * a5[:, -1]
    assert np.allclose(y2.reshape(-1, A_), expect, rtol=1e-12, atol=1e-13 * scale)


def test_symmetric_operator(oracle_mod):
    mesh = small_mesh(12, 10, 16)
    o = oracle_mod.Oracle(mesh, -1.0)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(o.N), rng.standard_normal(o.N)
    lhs, rhs = np.dot(o.ax(x), y), np.dot(x, o.ax(y))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def _harmonic_residual(oracle_mod, nr, nt, np_):
    """max |delsq(phi)| / dV for phi = r cos(theta) (harmonic), ghosts from
    the exact field: a consistent 2nd-order finite-volume Laplacian gives
    O(h^2)."""
    from p3d_driver import inputs
    mesh = inputs.global_mesh(nr, nt, np_, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(1.0,))
    o = oracle_mod.Oracle(mesh, -1.0)
    R, T, _ = np.meshgrid(mesh.rh, mesh.th, mesh.ph, indexing="ij")
    phi = (R * np.cos(T)).ravel(order="F")
    y = np.empty(o.N)
    oracle_mod.lib().orc_delsq(o.gp, oracle_mod._d(o.coefs().ravel()), oracle_mod._d(phi),
                               oracle_mod._d(y))
    dV = (mesh.rh[1:-1, None, None] ** 2 * mesh.drh[1:-1, None, None]
          * (mesh.sth[1:-1] * mesh.dth[1:-1])[None, :, None] * mesh.dph[None, None, 1:-1])
    Y = y.reshape(np_ - 2, nt - 2, nr - 2).transpose(2, 1, 0)
    # away from the poles, where the pole rows use the axis treatment
    return np.max(np.abs(Y / dV)[:, 2:-2, :])


def test_second_order_consistency(oracle_mod):
    e1 = _harmonic_residual(oracle_mod, 12, 22, 16)
    e2 = _harmonic_residual(oracle_mod, 22, 42, 30)
    assert e2 < e1 / 3.0


def test_rhs_closed_form(oracle_mod):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 5)
    b = o.rhs(br0)
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    Bm = b.reshape(C_, B_, A_)
    # b(2,j,k) = -(a3 * x(1)), x(1) = -br0*dr1, a3 = -sth dth dph r(1)^2/dr(1)
    sth, dth, dph = mesh.sth[1:-1], mesh.dth[1:-1], mesh.dph[1:-1]
    expect = (-(sth[None, :] * dth[None, :] * dph[:, None]) * mesh.r[0] ** 2 / mesh.dr[0]
              * br0[1:-1, 1:-1].T * mesh.dr[0])
    assert np.allclose(Bm[:, :, 0], expect, rtol=1e-13, atol=0)
    assert np.all(Bm[:, :, 1:] == 0.0)


def test_rhs_with_neumann_r1(oracle_mod):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, 1.0)
    br0 = random_br0(mesh, 6)
    br1 = random_br0(mesh, 7)
    b = o.rhs(br0, br1)
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    Bm = b.reshape(C_, B_, A_)
    a5 = o.coefs()[4].reshape(C_, B_, A_)[:, :, -1]
    drn = mesh.dr[mesh.nr - 2]
    assert np.allclose(Bm[:, :, -1], -(a5 * br1[1:-1, 1:-1].T * drn), rtol=1e-14, atol=0)
