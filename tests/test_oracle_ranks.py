"""Pins the oracle's restatement of the reference's CPU ifprec=2 solve on P
MPI ranks (orc_cgsolve_ranks; the bench's CPU baseline): per-rank diacsr +
Saad ilu0 + lusol, i.e. ILU(0) of the matrix with every coupling between
ranks dropped (src/pot3d.F90:4940-4967, 4684-4709, 5202-5220), inside the
global PCG (:4231-4369).

  - one rank: the single-rank ifprec=2 solve (orc_cgsolve) step for step;
  - P ranks: the first PCG step from x=0 is alpha M^-1 b, so its direction
    pins M^-1 against textbook dense ILU(0) of the rank-restricted matrix
    (Saad Alg. 10.4) and dense triangular solves;
  - P ranks: the same solution as one rank, with no fewer iterations;
  - decompose_dimension / rank_topology against hand-worked cases.
"""
import numpy as np
import pytest

from conftest import random_br0, small_mesh
from test_oracle_ilu import _dense_ilu0, _ilu_matrix


def _rank_restrict(M, mesh, los):
    """M with the couplings between ranks removed; los[d] are the rank
    starts per axis (0-based interior index, last = extent).  The phi wrap
    goes with more than one phi rank."""
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    idx = np.arange(A_ * B_ * C_)
    i, j, k = idx % A_, (idx // A_) % B_, idx // (A_ * B_)
    ri = np.searchsorted(los[0], i, side="right") - 1
    rj = np.searchsorted(los[1], j, side="right") - 1
    rk = np.searchsorted(los[2], k, side="right") - 1
    blk = ri + 1000 * rj + 1000000 * rk
    R = M.copy()
    R[blk[:, None] != blk[None, :]] = 0.0
    return R


def test_decompose_dimension_and_topology(oracle_mod):
    # 11 interior points over 3 ranks: 4, 4, 3 (the first ranks take one more)
    assert oracle_mod.decompose_dimension(13, 3) == [4, 4, 3]
    assert oracle_mod.decompose_dimension(10, 4) == [2, 2, 2, 2]
    # 151x301x601 over 16 ranks, worked by hand: points per rank and penalty
    # (2,2,4): 75.5, 150.5, 150.25 -> 5625 + 0.06 + 5588 = 11213 (the least);
    # (1,4,4): 151, 75.25, 150.25 -> 11364; (1,2,8): 151, 150.5, 75.125 -> 11430
    assert oracle_mod.rank_topology(151, 301, 601, 16) == (2, 2, 4)
    assert oracle_mod.rank_topology(10, 10, 10, 1) == (1, 1, 1)
    # too many ranks for the minimum points per rank (4 in r and theta)
    with pytest.raises(ValueError):
        oracle_mod.rank_topology(7, 7, 7, 64)


def test_one_rank_is_the_single_rank_solve(oracle_mod):
    mesh = small_mesh(13, 11, 18, drratio=2.0)
    o = oracle_mod.Oracle(mesh, -1.0)
    b = o.rhs(random_br0(mesh, 4))
    ref = o.cgsolve(b, ifprec=2, epscg=1e-10)
    got = o.cgsolve_ranks(b, (1, 1, 1), epscg=1e-10, stop_mode=0)
    assert got["ierr"] == 0 and got["ncg"] == ref["ncg"]
    assert np.max(np.abs(got["x"] - ref["x"])) <= 1e-12 * np.max(np.abs(ref["x"]))


@pytest.mark.parametrize("nprocs", [(2, 1, 1), (1, 2, 2), (2, 2, 3)])
def test_rank_preconditioner_is_dense_ilu0_of_the_restricted_matrix(oracle_mod, nprocs):
    mesh = small_mesh(9, 8, 12, drratio=1.3)
    o = oracle_mod.Oracle(mesh, -1.0)
    b = o.rhs(random_br0(mesh, 7))
    res = o.cgsolve_ranks(b, nprocs, epscg=1e-30, stop_mode=0, ncgmax=1)
    assert res["ncg"] == 1
    los = [np.array(v) for v in res["blocks"]]
    M = _rank_restrict(_ilu_matrix(o, mesh), mesh, los)
    if nprocs[2] > 1:  # the phi couplings at every rank boundary, the wrap too
        C_ = mesh.np - 2
        A_, B_ = mesh.nr - 2, mesh.nt - 2
        for p in range(A_ * B_):
            M[p, p + A_ * B_ * (C_ - 1)] = 0.0
            M[p + A_ * B_ * (C_ - 1), p] = 0.0
    LU = _dense_ilu0(M)
    L = np.tril(LU, -1) + np.eye(o.N)
    U = np.triu(LU)
    z = np.linalg.solve(U, np.linalg.solve(L, b.astype(np.float32).astype(np.float64)))
    x1 = res["x"]
    cos = float(x1 @ z) / (np.linalg.norm(x1) * np.linalg.norm(z))
    assert cos > 1 - 1e-10
    # a plausible mistake -- keeping the couplings between ranks -- fails
    LU0 = _dense_ilu0(_ilu_matrix(o, mesh))
    z0 = np.linalg.solve(np.triu(LU0), np.linalg.solve(np.tril(LU0, -1) + np.eye(o.N), b))
    assert float(x1 @ z0) / (np.linalg.norm(x1) * np.linalg.norm(z0)) < 1 - 1e-6


def test_ranks_reach_the_same_solution(oracle_mod):
    mesh = small_mesh(17, 19, 30, drratio=2.0)
    o = oracle_mod.Oracle(mesh, -1.0)
    b = o.rhs(random_br0(mesh, 2))
    one = o.cgsolve_ranks(b, (1, 1, 1), epscg=1e-12, stop_mode=1)
    many = o.cgsolve_ranks(b, (1, 2, 4), epscg=1e-12, stop_mode=1)
    assert one["ierr"] == 0 and many["ierr"] == 0
    assert many["ncg"] >= one["ncg"]
    assert np.max(np.abs(many["x"] - one["x"])) <= 1e-9 * np.max(np.abs(one["x"]))
