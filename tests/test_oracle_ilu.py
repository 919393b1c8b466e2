"""Pins the oracle's ILU0 path against textbook dense ILU(0) and dense
triangular solves on tiny grids.

Reference: getM src/pot3d.F90:5020-5100, diacsr :4870-5018, ilu0 (Saad IKJ)
:4717-4813, lu2luopt :4815-4868, lusol :5227-5302; the D-ILU restatement
(pivots only) is checked to be the same factorisation.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

from conftest import small_mesh


def _ilu_matrix(o, mesh):
    """Dense ILU matrix assembled from the coefficient planes: stored a with
    the r and theta boundary couplings dropped and the phi wrap kept."""
    a = o.coefs()
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    N = A_ * B_ * C_
    M = np.zeros((N, N))
    for k in range(C_):
        for j in range(B_):
            for i in range(A_):
                p = i + A_ * (j + B_ * k)
                M[p, p] = a[3, p]
                M[p, i + A_ * (j + B_ * ((k - 1) % C_))] += a[0, p]
                M[p, i + A_ * (j + B_ * ((k + 1) % C_))] += a[6, p]
                if j > 0:
                    M[p, p - A_] = a[1, p]
                if j < B_ - 1:
                    M[p, p + A_] = a[5, p]
                if i > 0:
                    M[p, p - 1] = a[2, p]
                if i < A_ - 1:
                    M[p, p + 1] = a[4, p]
    return M


def _dense_ilu0(M):
    """Saad, Iterative Methods for Sparse Linear Systems, Alg. 10.4 (IKJ),
    restricted to the nonzero pattern of M.  Returns the combined LU."""
    A = M.copy()
    S = M != 0
    n = A.shape[0]
    for i in range(1, n):
        for k in range(i):
            if not S[i, k]:
                continue
            A[i, k] = A[i, k] / A[k, k]
            js = np.nonzero(S[i, k + 1:])[0] + k + 1
            A[i, js] -= A[i, k] * A[k, js]
    return A


@pytest.fixture(scope="module")
def tiny(oracle_mod):
    mesh = small_mesh(7, 6, 10, drratio=1.3)
    o = oracle_mod.Oracle(mesh, -1.0)
    return mesh, o


def test_diacsr_is_the_ilu_matrix(tiny):
    mesh, o = tiny
    IA, JA, Av, dptr = o.csr()
    N = o.N
    nnz = IA[-1] - 1
    assert len(Av) == nnz + 1  # M = nnz+1, the reference's allocated length
    C = sp.csr_matrix((Av[:nnz].astype(np.float64), JA[:nnz] - 1, IA - 1), shape=(N, N)).toarray()
    M = _ilu_matrix(o, mesh).astype(np.float32).astype(np.float64)
    assert np.array_equal(C, M)
    for r in range(N):  # columns ascending within every row, diag pointer
        cols = JA[IA[r] - 1:IA[r + 1] - 1]
        assert np.all(np.diff(cols) > 0)
        assert JA[dptr[r] - 1] == r + 1


def test_ilu0_matches_dense_textbook(tiny):
    mesh, o = tiny
    IA, JA, LU, dptr, icode = o.ilu0()
    assert icode == 0
    N = o.N
    nnz = IA[-1] - 1
    got = sp.csr_matrix((LU[:nnz].astype(np.float64), JA[:nnz] - 1, IA - 1),
                        shape=(N, N)).toarray()
    M = _ilu_matrix(o, mesh).astype(np.float32).astype(np.float64)
    ref = _dense_ilu0(M)
    mask = M != 0
    assert np.allclose(got[mask], ref[mask], rtol=2e-5, atol=0)
    # fill stays on the pattern and changes only the diagonal + L scaling
    U_off = np.triu(mask, 1)
    assert np.array_equal(got[U_off], M[U_off].astype(np.float32).astype(np.float64))


def test_dilu_pivots_equal_ilu0_diagonal(tiny):
    mesh, o = tiny
    M = _ilu_matrix(o, mesh)
    ref = np.diag(_dense_ilu0(M))
    d = o.dilu_pivots()
    assert np.allclose(d, ref, rtol=1e-13, atol=0)


def test_lusol_matches_dense_triangular_solves(tiny):
    mesh, o = tiny
    M = _ilu_matrix(o, mesh).astype(np.float32).astype(np.float64)
    LU = _dense_ilu0(M)
    L = np.tril(LU, -1) + np.eye(o.N)
    U = np.triu(LU)
    rng = np.random.default_rng(2)
    r = rng.standard_normal(o.N).astype(np.float32)
    ref = sla.solve_triangular(U, sla.solve_triangular(L, r.astype(np.float64), lower=True,
                                                       unit_diagonal=True), lower=False)
    got = o.lusol(r)
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))


def test_dilu_apply_matches_dense(tiny):
    mesh, o = tiny
    M = _ilu_matrix(o, mesh)
    LU = _dense_ilu0(M)
    L = np.tril(LU, -1) + np.eye(o.N)
    U = np.triu(LU)
    rng = np.random.default_rng(4)
    r = rng.standard_normal(o.N).astype(np.float32)
    ref = np.linalg.solve(L @ U, r.astype(np.float64))
    got = o.dilu_apply(r)
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))
    # and the reference's FP32 lusol agrees with the D-ILU apply
    assert np.max(np.abs(got - o.lusol(r))) < 1e-4 * np.max(np.abs(ref))


def _block_restrict(M, mesh, blocks):
    """M with every coupling between two blocks of bi x bj x bk points
    removed (block starts at multiples of the extents); the phi wrap goes
    with more than one phi block."""
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    bi, bj, bk = blocks
    idx = np.arange(A_ * B_ * C_)
    i, j, k = idx % A_, (idx // A_) % B_, idx // (A_ * B_)
    blk = (i // bi) + 1000 * (j // bj) + 1000000 * (k // bk)
    R = M.copy()
    R[blk[:, None] != blk[None, :]] = 0.0
    return R


@pytest.mark.parametrize("blocks", [(3, 2, 4), (7, 2, 10), (5, 4, 3), (2, 6, 10)])
def test_block_jacobi_dilu_matches_dense(tiny, blocks):
    """Block-Jacobi D-ILU = textbook ILU(0) of the block-restricted matrix:
    pivots to 1e-13, apply to FP32 rounding."""
    mesh, o = tiny
    M = _block_restrict(_ilu_matrix(o, mesh), mesh, blocks)
    LU = _dense_ilu0(M)
    d = o.dilu_pivots(blocks)
    assert np.allclose(d, np.diag(LU), rtol=1e-13, atol=0)
    L = np.tril(LU, -1) + np.eye(o.N)
    U = np.triu(LU)
    rng = np.random.default_rng(9)
    r = rng.standard_normal(o.N).astype(np.float32)
    ref = np.linalg.solve(L @ U, r.astype(np.float64))
    got = o.dilu_apply(r, blocks=blocks)
    assert np.max(np.abs(got - ref)) < 1e-4 * np.max(np.abs(ref))
    # the unblocked extents give the plain D-ILU bit for bit
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    assert np.array_equal(o.dilu_pivots((A_, B_, C_)), o.dilu_pivots())
    assert np.array_equal(o.dilu_apply(r, blocks=(A_, B_, C_)), o.dilu_apply(r))


def test_block_is_the_references_per_rank_ilu(oracle_mod):
    """One block of the block-Jacobi D-ILU is the reference's ILU0 of that
    rank's local grid: diacsr on the local (nr, nt, np) drops exactly the
    couplings at the local r and theta boundaries (:4940-4956).  Checked on
    an r x theta block (one phi block keeps the wrap, as with nproc_p=1):
    the block's pivots equal the D-ILU pivots of the same sub-grid solved on
    its own, whose couplings to the outside are gone."""
    mesh = small_mesh(11, 10, 12, drratio=1.4)
    o = oracle_mod.Oracle(mesh, -1.0)
    A_, B_, C_ = mesh.nr - 2, mesh.nt - 2, mesh.np - 2
    blocks = (5, 4, C_)
    d = o.dilu_pivots(blocks).reshape(C_, B_, A_)
    a = o.coefs().reshape(7, C_, B_, A_)
    # the first block (i < 5, j < 4): rebuild its pivots from its own coefficients
    sub = a[:, :, :4, :5]
    ref = np.zeros((C_, 4, 5))
    for k in range(C_):
        for j in range(4):
            for i in range(5):
                dp = sub[3, k, j, i]
                if k == C_ - 1:
                    dp -= sub[6, k, j, i] / ref[0, j, i] * sub[0, 0, j, i]
                if k > 0:
                    dp -= sub[0, k, j, i] / ref[k - 1, j, i] * sub[6, k - 1, j, i]
                if j > 0:
                    dp -= sub[1, k, j, i] / ref[k, j - 1, i] * sub[5, k, j - 1, i]
                if i > 0:
                    dp -= sub[2, k, j, i] / ref[k, j, i - 1] * sub[4, k, j, i - 1]
                ref[k, j, i] = dp
    assert np.allclose(d[:, :4, :5], ref, rtol=1e-13, atol=0)
