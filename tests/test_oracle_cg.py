"""Pins the oracle's PCG (cgsolve/ernorm, src/pot3d.F90:4231-4485), its
post-solve phi, getb and magnetic energy (:6187-6321), against a dense
direct solve, closed forms, and the reference test suite's recorded runs."""
import math

import numpy as np
import pytest

from conftest import random_br0, small_mesh


@pytest.mark.parametrize("ifprec", [0, 1, 2, 3])
def test_cgsolve_matches_direct_solve(oracle_mod, ifprec):
    from test_oracle_operator import _assemble
    mesh = small_mesh(9, 8, 12)
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 11)
    b = o.rhs(br0)
    A = _assemble(o, mesh, -1.0).toarray()
    xref = np.linalg.solve(A, b)
    res = o.cgsolve(b, ifprec=ifprec, epscg=1e-13)
    assert res["ierr"] == 0
    assert np.max(np.abs(res["x"] - xref)) < 1e-9 * np.max(np.abs(xref))
    # on exit r holds the residual b - A x (to rounding)
    assert np.max(np.abs(res["r"] - (b - o.ax(res["x"])))) < 1e-10 * np.max(np.abs(b))


def test_zero_rhs_returns_zero(oracle_mod):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, -1.0)
    res = o.cgsolve(np.zeros(o.N), ifprec=1, x0=np.ones(o.N))
    assert res["ierr"] == 0 and res["ncg"] == 0 and np.all(res["x"] == 0)


def test_ncgmax_gives_ierr1(oracle_mod):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, -1.0)
    res = o.cgsolve(o.rhs(random_br0(mesh, 1)), ifprec=1, epscg=1e-12, ncgmax=3)
    assert res["ierr"] == 1 and res["ncg"] == 3


def test_energy_of_uniform_field(oracle_mod):
    """phi = r cos(theta) gives B = grad phi with |B| = 1, so the energy is
    half the shell volume (2 pi/3)(r1^3 - r0^3), to O(h^2)."""
    from p3d_driver import inputs

    def rel_err(nr, nt, np_):
        mesh = inputs.global_mesh(nr, nt, np_, 1.0, 2.0, rfrac=(0.0, 1.0), drratio=(1.0,))
        o = oracle_mod.Oracle(mesh, -1.0)
        R, T, _ = np.meshgrid(mesh.rh, mesh.th, mesh.ph, indexing="ij")
        phi = (R * np.cos(T)).ravel(order="F")
        w = o.energy(phi)
        # the energy sums the interior cells: shell between r(1) and r(nrm1)
        exact = 0.5 * (4.0 * math.pi / 3.0) * (mesh.r[-1] ** 3 - mesh.r[0] ** 3)
        return abs(w.sum() - exact) / exact

    e1 = rel_err(10, 20, 24)
    e2 = rel_err(18, 38, 46)
    assert e1 < 0.05 and e2 < e1 / 3.0


def test_fill_phi_boundary_values(oracle_mod):
    mesh = small_mesh()
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = random_br0(mesh, 3)
    x = np.random.default_rng(0).standard_normal(o.N)
    phi = o.fill_phi(x, br0).reshape(mesh.nr, mesh.nt, mesh.np, order="F")
    inner = phi[1:-1, 1:-1, 1:-1].transpose(2, 1, 0).ravel()
    assert np.array_equal(inner, x)
    assert np.allclose(phi[0, 1:-1, 1:-1], phi[1, 1:-1, 1:-1] - br0[1:-1, 1:-1] * mesh.dr[0])
    assert np.array_equal(phi[-1, 1:-1, 1:-1], -phi[-2, 1:-1, 1:-1])
    assert np.array_equal(phi[:, :, 0], phi[:, :, -2])
    assert np.array_equal(phi[:, :, -1], phi[:, :, 1])
    w = mesh.dph[1:-1] / (2.0 * math.pi)
    s0 = (phi[:, 1, 1:-1] * w).sum(axis=1)
    assert np.allclose(phi[:, 0, 1:-1], 2 * s0[:, None] - phi[:, 1, 1:-1], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("deck", [
    "potential_field_source_surface",
    pytest.param("potential_field_current_sheet", marks=pytest.mark.slow),
    pytest.param("open_field", marks=pytest.mark.slow),
])
def test_testsuite_deck_energies(oracle_mod, deck):
    """The reference's own pass rule: energies equal to 7 decimal places
    (bin/pot3d_validation.sh), plus the recorded iteration count and
    residual history (testsuite/<deck>/reference/pot3d.out)."""
    from p3d_driver import workloads
    w = workloads.testsuite_deck(deck)
    ref = workloads.testsuite_reference(deck)
    o = oracle_mod.Oracle(w.mesh, w.pm_r1)
    res = o.cgsolve(o.rhs(w.br0), ifprec=1, epscg=w.epscg, nhist=ref["iterations"] + 2)
    assert abs(res["ncg"] - ref["iterations"]) <= 1
    for it, val in ref["residual_history"][:-1]:
        assert abs(res["hist"][it] - val) <= 1e-6 * val
    e = o.energy(o.fill_phi(res["x"], w.br0))
    ee = ref["energy"]
    for got, key in zip(e, ("br2", "bt2", "bp2")):
        assert round(got, 7) == round(ee[key], 7)
