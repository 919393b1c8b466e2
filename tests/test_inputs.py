"""Driver-side inputs: genmesh/set_global_mesh against the mesh tables the
reference printed for its three test decks (8 decimals,
testsuite/*/reference/pot3d.out), and readbr's post-processing."""
import math

import numpy as np
import pytest

from pot3d_amd import inputs, workloads

DECKS = ["potential_field_source_surface", "potential_field_current_sheet", "open_field"]


@pytest.mark.parametrize("deck", DECKS)
def test_mesh_matches_reference_tables(deck):
    w = workloads.testsuite_deck(deck)
    ref = workloads.testsuite_reference(deck)["mesh"]
    m = w.mesh
    for coord, arr in (("r", m.r), ("t", m.t), ("p", m.p[:-1])):
        tab = np.array([v for _, v in ref[coord]])
        assert len(tab) == len(arr)
        assert np.max(np.abs(arr - tab)) < 6e-9


def test_half_mesh_relations():
    m = inputs.global_mesh(20, 17, 33, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(3.0,))
    assert np.allclose(m.rh[1:-1], 0.5 * (m.r[1:] + m.r[:-1]))
    assert np.allclose(m.dr, m.rh[1:] - m.rh[:-1])
    assert m.t[0] == 0.0 and m.t[-1] == inputs.PI
    assert m.dp[0] == m.dp[-2] and m.dph[-1] == m.dph[1]


def test_readbr_pole_periodic_and_flux_balance():
    m = inputs.global_mesh(10, 31, 61)
    tn = np.linspace(0, math.pi, 46)
    pn = np.linspace(0, 2 * math.pi, 91)
    T, P = np.meshgrid(tn, pn, indexing="ij")
    bn = 3.0 * np.cos(T) + np.sin(T) * np.cos(P) + 0.7
    br0 = inputs.readbr_interp(m, tn, pn, bn, option="ss")
    assert np.array_equal(br0[:, 0], br0[:, -2]) and np.array_equal(br0[:, -1], br0[:, 1])
    da = (m.sth[1:-1, None] * m.dth[1:-1, None]) * m.dph[None, 1:-1]
    assert abs(np.sum(br0[1:-1, 1:-1] * da)) < 1e-12 * np.sum(np.abs(br0[1:-1, 1:-1]) * da)
    w = m.dph[1:-1] / (2 * math.pi)
    s0 = np.sum(br0[1, 1:-1] * w)
    assert np.allclose(br0[0, :], 2 * s0 - br0[1, :])


def test_synthetic_br_is_seeded_and_balanced():
    m = inputs.global_mesh(10, 46, 91)
    a = inputs.synthetic_br0(m, 3)
    b = inputs.synthetic_br0(m, 3)
    c = inputs.synthetic_br0(m, 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.abs(a).max() > 40.0


def test_option_checks():
    assert inputs.pm_r1_for("potential") == 1.0
    assert inputs.pm_r1_for("ss") == -1.0 and inputs.pm_r1_for("open") == -1.0
    with pytest.raises(ValueError):
        inputs.pm_r1_for("bogus")
