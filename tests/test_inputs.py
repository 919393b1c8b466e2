"""Driver-side inputs: genmesh/set_global_mesh against the mesh tables the
reference printed for its three test decks (8 decimals,
testsuite/*/reference/pot3d.out), and the seeded synthetic Br maps."""
import math

import numpy as np
import pytest

from p3d_driver import inputs, workloads

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


def test_synthetic_br_is_seeded_and_balanced():
    m = inputs.global_mesh(10, 46, 91)
    a = workloads.synthetic_br0(m, 3)
    b = workloads.synthetic_br0(m, 3)
    c = workloads.synthetic_br0(m, 4)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.abs(a).max() > 40.0


def test_option_checks():
    assert inputs.pm_r1_for("potential") == 1.0
    assert inputs.pm_r1_for("ss") == -1.0 and inputs.pm_r1_for("open") == -1.0
    with pytest.raises(ValueError):
        inputs.pm_r1_for("bogus")
