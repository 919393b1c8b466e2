"""The oracle's readbr (oracle/readbr.py) pinned: against the flux balance
the reference printed for its three test decks (testsuite/*/reference/
pot3d.log) on the decks' own Br maps and meshes; interp against its literal
linear scan; bilinear functions reproduced; the seam, pole and zero-net-flux
identities."""
import math
import os

import numpy as np
import pytest

import readbr as R
from p3d_driver import inputs, workloads

DECKS = ["potential_field_source_surface", "potential_field_current_sheet", "open_field"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _deck(name):
    ref = workloads.testsuite_reference(name)
    nl = ref["namelist"]
    w = workloads.testsuite_deck(name)
    z = np.load(os.path.join(GOLDEN, f"testsuite_br_{name}.npz"))
    return w.mesh, z, nl, ref["readbr"]


@pytest.mark.parametrize("deck", DECKS)
def test_flux_balance_matches_reference_prints(deck):
    """Positive flux, negative flux and BR00 as the reference printed them
    (16 significant digits) for the deck's map on the deck's mesh."""
    mesh, z, nl, printed = _deck(deck)
    balance = R.balance_for(nl["option"], bool(nl["do_not_balance_flux"]))
    br0, flux = R.readbr(mesh, z["tn"], z["pn"], z["data"], balance, nl["option"] == "open")
    assert abs(flux["fluxp"] - printed["fluxp"]) <= 1e-13 * abs(printed["fluxp"])
    assert abs(flux["fluxm"] - printed["fluxm"]) <= 1e-13 * abs(printed["fluxm"])
    if printed["br00err"] is None:
        assert not balance and flux["br00err"] == 0.0
    else:
        # br00err is a difference of the two fluxes: relative to their size
        assert abs(flux["br00err"] - printed["br00err"]) <= 1e-13 * printed["fluxp"] / flux["area"]
    if nl["option"] == "open":
        assert np.all(br0 >= 0.0)


def test_interp_equals_its_linear_scan():
    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(0.1, 1.0, 40))
    pts = np.concatenate([rng.uniform(x[0], x[-1], 200), x])  # includes every node
    i, a = R.interp_table(x, pts)
    for q, v in enumerate(pts):
        si, sa = R.interp_scan(x, v)
        assert i[q] == si and a[q] == sa
    with pytest.raises(ValueError):
        R.interp_table(x, np.array([x[0] - 1e-9]))
    with pytest.raises(ValueError):
        R.intrp2d(x[::-1].copy(), x, np.zeros((40, 40)), x[:3], x[:3])


def test_bilinear_functions_are_reproduced():
    """f = a + b t + c p + d t p is bilinear on every cell: interpolation
    returns it to rounding."""
    tn = np.concatenate([[0.0], np.sort(np.random.default_rng(1).uniform(0, math.pi, 30)),
                         [math.pi]])
    pn = np.linspace(0.0, 2 * math.pi, 73)
    T, P = np.meshgrid(tn, pn, indexing="ij")
    f = 0.3 + 1.7 * T - 0.4 * P + 0.25 * T * P
    m = inputs.global_mesh(10, 25, 40)
    th, ph = m.th[1:-1], m.ph[1:-1]
    g = R.intrp2d(tn, pn, f, th, ph)
    TT, PP = np.meshgrid(th, ph, indexing="ij")
    exact = 0.3 + 1.7 * TT - 0.4 * PP + 0.25 * TT * PP
    assert np.max(np.abs(g - exact)) <= 1e-13 * np.max(np.abs(exact))


def test_seam_poles_and_zero_net_flux():
    m = inputs.global_mesh(10, 31, 61)
    tn = np.linspace(0, math.pi, 46)
    pn = np.linspace(0, 2 * math.pi, 91)
    T, P = np.meshgrid(tn, pn, indexing="ij")
    bn = 3.0 * np.cos(T) + np.sin(T) * np.cos(P) + 0.7
    br0, flux = R.readbr(m, tn, pn, bn, True, False)
    assert np.array_equal(br0[:, 0], br0[:, -2]) and np.array_equal(br0[:, -1], br0[:, 1])
    da = (m.sth[1:-1, None] * m.dth[1:-1, None]) * m.dph[None, 1:-1]
    assert abs(np.sum(br0[1:-1, 1:-1] * da)) < 1e-12 * np.sum(np.abs(br0[1:-1, 1:-1]) * da)
    w = m.dph[1:-1] / (2 * math.pi)
    s0 = np.sum(br0[1, 1:-1] * w)
    assert np.allclose(br0[0, :], 2 * s0 - br0[1, :])
    # the constant 0.7 has flux 0.7 * area; br00err is the map's mean over the sphere
    assert abs(flux["area"] - 4 * math.pi) < 1e-2
    assert abs(flux["br00err"] - (flux["fluxp"] + flux["fluxm"]) / flux["area"]) == 0.0
    # |Br| for 'open' after the balance
    b2, _ = R.readbr(m, tn, pn, bn, True, True)
    assert np.array_equal(b2, np.abs(br0))
