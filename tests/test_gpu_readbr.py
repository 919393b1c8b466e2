"""p3d_readbr (the driver's readbr + set_flux arithmetic on the GPU) against
the oracle's restatement (oracle/readbr.py) and the flux balance the
reference printed for its test decks (testsuite/*/reference/pot3d.log).

Bars: before the flux balance the interpolated, seamed and pole-filled br0
is bit-identical (same expression order, no FMA contraction).  The three
flux sums run over up to 7e5 terms, sequentially in the oracle (the
reference's loop order) and as a fixed tree on the GPU; the sequential
sum's rounding reaches about 1e-13 relative at that length (measured 2e-13
on the area of the 181x361 deck), so the sums agree to 1e-11 relative, br00
to 1e-11 of fluxp/area, and br0 after the balance to 1e-13 of its largest
magnitude."""
import math
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DECKS = ["potential_field_source_surface", "potential_field_current_sheet", "open_field"]
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cmp(mesh, tn, pn, bn, balance, abs_br, label, printed=None):
    import readbr as R
    from pot3d_amd import solver
    g, fg = solver.readbr(mesh, tn, pn, bn, balance, abs_br)
    o, fo = R.readbr(mesh, tn, pn, bn, balance, abs_br)
    d = float(np.max(np.abs(g - o)) / np.max(np.abs(o)))
    say(f"    {label}: br0 max rel diff {d:.2e}; fluxp {fg['fluxp']:.15e} "
        f"fluxm {fg['fluxm']:.15e} br00err {fg['br00err']:.6e}")
    for k in ("fluxp", "fluxm", "area"):
        assert abs(fg[k] - fo[k]) <= 1e-11 * abs(fo[k]), (k, fg[k], fo[k])
    assert abs(fg["br00err"] - fo["br00err"]) <= 1e-11 * abs(fo["fluxp"]) / fo["area"]
    assert d <= 1e-13
    if printed is not None:
        assert abs(fg["fluxp"] - printed["fluxp"]) <= 1e-12 * abs(printed["fluxp"])
        assert abs(fg["fluxm"] - printed["fluxm"]) <= 1e-12 * abs(printed["fluxm"])
    # without the balance and the sign flip the GPU result is bit-identical
    g0, _ = solver.readbr(mesh, tn, pn, bn, False, False)
    o0, _ = R.readbr(mesh, tn, pn, bn, False, False)
    assert np.array_equal(g0.view(np.uint64), o0.view(np.uint64))


@pytest.mark.parametrize("deck", DECKS)
def test_readbr_decks(deck):
    from p3d_driver import workloads
    say(f"[gpu] readbr deck {deck}")
    ref = workloads.testsuite_reference(deck)
    nl = ref["namelist"]
    w = workloads.testsuite_deck(deck)
    z = np.load(os.path.join(GOLDEN, f"testsuite_br_{deck}.npz"))
    balance = not (nl["option"] in ("ss", "open") and bool(nl["do_not_balance_flux"]))
    _cmp(w.mesh, z["tn"], z["pn"], z["data"], balance, nl["option"] == "open", deck,
         ref["readbr"])


@pytest.mark.parametrize("grid,nbip,option", [((51, 91, 181), 8, "ss"),
                                              ((301, 601, 1201), 16, "open")])
def test_readbr_synthetic_maps(grid, nbip, option):
    """The synthetic 1-degree maps of the benchmark configurations on their
    meshes (the largest: 601 x 1201 points)."""
    from p3d_driver import inputs, workloads
    say(f"[gpu] readbr synthetic {grid} {option}")
    d = workloads.DOMAINS["open" if option == "open" else "pfss"]
    mesh = inputs.global_mesh(grid[0], grid[1], grid[2], d["r0"], d["r1"], rfrac=(0.0, 1.0),
                              drratio=(workloads.PFSS_DRRATIO,))
    tn, pn, bn = inputs.synthetic_br_map(5, nbip)
    _cmp(mesh, tn, pn, bn, True, option == "open", f"synthetic {grid}")


def test_readbr_errors():
    """intrp2d's two errors come back as P3D_EINVAL before any device work."""
    from p3d_driver import inputs
    from pot3d_amd import _lib, solver
    say("[gpu] readbr errors")
    mesh = inputs.global_mesh(10, 21, 32)
    tn = np.linspace(0, math.pi, 11)
    pn = np.linspace(0, 2 * math.pi, 21)
    bn = np.ones((11, 21))
    bad = tn.copy()
    bad[4] = bad[3]
    with pytest.raises(RuntimeError, match="not monotonically increasing"):
        solver.readbr(mesh, bad, pn, bn)
    with pytest.raises(RuntimeError, match="outside the range"):
        solver.readbr(mesh, tn, pn[:-3], bn[:, :-3])
    assert _lib.last_error()
