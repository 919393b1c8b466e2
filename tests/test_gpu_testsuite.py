"""End-to-end on the GPU: the reference test suite's three decks
(testsuite/*/input), solved by libpot3d_hip.so and checked by the
reference's own rule -- energies equal to 7 decimal places
(bin/pot3d_validation.sh) -- plus the recorded CG iteration counts
(ifprec=1).  The ILU0 solve (-pc=2 in run_test_suite.sh) must give the
same energies."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DECKS = ["potential_field_source_surface", "potential_field_current_sheet", "open_field"]


@pytest.mark.parametrize("ifprec", [1, 2])
@pytest.mark.parametrize("deck", DECKS)
def test_deck(deck, ifprec):
    print(f"[gpu] testsuite deck={deck} ifprec={ifprec}", flush=True)
    sys.stdout.flush()
    from pot3d_amd import solver
    from p3d_driver import workloads
    w = workloads.testsuite_deck(deck)
    ref = workloads.testsuite_reference(deck)
    res = solver.potfld(w, ifprec=ifprec, ncgmax=3 * ref["iterations"])
    assert res.ierr == 0
    if ifprec == 1:
        assert abs(res.ncg - ref["iterations"]) <= 2
    else:
        assert res.ncg < ref["iterations"]
    for got, key in zip(res.energy, ("br2", "bt2", "bp2")):
        # bin/pot3d_validation.sh compares values rounded to 7 decimals; the
        # equivalent tolerance, robust to straddling a rounding boundary:
        assert abs(got - ref["energy"][key]) < 1e-7, (key, got, ref["energy"][key])
