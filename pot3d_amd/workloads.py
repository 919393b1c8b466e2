"""Named workloads: the reference test-suite decks and the synthetic PFSS/PFCS
configurations (grid, mesh options, Br map, boundary option).

A workload is everything the driver hands to the solve: the Mesh, br0(nt,np)
and pm_r1.  The test-suite decks reproduce testsuite/*/input/pot3d.dat of the
reference (their Br maps are committed fixtures); the synthetic ones are
seeded and built here.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass

import numpy as np

from . import inputs

_GOLDEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests",
                       "golden")


@dataclass
class Workload:
    name: str
    mesh: inputs.Mesh
    br0: np.ndarray          # (nt, np), C array holding Fortran br0(j,k) at [j-1,k-1]
    pm_r1: float
    option: str
    epscg: float = 1e-9
    ifprec: int = 1

    @property
    def N(self):
        return self.mesh.N


def _as_tuple(v):
    return tuple(v) if isinstance(v, (list, tuple)) else (v,)


def testsuite_deck(name: str) -> Workload:
    """One of 'potential_field_source_surface', 'potential_field_current_sheet',
    'open_field' (testsuite/<name>/input/pot3d.dat + its Br map)."""
    ref = json.load(open(os.path.join(_GOLDEN, "testsuite_reference.json")))[name]
    nl = ref["namelist"]
    mesh = inputs.global_mesh(int(nl["nr"]), int(nl["nt"]), int(nl["np"]), nl["r0"], nl["r1"],
                              rfrac=_as_tuple(nl["rfrac"]), drratio=_as_tuple(nl["drratio"]),
                              nfrmesh=int(nl["nfrmesh"]), tfrac=_as_tuple(nl["tfrac"]),
                              dtratio=_as_tuple(nl["dtratio"]), nftmesh=int(nl["nftmesh"]),
                              pfrac=_as_tuple(nl["pfrac"]), dpratio=_as_tuple(nl["dpratio"]),
                              nfpmesh=int(nl["nfpmesh"]), phishift=nl["phishift"])
    z = np.load(os.path.join(_GOLDEN, f"testsuite_br_{name}.npz"))
    br0 = inputs.readbr_interp(mesh, z["tn"], z["pn"], z["data"], option=nl["option"],
                               do_not_balance_flux=bool(nl["do_not_balance_flux"]))
    br0 = inputs.set_flux(br0, nl["option"])
    return Workload(name, mesh, br0, inputs.pm_r1_for(nl["option"]), nl["option"],
                    nl["epscg"], int(nl["ifprec"]))


def testsuite_reference(name: str) -> dict:
    return json.load(open(os.path.join(_GOLDEN, "testsuite_reference.json")))[name]


# Per-cell radial stretch 1.03 over the 49 interior cells of the 51-point
# configuration; kept as the segment ratio for the finer grids so that the
# grids nest in the mesh coordinate xi.
PFSS_DRRATIO = 1.03 ** 49


def pfss(nr: int, nt: int, np_: int, seed: int = 0, n_bipoles: int = 8, option: str = "ss",
         r0: float = 1.0, r1: float = 2.5, drratio: float = PFSS_DRRATIO, ifprec: int = 1,
         name: str | None = None) -> Workload:
    """Synthetic PFSS-style workload: geometric r mesh in [r0,r1], uniform theta
    and phi, seeded bipole + harmonic Br at r0."""
    mesh = inputs.global_mesh(nr, nt, np_, r0, r1, rfrac=(0.0, 1.0), drratio=(drratio,))
    br0 = inputs.synthetic_br0(mesh, seed, n_bipoles=n_bipoles, option=option)
    return Workload(name or f"pfss_{nr}x{nt}x{np_}_s{seed}", mesh, br0,
                    inputs.pm_r1_for(option), option, 1e-9, ifprec)


def config(idx: int, seed: int = 0) -> Workload:
    """The synthetic workloads, by number:
      1  51x91x181 PFSS, ifprec=1
      2  151x301x601 PFSS, ifprec=2 (ILU0)
      3  51x91x181 'potential' (Neumann dphi/dr=0 at r1), ILU0
      5  301x601x1201 'open', r in [2.5,10], 16 bipoles, ILU0"""
    if idx == 1:
        return pfss(51, 91, 181, seed, ifprec=1, name=f"cfg1_51x91x181_s{seed}")
    if idx == 2:
        return pfss(151, 301, 601, seed, ifprec=2, name=f"cfg2_151x301x601_s{seed}")
    if idx == 3:
        return pfss(51, 91, 181, seed, option="potential", ifprec=2,
                    name=f"cfg3_51x91x181_potential_s{seed}")
    if idx == 5:
        return pfss(301, 601, 1201, seed, n_bipoles=16, option="open", r0=2.5, r1=10.0,
                    drratio=1.03 ** 49, ifprec=2, name=f"cfg5_301x601x1201_s{seed}")
    raise ValueError(f"no workload {idx}")
