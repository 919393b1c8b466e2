"""Named workloads: the reference test-suite decks and the synthetic PFSS/PFCS
configurations (grid, mesh options, Br map, boundary option).

A workload is everything the driver hands to the solve: the Mesh, br0(nt,np)
and pm_r1.  The test-suite decks reproduce testsuite/*/input/pot3d.dat of the
reference (their Br maps are committed fixtures); the synthetic ones are
seeded and built here.  br0 is the driver's readbr + set_flux of the map:
by default the library's p3d_readbr on the GPU; the CPU tests plug in the
oracle's restatement with set_readbr.
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


# readbr backend: fn(mesh, tn, pn, bn, balance_flux, abs_br) -> br0(nt, np)
_READBR = None


def set_readbr(fn):
    """Selects how br0 is made from a map (None: the library's p3d_readbr)."""
    global _READBR
    _READBR = fn


def readbr(mesh: inputs.Mesh, tn, pn, bn, option: str, do_not_balance_flux: bool = False):
    """br0(nt, np) of the map bn(len(tn), len(pn)): readbr (interpolation,
    seam, poles, flux balance unless 'ss'/'open' with do_not_balance_flux,
    src/pot3d.F90:6751-6829) and set_flux (|Br| for 'open', :3806-3833)."""
    balance = not (option in ("ss", "open") and do_not_balance_flux)
    fn = _READBR
    if fn is None:
        from pot3d_amd import solver
        return solver.readbr(mesh, tn, pn, bn, balance, option == "open")[0]
    return fn(mesh, tn, pn, bn, balance, option == "open")


def synthetic_br0(mesh: inputs.Mesh, seed: int, n_bipoles: int = 8, option: str = "ss",
                  map_res_deg: float = 1.0):
    """br0 of the seeded synthetic map (inputs.synthetic_br_map, sampled every
    map_res_deg degrees) through readbr and set_flux."""
    nth = int(round(180.0 / map_res_deg)) + 1
    nph = int(round(360.0 / map_res_deg)) + 1
    tn, pn, bn = inputs.synthetic_br_map(seed, n_bipoles, nth, nph)
    return readbr(mesh, tn, pn, bn, option)


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
    br0 = readbr(mesh, z["tn"], z["pn"], z["data"], nl["option"],
                 bool(nl["do_not_balance_flux"]))
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
    br0 = synthetic_br0(mesh, seed, n_bipoles=n_bipoles, option=option)
    return Workload(name or f"pfss_{nr}x{nt}x{np_}_s{seed}", mesh, br0,
                    inputs.pm_r1_for(option), option, 1e-9, ifprec)


# The two problem families of the synthetic workloads: boundary option,
# radial extent and number of bipoles of the Br map.
DOMAINS = {
    "pfss": dict(option="ss", r0=1.0, r1=2.5, n_bipoles=8),
    "open": dict(option="open", r0=2.5, r1=10.0, n_bipoles=16),
}

# The coarse grid the warm-start network is trained on (51x91x181).
COARSE_GRID = (51, 91, 181)


def domain_workload(domain: str, grid, seed: int = 0, ifprec: int = 2) -> Workload:
    """The synthetic workload of one problem family on the grid (nr, nt, np)."""
    d = DOMAINS[domain]
    nr, nt, np_ = grid
    return pfss(nr, nt, np_, seed, n_bipoles=d["n_bipoles"], option=d["option"], r0=d["r0"],
                r1=d["r1"], ifprec=ifprec, name=f"{domain}_{nr}x{nt}x{np_}_s{seed}")


def config(idx: int, seed: int = 0) -> Workload:
    """The synthetic workloads, by number:
      1  51x91x181 PFSS, ifprec=1
      2  151x301x601 PFSS, ifprec=2 (ILU0)
      3  51x91x181 'potential' (Neumann dphi/dr=0 at r1), ILU0
      4  151x301x601 PFSS, ILU0, held-out seed for the warm start trained
         on 51x91x181 (workload 2 with that seed)
      5  301x601x1201 'open', r in [2.5,10], 16 bipoles, ILU0"""
    if idx == 1:
        return pfss(51, 91, 181, seed, ifprec=1, name=f"pfss_diag_51x91x181_s{seed}")
    if idx == 2:
        return pfss(151, 301, 601, seed, ifprec=2, name=f"pfss_151x301x601_s{seed}")
    if idx == 3:
        return pfss(51, 91, 181, seed, option="potential", ifprec=2,
                    name=f"potential_51x91x181_s{seed}")
    if idx == 4:
        return pfss(151, 301, 601, seed, ifprec=2, name=f"pfss_heldout_151x301x601_s{seed}")
    if idx == 5:
        return pfss(301, 601, 1201, seed, n_bipoles=16, option="open", r0=2.5, r1=10.0,
                    drratio=1.03 ** 49, ifprec=2, name=f"open_301x601x1201_s{seed}")
    raise ValueError(f"no workload {idx}")
