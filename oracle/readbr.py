"""Oracle (test infrastructure, never the product): numpy restatement of the
driver's Br-input arithmetic, the CPU side of the library's p3d_readbr.

  interp            src/pot3d.F90:7028-7072
  intrp2d           src/pot3d.F90:6830-6910 (bilinear; scale checks :6870-6889)
  readbr            src/pot3d.F90:6751-6829 (seam, m=0 poles, flux balance)
  set_flux          src/pot3d.F90:3806-3833 (|Br| for option 'open')

Pins (tests/test_oracle_readbr.py): the flux balance the reference printed
for its three test decks (testsuite/*/reference/pot3d.log: positive and
negative flux, BR00) on the decks' own Br maps and meshes; bilinear
functions reproduced exactly; interp against its literal linear scan
(including points that hit a node); the seam, pole and zero-net-flux
identities.
"""
from __future__ import annotations

import numpy as np

PI = 3.1415926535897932  # src/pot3d.F90:108


def interp_scan(x: np.ndarray, xv: float):
    """interp's loop as written: the first i with x(i) <= xv <= x(i+1)
    (0-based here), alpha=(xv-x(i))/(x(i+1)-x(i)); None when not found."""
    for i in range(len(x) - 1):
        if x[i] <= xv <= x[i + 1]:
            return i, (xv - x[i]) / (x[i + 1] - x[i])
    return None


def interp_table(x: np.ndarray, xv: np.ndarray):
    """interp for many points (vectorised; equal to interp_scan): the first
    interval containing xv is the smallest i with x(i+1) >= xv."""
    x = np.asarray(x, float)
    xv = np.asarray(xv, float)
    if np.any(xv < x[0]) or np.any(xv > x[-1]):
        raise ValueError("intrp2d: point outside the range of the scales")
    i = np.searchsorted(x, xv, side="left") - 1
    i = np.clip(i, 0, len(x) - 2)
    alpha = (xv - x[i]) / (x[i + 1] - x[i])
    return i, alpha


def intrp2d(xi: np.ndarray, yi: np.ndarray, fi: np.ndarray, x: np.ndarray, y: np.ndarray):
    """Bilinear interpolation of fi(len(xi), len(yi)) onto x by y, in the
    reference's expression order (:6891-6892)."""
    if np.any(np.diff(xi) <= 0) or np.any(np.diff(yi) <= 0):
        raise ValueError("intrp2d: scales are not monotonic")
    ii, ax = interp_table(xi, x)
    jj, ay = interp_table(yi, y)
    ax = ax[:, None]
    ay = ay[None, :]
    f00 = fi[np.ix_(ii, jj)]
    f01 = fi[np.ix_(ii, jj + 1)]
    f10 = fi[np.ix_(ii + 1, jj)]
    f11 = fi[np.ix_(ii + 1, jj + 1)]
    return ((1.0 - ax) * ((1.0 - ay) * f00 + ay * f01)
            + ax * ((1.0 - ay) * f10 + ay * f11))


def _seq_sum(a: np.ndarray) -> float:
    """Sequential FP64 sum in array order (the Fortran loop's order)."""
    return float(np.cumsum(a.ravel())[-1]) if a.size else 0.0


def readbr(mesh, tn, pn, bn, balance_flux: bool = True, abs_br: bool = False,
           pl: float = 2.0 * PI - 0.0):
    """readbr after rdhdf_2d, then set_flux: returns (br0, flux) with br0 a
    (nt, np) C array holding the Fortran br0(j,k) at [j-1, k-1] and flux =
    dict(fluxp, fluxm, area, br00err).  balance_flux False is the reference's
    ('ss' or 'open') with do_not_balance_flux; abs_br is option 'open'."""
    nt, np_ = mesh.nt, mesh.np
    br0 = np.zeros((nt, np_))
    br0[1:nt - 1, 1:np_ - 1] = intrp2d(np.asarray(tn, float), np.asarray(pn, float),
                                       np.asarray(bn, float), np.asarray(mesh.th)[1:nt - 1],
                                       np.asarray(mesh.ph)[1:np_ - 1])
    return finish(mesh, br0, balance_flux, abs_br, pl)


def finish(mesh, br0, balance_flux: bool = True, abs_br: bool = False,
           pl: float = 2.0 * PI - 0.0):
    """readbr's steps after the interpolation, on br0 (nt, np) whose interior
    is set (modified in place where possible); returns (br0, flux)."""
    nt, np_ = mesh.nt, mesh.np
    # periodic in phi (:6777-6778)
    br0[:, 0] = br0[:, np_ - 2]
    br0[:, np_ - 1] = br0[:, 1]
    # m=0 pole ghosts (:6784-6788): SUM over k=2..npm1 in order, times pl_i
    pl_i = 1.0 / pl
    dph = np.asarray(mesh.dph)
    sum0 = 0.0
    sum1 = 0.0
    for k in range(1, np_ - 1):
        sum0 = sum0 + br0[1, k] * dph[k]
        sum1 = sum1 + br0[nt - 2, k] * dph[k]
    sum0 = sum0 * pl_i
    sum1 = sum1 * pl_i
    br0[0, :] = 2.0 * sum0 - br0[1, :]
    br0[nt - 1, :] = 2.0 * sum1 - br0[nt - 2, :]
    # flux balance (:6792-6823): j outer, k inner, as the reference's loops
    sth, dth = np.asarray(mesh.sth), np.asarray(mesh.dth)
    da = (sth[1:nt - 1, None] * dth[1:nt - 1, None]) * dph[None, 1:np_ - 1]
    v = br0[1:nt - 1, 1:np_ - 1]
    fluxp = _seq_sum(np.where(v > 0.0, v * da, 0.0))
    fluxm = _seq_sum(np.where(v > 0.0, 0.0, v * da))
    area = _seq_sum(da)
    br00err = 0.0
    if balance_flux:
        br00err = (fluxp + fluxm) / area
        br0 = br0 - br00err
    if abs_br:
        br0 = np.abs(br0)
    return br0, dict(fluxp=fluxp, fluxm=fluxm, area=area, br00err=br00err)


def balance_for(option: str, do_not_balance_flux: bool) -> bool:
    """readbr balances the flux unless the option is 'ss' or 'open' with
    do_not_balance_flux set (:6812-6813)."""
    return not (option in ("ss", "open") and do_not_balance_flux)
