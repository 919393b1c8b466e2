"""Driver-side inputs for the solve: meshes and the photospheric Br maps.

POT3D's Fortran driver (not the solve path) builds the 1-D mesh arrays and
reads the Br map before calling the solver, whose C ABI takes them as given.
This module restates the parts of the driver that the tests, the bench, the
smoke run and the warm-start training need to produce the same inputs the
reference would hand to its solver:

* ``genmesh``            -- src/pot3d.F90:2864-3372 (+ map_xi_to_c :3374-3470,
                            map_c_to_xi :3472-3545, filters :3547-3700, fold)
* ``global_mesh``        -- set_global_mesh, src/pot3d.F90:2583-2706
* ``pm_r1_for``          -- the r1 boundary sign of each option (:1250-1285)
* ``synthetic_br_map``   -- the seeded bipole + spherical-harmonic Br map
                            of the synthetic workloads (no reference
                            counterpart; the reference reads an HDF5 map).

readbr's interpolation, seam, pole and flux-balance arithmetic runs in the
library (p3d_readbr); workloads.py hands it the maps.

Everything is float64 and follows the reference's operation order where
it affects rounding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = 3.1415926535897932  # src/pot3d.F90:108
PL = 2.0 * PI - 0.0      # pl = p1-p0, src/pot3d.F90:192-196
PL_I = 1.0 / PL


# --------------------------------------------------------------------------
# genmesh and helpers
# --------------------------------------------------------------------------
def _fold(x0: float, x1: float, x: float) -> float:
    """fold: src/pot3d.F90:3702-3752 (Fortran MOD keeps the sign of x-x0)."""
    if x0 == x1:
        return x
    xl = x1 - x0
    v = math.fmod(x - x0, xl) + x0
    if v < x0:
        v = v + xl
    if v >= x1:
        v = v - xl
    return v


def _map_xi_to_c(periodic, ns, xi, cs, a, r, cshft, xi_shift, xiv):
    """map_xi_to_c: src/pot3d.F90:3374-3470."""
    eps = 1.0e-5
    if periodic:
        xiv_p = _fold(0.0, 1.0, xiv + xi_shift)
    else:
        xiv_p = xiv
    i = 0
    while i < ns:
        if xi[i] <= xiv_p <= xi[i + 1]:
            break
        i += 1
    if i >= ns:
        raise ValueError("map_xi_to_c: xi value outside the segments")
    d = xiv_p - xi[i]
    d1 = xi[i + 1] - xi[i]
    da = a[i] * d
    da1 = a[i] * d1
    if abs(da1) < eps:
        fac = (d * (1.0 + 0.5 * da)) / (d1 * (1.0 + 0.5 * da1))
    else:
        fac = (math.exp(da) - 1.0) / (r[i] - 1.0)
    cv = cs[i] + (cs[i + 1] - cs[i]) * fac
    if periodic:
        cv = cv - cshft
        cv = _fold(cs[0], cs[ns], cv)
    return cv


def _map_c_to_xi(periodic, ns, xi, cs, a, r, cv):
    """map_c_to_xi: src/pot3d.F90:3472-3545."""
    eps = 1.0e-5
    cv_p = _fold(cs[0], cs[ns], cv) if periodic else cv
    i = 0
    while i < ns:
        if cs[i] <= cv_p <= cs[i + 1]:
            break
        i += 1
    if i >= ns:
        raise ValueError("map_c_to_xi: value outside the segments")
    d = (cv_p - cs[i]) / (cs[i + 1] - cs[i])
    da = (r[i] - 1.0) * d
    if abs(da) < eps:
        fac = d * (xi[i + 1] - xi[i])
    else:
        fac = math.log(da + 1.0) / a[i]
    return xi[i] + fac


def _filter_coord(c):
    """filter_coord (1,2,1)/4, end points fixed: src/pot3d.F90:3547-3590."""
    ff = c.copy()
    for i in range(1, len(c) - 1):
        c[i] = 0.25 * (ff[i - 1] + 2.0 * ff[i] + ff[i + 1])


def _filter_coord_periodic(xl, c):
    """filter_coord_periodic: src/pot3d.F90:3592-3650."""
    n = len(c)
    f1old = c[0]
    ff = np.empty(n + 2)
    ff[1:n + 1] = c
    ff[0] = c[n - 2] - xl
    ff[n + 1] = c[1] + xl
    for i in range(1, n + 1):
        c[i - 1] = 0.25 * (ff[i - 1] + 2.0 * ff[i] + ff[i + 1])
    f1new = c[0]
    for i in range(n):
        c[i] = c[i] - f1new + f1old


def genmesh(nc: int, c0: float, c1: float, frac=(0.0,), dratio=(1.0,), nfilt: int = 0,
            periodic: bool = False, c_shift: float = 0.0) -> np.ndarray:
    """Restatement of genmesh (src/pot3d.F90:2864-3372): nc points in [c0,c1],
    piecewise-geometric segments given by ``frac``/``dratio``, optional
    (1,2,1) filtering.  Returns float64 array c(1:nc)."""
    nseg = 30
    fr = np.zeros(nseg)
    fr[:len(frac)] = frac
    dr_ = np.zeros(nseg)
    dr_[:len(dratio)] = dratio
    if nc < 2 or c0 >= c1:
        raise ValueError("genmesh: invalid mesh size or interval")
    nf = 1
    while nf <= nseg - 1:
        if fr[nf] == 0.0:
            break
        nf += 1
    if nf == 1 and fr[0] == 0.0:
        ns = 1
        cs = [c0, c1]
        rr = [1.0]
    else:
        for i in range(1, nf):
            if fr[i] < fr[i - 1]:
                raise ValueError("genmesh: FRAC must increase monotonically")
        if periodic:
            if fr[0] < 0.0 or fr[nf - 1] > 1.0:
                raise ValueError("genmesh: FRAC must be in [0,1]")
            if fr[0] == 0.0 and fr[nf - 1] == 1.0:
                raise ValueError("genmesh: periodic FRAC cannot contain both 0 and 1")
        else:
            if fr[0] != 0.0 or fr[nf - 1] != 1.0:
                raise ValueError("genmesh: FRAC must start at 0 and end at 1")
        nr_ = nf if periodic else nf - 1
        for i in range(nr_):
            if dr_[i] <= 0.0:
                raise ValueError("genmesh: DRATIO not set or not positive")
        ns = nf - 1
        if periodic:
            if fr[0] != 0.0:
                ns += 1
            if fr[nf - 1] != 1.0:
                ns += 1
        cs = [0.0] * (ns + 1)
        rr = [0.0] * ns
        if periodic:
            if fr[0] != 0.0:
                cs[0] = c0
                for m in range(nf):
                    cs[m + 1] = c0 + (c1 - c0) * fr[m]
                if fr[nf - 1] != 1.0:
                    alpha = (1.0 - fr[nf - 1]) / (fr[0] + 1.0 - fr[nf - 1])
                    rr[0] = dr_[nr_ - 1] / (1.0 + alpha * (dr_[nr_ - 1] - 1.0))
                    for m in range(nr_):
                        rr[m + 1] = dr_[m]
                    cs[ns] = c1
                    rr[ns - 1] = 1.0 + alpha * (dr_[nr_ - 1] - 1.0)
                else:
                    rr[0] = dr_[nr_ - 1]
                    for m in range(nr_ - 1):
                        rr[m + 1] = dr_[m]
            else:
                for m in range(nf):
                    cs[m] = c0 + (c1 - c0) * fr[m]
                for m in range(nr_):
                    rr[m] = dr_[m]
                cs[ns] = c1
        else:
            for m in range(nf):
                cs[m] = c0 + (c1 - c0) * fr[m]
            for m in range(nr_):
                rr[m] = dr_[m]
    eps = 1.0e-5
    f = [0.0] * ns
    for i in range(ns):
        d = rr[i] - 1.0
        if abs(d) < eps:
            f[i] = (cs[i + 1] - cs[i]) * (1.0 + 0.5 * d)
        else:
            f[i] = (cs[i + 1] - cs[i]) * math.log(rr[i]) / d
    fac = 0.0
    for i in range(ns - 1, -1, -1):
        fac = fac / rr[i] + f[i]
    d = f[0] / fac
    xi = [0.0] * (ns + 1)
    for i in range(1, ns):
        xi[i] = xi[i - 1] + d
        if i + 1 < ns:
            d = d * f[i] / (f[i - 1] * rr[i - 1])
    xi[ns] = 1.0
    a = [math.log(rr[i]) / (xi[i + 1] - xi[i]) for i in range(ns)]
    if periodic:
        cshft = -c_shift
        xi_shift = _map_c_to_xi(True, ns, xi, cs, a, rr, cshft)
    else:
        cshft = 0.0
        xi_shift = 0.0
    dxi = 1.0 / (nc - 1.0)
    c = np.empty(nc)
    c[0] = c0
    for j in range(2, nc):
        xiv = (j - 1) * dxi
        c[j - 1] = _map_xi_to_c(periodic, ns, xi, cs, a, rr, cshft, xi_shift, xiv)
    c[nc - 1] = c1
    for _ in range(nfilt):
        if periodic:
            _filter_coord_periodic(c1 - c0, c)
        else:
            _filter_coord(c)
    return c


# --------------------------------------------------------------------------
# Global mesh (one rank: local == global)
# --------------------------------------------------------------------------
@dataclass
class Mesh:
    """The 1-D mesh arrays the solver takes (0-based copies of the
    reference's 1-based arrays).  Dimensions as in src/pot3d.F90:2530-2560:
    r,dr: nr-1;  rh,drh: nr;  t,dt,st: nt-1;  th,dth,sth: nt;  p,dp,ph,dph: np."""
    nr: int
    nt: int
    np: int
    r: np.ndarray
    dr: np.ndarray
    rh: np.ndarray
    drh: np.ndarray
    t: np.ndarray
    dt: np.ndarray
    th: np.ndarray
    dth: np.ndarray
    p: np.ndarray
    dp: np.ndarray
    ph: np.ndarray
    dph: np.ndarray
    st: np.ndarray = field(init=False)
    sth: np.ndarray = field(init=False)

    def __post_init__(self):
        self.st = np.sin(self.t)
        self.sth = np.sin(self.th)

    @property
    def N(self) -> int:
        return (self.nr - 2) * (self.nt - 2) * (self.np - 2)


def global_mesh(nr: int, nt: int, np_: int, r0: float = 1.0, r1: float = 2.5,
                rfrac=(0.0, 1.0), drratio=(1.0,), nfrmesh: int = 0,
                tfrac=(0.0,), dtratio=(1.0,), nftmesh: int = 0,
                pfrac=(0.0,), dpratio=(1.0,), nfpmesh: int = 0,
                phishift: float = 0.0) -> Mesh:
    """set_global_mesh: src/pot3d.F90:2583-2706 (axisymmetric=.false.)."""
    nrm1, ntm1, npm1 = nr - 1, nt - 1, np_ - 1
    r = genmesh(nrm1, r0, r1, rfrac, drratio, nfrmesh, False, 0.0)
    rh = np.empty(nr)
    drh = np.empty(nr)
    for i in range(2, nrm1 + 1):
        rh[i - 1] = 0.5 * (r[i - 1] + r[i - 2])
        drh[i - 1] = r[i - 1] - r[i - 2]
    rh[0] = rh[1] - drh[1]
    rh[nr - 1] = rh[nrm1 - 1] + drh[nrm1 - 1]
    drh[0] = drh[1]
    drh[nr - 1] = drh[nrm1 - 1]
    dr = np.array([rh[i] - rh[i - 1] for i in range(1, nrm1 + 1)])

    t = genmesh(ntm1, 0.0, PI, tfrac, dtratio, nftmesh, False, 0.0)
    th = np.empty(nt)
    dth = np.empty(nt)
    for j in range(2, ntm1 + 1):
        th[j - 1] = 0.5 * (t[j - 1] + t[j - 2])
        dth[j - 1] = t[j - 1] - t[j - 2]
    th[0] = th[1] - dth[1]
    th[nt - 1] = th[ntm1 - 1] + dth[ntm1 - 1]
    dth[0] = dth[1]
    dth[nt - 1] = dth[ntm1 - 1]
    dt = np.array([th[j] - th[j - 1] for j in range(1, ntm1 + 1)])

    p = np.empty(np_)
    p[:npm1] = genmesh(npm1, 0.0, 2.0 * PI, pfrac, dpratio, nfpmesh, True, phishift)
    p[np_ - 1] = p[1] + PL
    ph = np.empty(np_)
    dph = np.empty(np_)
    for k in range(2, np_ + 1):
        ph[k - 1] = 0.5 * (p[k - 1] + p[k - 2])
        dph[k - 1] = p[k - 1] - p[k - 2]
    ph[0] = ph[npm1 - 1] - PL
    dph[0] = dph[npm1 - 1]
    dp = np.empty(np_)
    for k in range(1, npm1 + 1):
        dp[k - 1] = ph[k] - ph[k - 1]
    dp[np_ - 1] = dp[1]
    dph[np_ - 1] = dph[1]
    dp[0] = dp[npm1 - 1]
    return Mesh(nr, nt, np_, r, dr, rh, drh, t, dt, th, dth, p, dp, ph, dph)


def geometric_r_ratio(per_cell_ratio: float, nr: int) -> float:
    """genmesh's drratio is the whole-segment ratio; a per-cell stretch q over
    nr-2 cells of the main mesh (nr-1 points) is drratio = q**(nr-2)."""
    return per_cell_ratio ** (nr - 2)


# --------------------------------------------------------------------------
# Br boundary map
# --------------------------------------------------------------------------
def pm_r1_for(option: str) -> float:
    """check_input: 'potential' -> +1 (dphi/dr=0 at r1); 'open'/'ss' -> -1
    (phi=0 at r1).  src/pot3d.F90:1250-1285."""
    if option == "potential":
        return 1.0
    if option in ("open", "ss"):
        return -1.0
    raise ValueError(f"invalid OPTION {option!r}; allowed: 'potential', 'open', 'ss'")


def synthetic_br_map(seed: int, n_bipoles: int = 8, ntheta: int = 181, nphi: int = 361,
                     lmax: int = 15, bg_rms: float = 5.0):
    """Seeded synthetic photospheric Br on a 1-degree (theta,phi) map.

    * ``n_bipoles`` pairs of opposite-sign Gaussian spots: sigma ~ U[2,6] deg,
      peak |B| ~ U[50,500] G, latitude ~ U[-60,60] deg, longitude ~ U[0,360);
      the two spots sit 3 sigma apart along a Joy's-law tilted axis
      (tilt = 0.5*|lat|, leading spot equatorward).
    * a background sum_{l=1..lmax} sum_m c_lm Y_lm (real harmonics) with
      c_lm ~ N(0,1)/l, scaled to ``bg_rms`` gauss rms.
    * the area-weighted mean is removed (zero net flux).
    Returns (theta scale, phi scale, map[theta,phi])."""
    rng = np.random.default_rng(seed)
    th = np.linspace(0.0, PI, ntheta)
    ph = np.linspace(0.0, 2.0 * PI, nphi)
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    xyz = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)])
    br = np.zeros_like(TH)
    deg = PI / 180.0
    for _ in range(n_bipoles):
        sig = rng.uniform(2.0, 6.0) * deg
        amp = rng.uniform(50.0, 500.0)
        lat = rng.uniform(-60.0, 60.0) * deg
        lon = rng.uniform(0.0, 360.0) * deg
        sign = rng.choice([-1.0, 1.0])
        tilt = 0.5 * abs(lat) * (1.0 if lat >= 0 else -1.0)
        half = 1.5 * sig
        for s, fac in ((+1.0, sign), (-1.0, -sign)):
            dlon = s * half * math.cos(tilt) / max(math.cos(lat), 0.2)
            dlat = -s * half * math.sin(tilt)
            c_th = PI / 2 - (lat + dlat)
            c_ph = lon + dlon
            c = np.array([math.sin(c_th) * math.cos(c_ph), math.sin(c_th) * math.sin(c_ph),
                          math.cos(c_th)])
            ang = np.arccos(np.clip(np.tensordot(c, xyz, axes=1), -1.0, 1.0))
            br += fac * amp * np.exp(-0.5 * (ang / sig) ** 2)
    # real spherical-harmonic background (unnormalised associated Legendre
    # via scipy, each term normalised numerically on the map).  Y_lm is
    # separable, P_l^m(cos th) trig(m ph), so the sum is one (theta x terms)
    # by (terms x phi) product.
    from scipy.special import lpmv
    st_ = np.sin(th)
    cols_t, cols_p, coef = [], [], []
    for l in range(1, lmax + 1):
        for m in range(0, l + 1):
            plm = lpmv(m, l, np.cos(th))
            for trig in ((np.cos,) if m == 0 else (np.cos, np.sin)):
                tp = trig(m * ph)
                mean = float(np.sum(plm * plm * st_) * np.sum(tp * tp)) / (th.size * ph.size)
                nrm = math.sqrt(mean + 1e-300)
                cols_t.append(plm)
                cols_p.append(tp)
                coef.append(rng.normal() / l / nrm)
    bg = (np.stack(cols_t, 1) * np.asarray(coef)) @ np.stack(cols_p, 0)
    w = np.sin(TH)
    bg *= bg_rms / math.sqrt(float(np.sum(bg * bg * w) / np.sum(w)))
    br += bg
    br -= float(np.sum(br * w) / np.sum(w))
    return th, ph, br
