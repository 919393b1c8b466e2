"""Host-side mirror of POT3D's solve path over libpot3d_hip.so.

Names follow the reference (src/pot3d.F90): ``readbr`` (:6751-6829, with
set_flux :3806-3833), ``potfld`` (:3866-3976), ``solve`` (:4169-4229,
returning ierr), ``ax`` (:5102-5150), ``prec_inv`` (:5152-5225), ``getb`` /
``magnetic_energy`` (:6187-6321).  Vectors are
numpy float64 arrays in the reference's packed layout (N=(nr-2)(nt-2)(np-2),
r fastest); phi is the padded (nr,nt,np) Fortran-ordered array, returned
flat.  Every call runs on the GPU through the C ABI; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def readbr(mesh, tn, pn, bn, balance_flux: bool = True, abs_br: bool = False,
           pl: float = 2.0 * 3.1415926535897932 - 0.0):
    """The driver's readbr after the HDF5 read, and set_flux, on the GPU
    (p3d_readbr): the map bn(len(tn), len(pn)) on its theta/phi scales
    interpolated onto the mesh, seam, m=0 poles, flux balance (balance_flux)
    and |Br| (abs_br, option 'open').  Returns (br0 as an (nt, np) C array
    holding the Fortran br0(j,k) at [j-1, k-1], dict(fluxp, fluxm, area,
    br00err))."""
    tn = _f64(tn)
    pn = _f64(pn)
    bnf = np.asfortranarray(bn, dtype=np.float64)
    assert bnf.shape == (tn.size, pn.size)
    nt, np_ = mesh.nt, mesh.np
    br0 = np.empty((np_, nt))  # Fortran (nt, np) = C (np, nt)
    flux = np.zeros(4)
    arrs = [_f64(a) for a in (mesh.th, mesh.ph, mesh.sth, mesh.dth, mesh.dph)]
    check(_lib.load().p3d_readbr(nt, np_, *(_ptr(a) for a in arrs), float(pl), tn.size, pn.size,
                                 _ptr(tn), _ptr(pn), bnf.ctypes.data_as(_dp), int(balance_flux),
                                 int(abs_br), _ptr(br0), _ptr(flux)))
    return np.ascontiguousarray(br0.T), dict(zip(("fluxp", "fluxm", "area", "br00err"),
                                                 map(float, flux)))


class Pot3dSolver:
    """One rank's solver: the matrix-free operator for the given mesh,
    pm_r1 (+1 'potential', -1 'open'/'ss') and ifprec (0 none, 1 diagonal,
    2 ILU0).  ILU0 is factorised on the GPU at construction."""

    def __init__(self, mesh, pm_r1: float, ifprec: int = 1):
        L = _lib.load()
        self.mesh = mesh
        self.nr, self.nt, self.np = mesh.nr, mesh.nt, mesh.np
        self.N = (self.nr - 2) * (self.nt - 2) * (self.np - 2)
        self.ifprec = ifprec
        arrs = [_f64(v) for v in (mesh.r, mesh.dr, mesh.rh, mesh.drh, mesh.st, mesh.dt,
                                   mesh.sth, mesh.dth, mesh.dp, mesh.dph)]
        h = C.c_void_p()
        check(L.p3d_create(C.byref(h), self.nr, self.nt, self.np, *[_ptr(a) for a in arrs],
                           float(pm_r1), int(ifprec)))
        self._h = h
        self._L = L

    # -- lifecycle ----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.p3d_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_option(self, key: int, value: float):
        check(self._L.p3d_set_option(self._h, int(key), float(value)))

    # -- boundary data --------------------------------------------------------
    def set_br0(self, br0: np.ndarray):
        """br0(nt,np): C array holding Fortran br0(j,k) at [j-1,k-1]."""
        b = np.asfortranarray(br0, dtype=np.float64)
        assert b.shape == (self.nt, self.np)
        check(self._L.p3d_set_br0(self._h, b.ctypes.data_as(_dp)))

    def set_br1(self, br1: np.ndarray | None):
        if br1 is None:
            check(self._L.p3d_set_br1(self._h, None))
            return
        b = np.asfortranarray(br1, dtype=np.float64)
        check(self._L.p3d_set_br1(self._h, b.ctypes.data_as(_dp)))

    def rhs(self) -> np.ndarray:
        out = np.empty(self.N)
        check(self._L.p3d_rhs(self._h, _ptr(out)))
        return out

    # -- operations -------------------------------------------------------------
    def ax(self, x: np.ndarray) -> np.ndarray:
        y = np.empty(self.N)
        check(self._L.p3d_apply_A(self._h, _ptr(_f64(x)), _ptr(y)))
        return y

    def prec_inv(self, x: np.ndarray) -> np.ndarray:
        y = np.array(x, dtype=np.float64)
        check(self._L.p3d_apply_prec(self._h, _ptr(y)))
        return y

    def ilu_dinv(self) -> np.ndarray:
        d = np.empty(self.N, dtype=np.float32)
        check(self._L.p3d_ilu_dinv(self._h, d.ctypes.data_as(_fp)))
        return d

    def solve(self, x: np.ndarray, rhs: np.ndarray | None = None, epscg: float = 1e-9,
              ncgmax: int = 1000000, ncghist: int = 0):
        """The reference's solve(x,rhs,N,ierr): x (guess in, solution out) and
        rhs (in; residual out) are updated in place.  Returns
        (ierr, ncg, epsn) with ierr 0 converged / 1 ncgmax reached."""
        assert x.dtype == np.float64 and x.flags.c_contiguous and x.size == self.N
        if rhs is not None:
            assert rhs.dtype == np.float64 and rhs.flags.c_contiguous and rhs.size == self.N
        ncg = C.c_int(0)
        epsn = C.c_double(0.0)
        rc = self._L.p3d_solve(self._h, _ptr(x), _ptr(rhs), float(epscg), int(ncgmax),
                               int(ncghist), C.byref(ncg), C.byref(epsn))
        if rc == _lib.P3D_ENOCONV:
            return 1, ncg.value, epsn.value
        check(rc)
        return 0, ncg.value, epsn.value

    def solve_dev(self, x_ptr: int, rhs_ptr: int | None, epscg: float = 1e-9,
                  ncgmax: int = 1000000):
        """Device-resident solve: x_ptr / rhs_ptr are device addresses of N
        doubles (e.g. torch tensor data_ptr()).  Returns (ierr, ncg, epsn)."""
        ncg = C.c_int(0)
        epsn = C.c_double(0.0)
        rc = self._L.p3d_solve_dev(self._h, C.c_void_p(x_ptr),
                                   C.c_void_p(rhs_ptr) if rhs_ptr else None, float(epscg),
                                   int(ncgmax), C.byref(ncg), C.byref(epsn))
        if rc == _lib.P3D_ENOCONV:
            return 1, ncg.value, epsn.value
        check(rc)
        return 0, ncg.value, epsn.value

    def stats(self) -> dict:
        st = _lib.Stats()
        check(self._L.p3d_get_stats(self._h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in _lib.Stats._fields_ if f != "pad_"}

    def history(self, n: int) -> np.ndarray:
        h = np.zeros(max(n, 1))
        m = self._L.p3d_history(self._h, _ptr(h), int(n))
        if m < 0:
            check(m)
        return h[:m]

    def time_kernel(self, which: int, reps: int = 20) -> float:
        ms = C.c_double(0.0)
        check(self._L.p3d_time_kernel(self._h, int(which), int(reps), C.byref(ms)))
        return ms.value

    # -- post-solve ---------------------------------------------------------------
    def fill_phi(self, x: np.ndarray | None = None) -> np.ndarray:
        phi = np.empty(self.nr * self.nt * self.np)
        check(self._L.p3d_fill_phi(self._h, _ptr(_f64(x)) if x is not None else None, _ptr(phi)))
        return phi

    def magnetic_energy(self, phi: np.ndarray | None = None) -> np.ndarray:
        w = np.empty(3)
        check(self._L.p3d_energy(self._h, _ptr(_f64(phi)) if phi is not None else None, _ptr(w),
                                 None, None, None))
        return w

    def getb(self, phi: np.ndarray | None = None):
        nr, nt, np_ = self.nr, self.nt, self.np
        br = np.empty((nr - 1) * nt * np_)
        bt = np.empty(nr * (nt - 1) * np_)
        bp = np.empty(nr * nt * (np_ - 1))
        w = np.empty(3)
        check(self._L.p3d_energy(self._h, _ptr(_f64(phi)) if phi is not None else None, _ptr(w),
                                 _ptr(br), _ptr(bt), _ptr(bp)))
        return br, bt, bp


@dataclass
class PotfldResult:
    x: np.ndarray
    phi: np.ndarray
    energy: np.ndarray
    ierr: int
    ncg: int
    epsn: float
    stats: dict


def potfld(workload, ifprec: int | None = None, epscg: float | None = None, mixed: bool = False,
           x0: np.ndarray | None = None, stop: int = 0, ncgmax: int = 1000000,
           solver: Pot3dSolver | None = None) -> PotfldResult:
    """The reference's potfld (src/pot3d.F90:3866-3976) for a Workload:
    boundary-trick rhs, PCG solve from x0 (zero by default, :3924), phi with
    boundary values and seam, then the energy diagnostic."""
    ip = workload.ifprec if ifprec is None else ifprec
    own = solver is None
    s = solver or Pot3dSolver(workload.mesh, workload.pm_r1, ip)
    try:
        s.set_br0(workload.br0)
        s.set_option(_lib.OPT_MIXED, 1.0 if mixed else 0.0)
        s.set_option(_lib.OPT_STOP, float(stop))
        x = np.zeros(s.N) if x0 is None else np.array(x0, dtype=np.float64)
        ierr, ncg, epsn = s.solve(x, None, epscg or workload.epscg, ncgmax)
        phi = s.fill_phi()
        w = s.magnetic_energy()
        return PotfldResult(x, phi, w, ierr, ncg, epsn, s.stats())
    finally:
        if own:
            s.close()
