"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product path
(pot3d_amd) never imports this module.

All arrays are float64 unless stated; packed vectors are length
N=(nr-2)(nt-2)(np-2) in the reference's r-fastest order; padded arrays are
(nr,nt,np) in Fortran order (use ``order='F'`` views); br0 is (nt,np)
Fortran order.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


class _Grid(C.Structure):
    _fields_ = [("nr", C.c_int), ("nt", C.c_int), ("np", C.c_int),
                ("r", _dp), ("dr", _dp), ("rh", _dp), ("drh", _dp),
                ("st", _dp), ("dt", _dp), ("sth", _dp), ("dth", _dp),
                ("dp", _dp), ("dph", _dp), ("pm_r1", C.c_double)]


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        G = C.POINTER(_Grid)
        sig = {
            "orc_coefs": (None, [G, _dp]),
            "orc_set_boundary_points": (None, [G, _dp, _dp, C.c_double, _dp]),
            "orc_seam": (None, [G, _dp]),
            "orc_unpack": (None, [G, _dp, _dp]),
            "orc_delsq": (None, [G, _dp, _dp, _dp]),
            "orc_ax": (None, [G, _dp, _dp, _dp, _dp]),
            "orc_rhs": (None, [G, _dp, _dp, _dp, _dp, _dp]),
            "orc_fill_phi": (None, [G, _dp, _dp, _dp, _dp]),
            "orc_diag_prec": (None, [G, _dp, _fp]),
            "orc_getM": (C.c_int64, [G, _ip]),
            "orc_diacsr": (None, [G, _dp, _ip, _fp, _ip, _ip]),
            "orc_ilu0": (C.c_int, [C.c_int64, _fp, _ip, _ip, _ip]),
            "orc_lu2luopt": (None, [C.c_int64, _fp, _fp, _ip, _ip, _ip, _ip, _ip, _ip]),
            "orc_lusol": (None, [C.c_int64, _fp, _fp, _ip, _ip, _ip, _fp]),
            "orc_dilu_pivots": (None, [G, _dp, _dp]),
            "orc_dilu_apply": (None, [G, _dp, _fp, _fp]),
            "orc_dilu_pivots_blk": (None, [G, _dp, C.c_int, C.c_int, C.c_int, _dp]),
            "orc_dilu_apply_blk": (None, [G, _dp, _fp, C.c_int, C.c_int, C.c_int, _fp]),
            "orc_prec_apply": (C.c_int, [G, _dp, C.c_int, _dp]),
            "orc_cgsolve": (C.c_int, [G, _dp, C.c_int, C.c_double, C.c_int, _dp, _dp,
                                      _ip, _dp, _dp, C.c_int]),
            "orc_getb": (None, [G, _dp, _dp, _dp, _dp]),
            "orc_energy": (None, [G, _dp, _dp, _dp, _dp]),
            "orc_phi_energy": (None, [G, _dp, _dp]),
            "orc_dot": (C.c_double, [_dp, _dp, C.c_int64]),
            "orc_cgsolve_ranks": (C.c_int, [G, _dp, _ip, _ip, _ip, _ip, C.c_int, C.c_double,
                                            C.c_int, _dp, _dp, _ip, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def rank_topology(nr: int, nt: int, np_: int, nproc: int):
    """The processor topology (nproc_r, nproc_t, nproc_p) the reference picks
    for nproc MPI ranks when none is prescribed (set_proc_topology,
    src/pot3d.F90:1386-1530): among the factorisations fr*ft*fp = nproc, in
    the order fp outer, ft, fr inner, the first one minimising
    (nr/fr - nt/ft)^2 + (nt/ft - np/fp)^2 + (np/fp - nr/fr)^2 with at least 4
    points per rank in r and theta and 3 in phi."""
    facs = [f for f in range(1, nproc + 1) if nproc % f == 0]
    best, best_pen = None, None
    for fp in facs:
        for ft in facs:
            for fr in facs:
                if fr * ft * fp != nproc:
                    continue
                a, b, c = nr / fr, nt / ft, np_ / fp
                pen = (a - b) ** 2 + (b - c) ** 2 + (c - a) ** 2
                if a < 4 or b < 4 or c < 3:
                    continue
                if best_pen is None or pen < best_pen:
                    best, best_pen = (fr, ft, fp), pen
    if best is None:
        raise ValueError(f"no rank topology for {nproc} ranks on {nr}x{nt}x{np_}")
    return best


def decompose_dimension(nx: int, nproc: int):
    """Interior points per rank along one axis (decompose_dimension,
    src/pot3d.F90:2116-2167): nx-2 points as equal as possible, the first
    ranks taking one more."""
    m = nx - 2
    av, rem = divmod(m, nproc)
    return [av + 1 if q < rem else av for q in range(nproc)]


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _f(a):
    return a.ctypes.data_as(_fp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Oracle:
    """The oracle bound to one mesh (a p3d_driver.inputs.Mesh) and pm_r1."""

    def __init__(self, mesh, pm_r1: float):
        self.mesh = mesh
        self.nr, self.nt, self.np = mesh.nr, mesh.nt, mesh.np
        self.N = (self.nr - 2) * (self.nt - 2) * (self.np - 2)
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                      (mesh.r, mesh.dr, mesh.rh, mesh.drh, mesh.st, mesh.dt, mesh.sth, mesh.dth,
                       mesh.dp, mesh.dph)]
        k = self._keep
        self.g = _Grid(self.nr, self.nt, self.np, _d(k[0]), _d(k[1]), _d(k[2]), _d(k[3]),
                       _d(k[4]), _d(k[5]), _d(k[6]), _d(k[7]), _d(k[8]), _d(k[9]), float(pm_r1))
        self._a = None

    @property
    def gp(self):
        return C.byref(self.g)

    def padded(self):
        return np.zeros(self.nr * self.nt * self.np)

    # -- operator ---------------------------------------------------------
    def coefs(self) -> np.ndarray:
        """a(2:nrm1,2:ntm1,2:npm1,7) as an array of shape (7, N)."""
        if self._a is None:
            a = np.empty(7 * self.N)
            lib().orc_coefs(self.gp, _d(a))
            self._a = a
        return self._a.reshape(7, self.N)

    def ax(self, x: np.ndarray) -> np.ndarray:
        a = self.coefs()
        y = np.empty(self.N)
        xa = self.padded()
        lib().orc_ax(self.gp, _d(a), _d(np.ascontiguousarray(x, float)), _d(y), _d(xa))
        return y

    def rhs(self, br0: np.ndarray, br1: np.ndarray | None = None) -> np.ndarray:
        a = self.coefs()
        b = np.empty(self.N)
        xa = self.padded()
        b0 = np.asfortranarray(br0, dtype=np.float64)
        b1 = np.asfortranarray(br1, dtype=np.float64) if br1 is not None else None
        lib().orc_rhs(self.gp, _d(a), b0.ctypes.data_as(_dp),
                      b1.ctypes.data_as(_dp) if b1 is not None else None, _d(b), _d(xa))
        return b

    def set_boundary_points(self, xpad: np.ndarray, br0, vmask: float, br1=None):
        b0 = np.asfortranarray(br0, dtype=np.float64) if br0 is not None else None
        b1 = np.asfortranarray(br1, dtype=np.float64) if br1 is not None else None
        lib().orc_set_boundary_points(self.gp, b0.ctypes.data_as(_dp) if b0 is not None else None,
                                      b1.ctypes.data_as(_dp) if b1 is not None else None,
                                      float(vmask), _d(xpad))

    def fill_phi(self, x: np.ndarray, br0: np.ndarray, br1=None) -> np.ndarray:
        """phi(nr,nt,np) Fortran-ordered flat array."""
        phi = self.padded()
        b0 = np.asfortranarray(br0, dtype=np.float64)
        b1 = np.asfortranarray(br1, dtype=np.float64) if br1 is not None else None
        lib().orc_fill_phi(self.gp, _d(np.ascontiguousarray(x, float)), b0.ctypes.data_as(_dp),
                           b1.ctypes.data_as(_dp) if b1 is not None else None, _d(phi))
        return phi

    def energy(self, phi: np.ndarray) -> np.ndarray:
        w = np.empty(3)
        lib().orc_phi_energy(self.gp, _d(np.ascontiguousarray(phi, float)), _d(w))
        return w

    def getb(self, phi: np.ndarray):
        nr, nt, np_ = self.nr, self.nt, self.np
        br = np.empty((nr - 1) * nt * np_)
        bt = np.empty(nr * (nt - 1) * np_)
        bp = np.empty(nr * nt * (np_ - 1))
        lib().orc_getb(self.gp, _d(np.ascontiguousarray(phi, float)), _d(br), _d(bt), _d(bp))
        return br, bt, bp

    # -- preconditioners --------------------------------------------------
    def diag_prec(self) -> np.ndarray:
        ai = np.empty(self.N, dtype=np.float32)
        lib().orc_diag_prec(self.gp, _d(self.coefs().ravel()), _f(ai))
        return ai

    def csr(self):
        """getM + diacsr: (IA, JA, Acsr, Adptr), 1-based; IA has N+1 entries,
        Acsr/JA have M = nnz+1 entries (the last one unused)."""
        IA = np.empty(self.N + 1, dtype=np.int32)
        M = lib().orc_getM(self.gp, _i(IA))
        Av = np.zeros(M, dtype=np.float32)
        JA = np.zeros(M, dtype=np.int32)
        dptr = np.empty(self.N, dtype=np.int32)
        lib().orc_diacsr(self.gp, _d(self.coefs().ravel()), _i(IA), _f(Av), _i(JA), _i(dptr))
        return IA, JA, Av, dptr

    def ilu0(self):
        """Saad ILU(0) of the diacsr matrix: returns (IA, JA, LU, dptr, icode)."""
        IA, JA, Av, dptr = self.csr()
        icode = lib().orc_ilu0(self.N, _f(Av), _i(JA), _i(IA), _i(dptr))
        return IA, JA, Av, dptr, icode

    def lusol(self, rhs32: np.ndarray) -> np.ndarray:
        IA, JA, Av, dptr, _ = self.ilu0()
        M = len(Av)
        LU = np.zeros(M, dtype=np.float32)
        LUJA = np.zeros(M, dtype=np.int32)
        N1 = np.empty(self.N, dtype=np.int32)
        N2 = np.empty(self.N, dtype=np.int32)
        lib().orc_lu2luopt(self.N, _f(LU), _f(Av), _i(IA), _i(JA), _i(LUJA), _i(dptr), _i(N1), _i(N2))
        lud = (1.0 / Av[dptr - 1].astype(np.float64)).astype(np.float32)
        x = np.array(rhs32, dtype=np.float32)
        lib().orc_lusol(self.N, _f(x), _f(LU), _i(LUJA), _i(N1), _i(N2), _f(lud))
        return x

    def dilu_pivots(self, blocks=None) -> np.ndarray:
        """D-ILU pivots; blocks=(bi, bj, bk) block extents in points selects
        the block-Jacobi variant (couplings between blocks dropped)."""
        d = np.empty(self.N)
        if blocks is None:
            lib().orc_dilu_pivots(self.gp, _d(self.coefs().ravel()), _d(d))
        else:
            bi, bj, bk = (int(v) for v in blocks)
            lib().orc_dilu_pivots_blk(self.gp, _d(self.coefs().ravel()), bi, bj, bk, _d(d))
        return d

    def dilu_dinv(self, blocks=None) -> np.ndarray:
        return (1.0 / self.dilu_pivots(blocks)).astype(np.float32)

    def dilu_apply(self, r32: np.ndarray, dinv: np.ndarray | None = None,
                   blocks=None) -> np.ndarray:
        if dinv is None:
            dinv = self.dilu_dinv(blocks)
        x = np.array(r32, dtype=np.float32)
        if blocks is None:
            lib().orc_dilu_apply(self.gp, _d(self.coefs().ravel()),
                                 _f(np.ascontiguousarray(dinv)), _f(x))
        else:
            bi, bj, bk = (int(v) for v in blocks)
            lib().orc_dilu_apply_blk(self.gp, _d(self.coefs().ravel()),
                                     _f(np.ascontiguousarray(dinv)), bi, bj, bk, _f(x))
        return x

    def prec_apply(self, x: np.ndarray, ifprec: int) -> np.ndarray:
        y = np.array(x, dtype=np.float64)
        lib().orc_prec_apply(self.gp, _d(self.coefs().ravel()), int(ifprec), _d(y))
        return y

    # -- solve --------------------------------------------------------------
    def cgsolve_ranks(self, rhs: np.ndarray, nprocs, epscg: float = 1e-9, stop_mode: int = 1,
                      ncgmax: int = 1000000, x0: np.ndarray | None = None):
        """The reference's CPU ifprec=2 solve on nprocs=(fr, ft, fp) MPI
        ranks, one OpenMP thread per rank: per-rank diacsr/ilu0/lusol
        (block-Jacobi ILU0), global operator and dots.  stop_mode 1 stops on
        ||r||2/||b||2, 0 on the reference's epsn.  Returns dict(x, r, ierr,
        ncg, epsn, t_setup, t_iter, blocks)."""
        nb = np.array(nprocs, dtype=np.int32)
        los = []
        for n, f in zip((self.nr, self.nt, self.np), nprocs):
            los.append(np.concatenate([[0], np.cumsum(decompose_dimension(n, int(f)))])
                       .astype(np.int32))
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        r = np.array(rhs, dtype=np.float64)
        ncg = C.c_int(0)
        epsn = C.c_double(0.0)
        t = np.zeros(2)
        ierr = lib().orc_cgsolve_ranks(self.gp, _d(self.coefs().ravel()), _i(nb), _i(los[0]),
                                       _i(los[1]), _i(los[2]), int(stop_mode), float(epscg),
                                       int(ncgmax), _d(x), _d(r), C.byref(ncg), C.byref(epsn),
                                       _d(t))
        return dict(x=x, r=r, ierr=ierr, ncg=ncg.value, epsn=epsn.value, t_setup=float(t[0]),
                    t_iter=float(t[1]), blocks=[list(map(int, lo)) for lo in los])


    def cgsolve(self, rhs: np.ndarray, ifprec: int = 1, epscg: float = 1e-9,
                ncgmax: int = 1000000, x0: np.ndarray | None = None, nhist: int = 0):
        """Returns dict(x, r, ierr, ncg, epsn, hist)."""
        x = np.zeros(self.N) if x0 is None else np.array(x0, dtype=np.float64)
        r = np.array(rhs, dtype=np.float64)
        ncg = C.c_int(0)
        epsn = C.c_double(0.0)
        hist = np.zeros(max(nhist, 1))
        ierr = lib().orc_cgsolve(self.gp, _d(self.coefs().ravel()), int(ifprec), float(epscg),
                                 int(ncgmax), _d(x), _d(r), C.byref(ncg), C.byref(epsn),
                                 _d(hist) if nhist > 0 else None, int(nhist))
        return dict(x=x, r=r, ierr=ierr, ncg=ncg.value, epsn=epsn.value,
                    hist=hist[:min(nhist, ncg.value + 1)] if nhist > 0 else None)
