"""Learned warm start for the PCG solve (libpot3d_hip.so, p3d_nn_*).

The reference has no learned part: its initial guess is zero
(src/pot3d.F90:3922-3925).  Here a BF16 3-D convolutional encoder-decoder
maps the r0 boundary field br0 (and the radial grid) to the potential on the
solver's interior points.  It is trained on the GPU by the library's own
forward, backward and Adam kernels, on (br0, x) pairs that the plain PCG
solve produces.  Its output is only ever the initial guess of the same PCG
solve to the same tolerance; the solver replaces it by zero when it does
not lower the residual below ||b|| (P3D_OPT_GUESS_CHECK).

Networks are trained on a coarse grid and their prediction is carried to a
finer grid by trilinear prolongation of the padded coarse potential (with
its boundary values and seam, p3d_fill_phi) on the GPU.

The area-weighted mean m of br0 (the monopole, large for the 'open'
problem's |Br| maps) is a global quantity that the network's receptive field
does not span.  With set_monopole the network sees br0 - m and learns
x - m x_m, where x_m is the PCG solution for br0 = 1 on the same grid: the
problem is linear in br0, so the prediction m x_m + net(br0 - m) carries
the monopole exactly.

Every computation runs through the C ABI; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import check

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _fortran_br(br0, nt, np_):
    b = np.asfortranarray(br0, dtype=np.float64)
    assert b.shape == (nt, np_)
    return b


class WarmStartNet:
    """The network for the interior grid of ``mesh`` (nr, nt, np)."""

    def __init__(self, mesh, seed: int = 0):
        L = _lib.load()
        self.mesh = mesh
        self.nr, self.nt, self.np = mesh.nr, mesh.nt, mesh.np
        self.N = (self.nr - 2) * (self.nt - 2) * (self.np - 2)
        rh = np.ascontiguousarray(mesh.rh, dtype=np.float64)
        h = C.c_void_p()
        check(L.p3d_nn_create(C.byref(h), self.nr, self.nt, self.np, rh.ctypes.data_as(_dp),
                              int(seed) & 0xFFFFFFFF))
        self._h = h
        self._L = L
        self.nparams = int(L.p3d_nn_num_params(h))
        self.nsamples = 0
        self.bscale = 1.0
        self.phiscale = 1.0
        self._mono = None  # (x_m, area weights of br0's interior) with the monopole split

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self._L.p3d_nn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, bscale: float, phiscale: float):
        check(self._L.p3d_nn_set_scales(self._h, float(bscale), float(phiscale)))
        self.bscale, self.phiscale = float(bscale), float(phiscale)

    def params(self) -> np.ndarray:
        p = np.empty(self.nparams, np.float32)
        check(self._L.p3d_nn_get_params(self._h, p.ctypes.data_as(_fp)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        check(self._L.p3d_nn_set_params(self._h, p.ctypes.data_as(_fp)))

    def set_monopole(self, x_m: np.ndarray, area: np.ndarray):
        """Split br0's area-weighted mean off (see the module docstring):
        x_m the packed solution for br0 = 1, area the (nt-2, np-2) weights of
        br0's interior points (area_weights)."""
        x_m = np.ascontiguousarray(x_m, dtype=np.float64)
        area = np.asarray(area, dtype=np.float64)
        assert x_m.size == self.N and area.shape == (self.nt - 2, self.np - 2)
        self._mono = (x_m, area)

    @property
    def monopole(self) -> bool:
        return self._mono is not None

    def mean_br(self, br0: np.ndarray) -> float:
        """Area-weighted mean of br0 over its interior points."""
        area = self._mono[1]
        return float(np.sum(area * np.asarray(br0)[1:-1, 1:-1]) / np.sum(area))

    def split(self, br0: np.ndarray, x: np.ndarray | None = None):
        """(br0 - m, x - m x_m) with the monopole split, else the inputs."""
        if self._mono is None:
            return br0, x
        m = self.mean_br(br0)
        return (np.asarray(br0) - m,
                None if x is None else np.asarray(x) - m * self._mono[0])

    def add_sample(self, br0: np.ndarray, x: np.ndarray):
        br0, x = self.split(br0, x)
        b = _fortran_br(br0, self.nt, self.np)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        assert xx.size == self.N
        check(self._L.p3d_nn_add_sample(self._h, b.ctypes.data_as(_dp), xx.ctypes.data_as(_dp)))
        self.nsamples += 1

    def loss_grad(self, br0: np.ndarray, x: np.ndarray, grads: bool = True):
        br0, x = self.split(br0, x)
        b = _fortran_br(br0, self.nt, self.np)
        xx = np.ascontiguousarray(x, dtype=np.float64)
        loss = C.c_double(0.0)
        g = np.empty(self.nparams, np.float32) if grads else None
        check(self._L.p3d_nn_loss_grad(self._h, b.ctypes.data_as(_dp), xx.ctypes.data_as(_dp),
                                       C.byref(loss), g.ctypes.data_as(_fp) if grads else None))
        return loss.value, g

    def set_augment(self, phi_roll: bool, theta_flip: bool, sign_flip: bool):
        check(self._L.p3d_nn_set_augment(self._h, int(phi_roll), int(theta_flip),
                                         int(sign_flip)))

    def save(self, path: str, **meta):
        """Parameters, scales and the grid (plus ``meta``) in an .npz file;
        with the monopole split, monopole = 1 (x_m is recomputed on load)."""
        if self._mono is not None:
            meta = dict(meta, monopole=1)
        np.savez_compressed(path, params=self.params(), bscale=self.bscale,
                            phiscale=self.phiscale, grid=np.array([self.nr, self.nt, self.np]),
                            **meta)

    @classmethod
    def load(cls, path: str, mesh, pm_r1: float | None = None):
        """A network for ``mesh`` with the parameters and scales of ``path``
        (written by save); returns (net, the file's other entries).  A network
        trained with the monopole split needs the problem's pm_r1 (the outer
        boundary condition) to solve for x_m."""
        z = np.load(path, allow_pickle=False)
        grid = tuple(int(v) for v in z["grid"])
        if grid != (mesh.nr, mesh.nt, mesh.np):
            raise ValueError(f"{path} holds a network for {grid}, not "
                             f"{(mesh.nr, mesh.nt, mesh.np)}")
        net = cls(mesh, seed=0)
        net.set_params(z["params"])
        net.set_scales(float(z["bscale"]), float(z["phiscale"]))
        if "monopole" in z.files and int(z["monopole"]):
            if pm_r1 is None:
                net.close()
                raise ValueError(f"{path} splits the monopole off: pass the problem's pm_r1")
            net.set_monopole(monopole_response(mesh, pm_r1), area_weights(mesh))
        meta = {k: z[k] for k in z.files if k not in ("params", "bscale", "phiscale", "grid")}
        return net, meta

    def train(self, nsteps: int, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(nsteps, 1), np.float32)
        check(self._L.p3d_nn_train(self._h, int(nsteps), float(lr), int(seed) & 0xFFFFFFFF,
                                   losses.ctypes.data_as(_fp)))
        return losses[:nsteps]

    def predict(self, br0: np.ndarray) -> np.ndarray:
        bs, _ = self.split(br0)
        b = _fortran_br(bs, self.nt, self.np)
        x = np.empty(self.N)
        check(self._L.p3d_nn_predict(self._h, b.ctypes.data_as(_dp), x.ctypes.data_as(_dp), None))
        if self._mono is not None:
            x += self.mean_br(br0) * self._mono[0]
        return x

    def predict_dev(self, br0: np.ndarray, x_ptr: int):
        if self._mono is not None:
            raise ValueError("predict_dev has no monopole split: use predict")
        b = _fortran_br(br0, self.nt, self.np)
        check(self._L.p3d_nn_predict(self._h, b.ctypes.data_as(_dp), None, C.c_void_p(x_ptr)))


def area_weights(mesh) -> np.ndarray:
    """Cell areas sin(theta) dtheta dphi of br0's interior points (nt-2, np-2):
    the weights of the driver's flux sums."""
    return (np.asarray(mesh.sth)[1:-1, None] * np.asarray(mesh.dth)[1:-1, None]
            * np.asarray(mesh.dph)[None, 1:-1])


def monopole_response(mesh, pm_r1: float, epscg: float = 1e-12) -> np.ndarray:
    """x_m: the packed FP64 PCG solution for br0 = 1 on ``mesh`` (ILU0,
    ||b-Ax||2/||b||2 <= epscg)."""
    from .solver import Pot3dSolver
    with Pot3dSolver(mesh, pm_r1, 2) as s:
        s.set_br0(np.ones((mesh.nt, mesh.np)))
        s.set_option(_lib.OPT_STOP, 1)
        x = np.zeros(s.N)
        ierr, ncg, _ = s.solve(x, None, epscg, ncgmax=100000)
        if ierr != 0:
            raise RuntimeError(f"monopole response did not converge ({ncg} iterations)")
    return x


# -- prolongation ------------------------------------------------------------
def _axis_table(coarse: np.ndarray, fine: np.ndarray):
    """Lower bracketing index into the coarse half mesh and the weight of the
    upper point, for each fine coordinate."""
    i0 = np.searchsorted(coarse, fine, side="right") - 1
    i0 = np.clip(i0, 0, coarse.size - 2)
    w = (fine - coarse[i0]) / (coarse[i0 + 1] - coarse[i0])
    return np.ascontiguousarray(i0, np.int32), np.ascontiguousarray(w, np.float64)


@dataclass
class ProlongTables:
    nrc: int
    ntc: int
    npc: int
    af: int
    bf: int
    cf: int
    arrays: list = field(default_factory=list)  # ir, wr, it, wt, ip, wp

    def args(self):
        ir, wr, it, wt, ip, wp = self.arrays
        return (ir.ctypes.data_as(_ip), wr.ctypes.data_as(_dp), it.ctypes.data_as(_ip),
                wt.ctypes.data_as(_dp), ip.ctypes.data_as(_ip), wp.ctypes.data_as(_dp))


def prolong_tables(cmesh, fmesh) -> ProlongTables:
    """Trilinear tables from the padded half mesh (rh, th, ph) of ``cmesh``
    to the interior half-mesh points of ``fmesh``; both meshes span the same
    domain, so every fine interior point lies inside the coarse padded box."""
    ir, wr = _axis_table(np.asarray(cmesh.rh), np.asarray(fmesh.rh)[1:-1])
    it, wt = _axis_table(np.asarray(cmesh.th), np.asarray(fmesh.th)[1:-1])
    ip, wp = _axis_table(np.asarray(cmesh.ph), np.asarray(fmesh.ph)[1:-1])
    return ProlongTables(cmesh.nr, cmesh.nt, cmesh.np, fmesh.nr - 2, fmesh.nt - 2,
                         fmesh.np - 2, [ir, wr, it, wt, ip, wp])


def prolong(tab: ProlongTables, phic: np.ndarray) -> np.ndarray:
    """Fine packed x from the coarse padded phi (host arrays)."""
    phic = np.ascontiguousarray(phic, dtype=np.float64)
    assert phic.size == tab.nrc * tab.ntc * tab.npc
    xf = np.empty(tab.af * tab.bf * tab.cf)
    check(_lib.load().p3d_nn_prolong(tab.nrc, tab.ntc, tab.npc, phic.ctypes.data_as(_dp),
                                     tab.af, tab.bf, tab.cf, *tab.args(),
                                     xf.ctypes.data_as(_dp)))
    return xf


def prolong_dev(tab: ProlongTables, phic_ptr: int, xf_ptr: int):
    """Same on device addresses."""
    check(_lib.load().p3d_nn_prolong_dev(tab.nrc, tab.ntc, tab.npc, C.c_void_p(phic_ptr),
                                         tab.af, tab.bf, tab.cf, *tab.args(),
                                         C.c_void_p(xf_ptr)))


def prolong_dev_noalloc(tab: ProlongTables, phic_ptr: int, xf_ptr: int):
    """prolong_dev through the entry point that allocates nothing (tables
    staged in the library's fixed device array; af+bf+cf <= 8192)."""
    check(_lib.load().p3d_nn_prolong_dev_noalloc(tab.nrc, tab.ntc, tab.npc, C.c_void_p(phic_ptr),
                                                 tab.af, tab.bf, tab.cf, *tab.args(),
                                                 C.c_void_p(xf_ptr)))


# -- training data -------------------------------------------------------------
def _rms_interior(br0):
    b = np.asarray(br0)[1:-1, 1:-1]
    return float(np.sqrt(np.mean(b * b)))


def fit_scales(brs, xs):
    """Per-sample normalisation (bscale = 0): each br0 is scaled by its own
    interior rms s; the output scale is the mean ratio rms(x) / s over the
    training pairs, so targets have about unit rms."""
    ratios = []
    for b, x in zip(brs, xs):
        s = _rms_interior(b)
        if s > 0:
            ratios.append(float(np.sqrt(np.mean(np.square(x)))) / s)
    ratio = float(np.mean(ratios)) if ratios else 1.0
    return 0.0, (ratio if ratio > 0 else 1.0)


def generate_pairs(solver, brs, epscg: float = 1e-9, ncgmax: int = 100000):
    """Solves the plain PCG for each br0 (solver: a Pot3dSolver); returns the
    solutions and total iterations."""
    xs, iters = [], 0
    for b in brs:
        solver.set_br0(b)
        x = np.zeros(solver.N)
        ierr, ncg, _ = solver.solve(x, None, epscg, ncgmax)
        if ierr != 0:
            raise RuntimeError(f"training solve did not converge ({ncg} iterations)")
        xs.append(x)
        iters += ncg
    return xs, iters


def train_network(net: WarmStartNet, brs, xs, epochs: int = 4, lr: float = 1e-3,
                  seed: int = 0, set_scales: bool = True):
    """Stores the pairs on the device and trains epochs x len(brs) Adam
    steps, the learning rate stepping down to lr/4 and lr/16 for the last
    40% and 15% of the steps; returns the per-step losses and the training
    time."""
    if set_scales:  # on what the network sees (after the monopole split, if any)
        net.set_scales(*fit_scales((net.split(b)[0] for b in brs),
                                   (net.split(b, x)[1] for b, x in zip(brs, xs))))
    for b, x in zip(brs, xs):
        net.add_sample(b, x)
    total = epochs * len(brs)
    n1 = int(0.6 * total)
    n2 = int(0.25 * total)
    phases = [(n1, lr), (n2, lr / 4), (total - n1 - n2, lr / 16)]
    t0 = time.perf_counter()
    losses = [net.train(n, a, seed + q) for q, (n, a) in enumerate(phases) if n > 0]
    return np.concatenate(losses), time.perf_counter() - t0
