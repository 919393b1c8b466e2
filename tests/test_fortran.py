"""The Fortran side of the drop-in: pot3d_amd/fortran/pot3d_hip_mod.f90
declares the C ABI with iso_c_binding, and a slim potfld driver
(p3d_fortran_demo.f90) links against libpot3d_hip.so."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "pot3d_amd", "fortran")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
DEMO = os.path.join(FDIR, "p3d_fortran_demo")


@pytest.mark.skipif(not os.path.exists(FLANG), reason="ROCm flang not installed")
def test_fortran_module_compiles_and_links():
    subprocess.run(["make", "-s", "-C", FDIR], check=True)
    nm = subprocess.run(["nm", DEMO], capture_output=True, text=True, check=True).stdout
    for sym in ("p3d_create", "p3d_solve", "p3d_fill_phi", "p3d_energy"):
        assert f"U {sym}" in nm


def _demo_mesh():
    """The uniform meshes p3d_fortran_demo.f90 builds (nr=17, nt=19, np=28)."""
    from p3d_driver import inputs
    nr, nt, np_ = 17, 19, 28
    pi = inputs.PI
    r = np.array([1.0 + 1.5 * (i - 1) / (nr - 2) for i in range(1, nr)])
    t = np.array([pi * (j - 1) / (nt - 2) for j in range(1, nt)])
    p = np.empty(np_)
    p[:np_ - 1] = [2.0 * pi * (k - 1) / (np_ - 2) for k in range(1, np_)]
    p[np_ - 1] = p[1] + 2.0 * pi

    def half(c, n):
        h = np.empty(n)
        dh = np.empty(n)
        h[1:n - 1] = 0.5 * (c[1:] + c[:-1])
        dh[1:n - 1] = c[1:] - c[:-1]
        h[0] = h[1] - dh[1]
        h[n - 1] = h[n - 2] + dh[n - 2]
        dh[0] = dh[1]
        dh[n - 1] = dh[n - 2]
        return h, dh

    rh, drh = half(r, nr)
    th, dth = half(t, nt)
    dr = rh[1:] - rh[:-1]
    dt = th[1:] - th[:-1]
    ph = np.empty(np_)
    dph = np.empty(np_)
    ph[1:] = 0.5 * (p[1:] + p[:-1])
    dph[1:] = p[1:] - p[:-1]
    ph[0] = ph[np_ - 2] - 2.0 * pi
    dph[0] = dph[np_ - 2]
    dp = np.empty(np_)
    dp[:np_ - 1] = ph[1:] - ph[:-1]
    dp[np_ - 1] = dp[1]
    dph[np_ - 1] = dph[1]
    dp[0] = dp[np_ - 2]
    return inputs.Mesh(nr, nt, np_, r, dr, rh, drh, t, dt, th, dth, p, dp, ph, dph)


@pytest.mark.gpu
def test_fortran_demo_solves_on_gpu(oracle_mod):
    print("[gpu] fortran demo", flush=True)
    if not os.path.exists(DEMO):
        pytest.fail("p3d_fortran_demo not built (run __graft_entry__.build())")
    out = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    line = [l for l in out.stdout.splitlines() if l.startswith("FORTRAN_DEMO")][0].split()
    ncg, w = int(line[1]), np.array([float(v) for v in line[2:5]])
    mesh = _demo_mesh()
    o = oracle_mod.Oracle(mesh, -1.0)
    br0 = np.cos(mesh.th)[:, None] * np.ones((1, mesh.np))
    res = o.cgsolve(o.rhs(br0), ifprec=3, epscg=1e-9)
    wref = o.energy(o.fill_phi(res["x"], br0))
    assert abs(ncg - res["ncg"]) <= 2
    assert np.allclose(w, wref, rtol=1e-7, atol=1e-12)
