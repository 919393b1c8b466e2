import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libpot3d_hip.so kernels)")
    config.addinivalue_line("markers", "slow: long CPU run (set P3D_SLOW=1 to include)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("P3D_SLOW", "0") == "1":
        return
    skip = pytest.mark.skip(reason="slow CPU test; set P3D_SLOW=1")
    for it in items:
        if "slow" in it.keywords and "gpu" not in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle as O
    O.lib()
    return O


# The synthetic workloads in the tests take br0 from the oracle's readbr
# (the CPU restatement); the GPU's p3d_readbr is checked against it by
# tests/test_gpu_readbr.py.
import readbr as _oracle_readbr  # noqa: E402
from p3d_driver import workloads as _workloads  # noqa: E402

_workloads.set_readbr(lambda mesh, tn, pn, bn, balance, abs_br:
                      _oracle_readbr.readbr(mesh, tn, pn, bn, balance, abs_br)[0])


def small_mesh(nr=9, nt=8, np_=12, drratio=1.5, r0=1.0, r1=2.5):
    from p3d_driver import inputs
    return inputs.global_mesh(nr, nt, np_, r0, r1, rfrac=(0.0, 1.0), drratio=(drratio,))


def random_br0(mesh, seed=0):
    """A smooth-ish random br0(nt,np) with the driver's periodic/pole
    treatment and flux balance applied (readbr's post-processing)."""
    rng = np.random.default_rng(seed)
    br0 = np.zeros((mesh.nt, mesh.np))
    br0[1:-1, 1:-1] = rng.standard_normal((mesh.nt - 2, mesh.np - 2))
    return _oracle_readbr.finish(mesh, br0, True, False)[0]
