"""p3d_nn_prolong_dev_noalloc (the prolongation that allocates nothing: its
tables are staged in a fixed device array of the library) against
p3d_nn_prolong_dev (tables in buffers allocated per call), through the C ABI:
bit for bit on two calls in a row, the second on a smaller grid (the fixed
array keeps the first call's larger tables, which the second must not
read), and the size limit rejected before any device work."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _tables(coarse, fine):
    from pot3d_amd import warmstart as W
    from p3d_driver import inputs
    cm = inputs.global_mesh(*coarse, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(2.5,))
    fm = inputs.global_mesh(*fine, 1.0, 2.5, rfrac=(0.0, 1.0), drratio=(2.5,))
    return W.prolong_tables(cm, fm)


def test_noalloc_prolong_matches_allocating_entry_point():
    from pot3d_amd import warmstart as W
    from pot3d_amd._hip import Hip
    hip = Hip()
    rng = np.random.default_rng(7)
    # the second call is smaller on every axis and in total
    for coarse, fine in (((26, 46, 91), (51, 91, 181)), ((11, 17, 24), (21, 31, 48))):
        tab = _tables(coarse, fine)
        nc = tab.nrc * tab.ntc * tab.npc
        nf = tab.af * tab.bf * tab.cf
        say(f"[gpu] no-alloc prolongation {coarse} -> {fine}: {nc} -> {nf} points")
        phic = rng.standard_normal(nc)
        d_c = hip.to_device(phic)
        d_ref = hip.malloc(8 * nf)
        d_out = hip.malloc(8 * nf)
        try:
            hip.memset(d_out, 0xFF, 8 * nf)  # NaN pattern: every point must be written
            W.prolong_dev(tab, d_c, d_ref)
            W.prolong_dev_noalloc(tab, d_c, d_out)
            ref = np.empty(nf)
            out = np.empty(nf)
            hip.d2h(ref, d_ref)
            hip.d2h(out, d_out)
        finally:
            hip.free(d_c)
            hip.free(d_ref)
            hip.free(d_out)
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
        # and both against the host entry point (same kernel on staged copies)
        assert np.array_equal(W.prolong(tab, phic), ref)


def test_noalloc_prolong_rejects_tables_beyond_the_limit():
    """af+bf+cf > 8192 is refused on the host; no kernel runs."""
    from pot3d_amd import _lib
    L = _lib.load()
    af, bf, cf = 4000, 4000, 200
    ir = np.zeros(af, np.int32)
    it = np.zeros(bf, np.int32)
    ip = np.zeros(cf, np.int32)
    wr, wt, wp = np.zeros(af), np.zeros(bf), np.zeros(cf)
    _ip = C.POINTER(C.c_int)
    _dp = C.POINTER(C.c_double)
    # the pointers are never dereferenced on the device: the call fails first
    rc = L.p3d_nn_prolong_dev_noalloc(4, 4, 4, C.c_void_p(16), af, bf, cf,
                                      ir.ctypes.data_as(_ip), wr.ctypes.data_as(_dp),
                                      it.ctypes.data_as(_ip), wt.ctypes.data_as(_dp),
                                      ip.ctypes.data_as(_ip), wp.ctypes.data_as(_dp),
                                      C.c_void_p(16))
    assert rc == _lib.P3D_EINVAL
    assert "no-allocation limit" in _lib.last_error()
