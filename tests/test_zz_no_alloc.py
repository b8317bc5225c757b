"""rc_box_field_noalloc (box fields staged in the library's static device
array, no allocation) against rc_box_field on the same state: bit for bit, on
two calls in a row, the second with a smaller box, with device, pinned-host
and NULL outputs."""
import ctypes as C

import numpy as np
import pytest

from rheocell_amd import confgen
from rheocell_amd.conf import RCexception
from rheocell_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _hip():
    # the HIP runtime librheocell.so itself links (torch bundles another one,
    # and two runtimes cannot share the GPU in one process)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipHostFree.argtypes = [C.c_void_p]
    return hip


def _noalloc(sim, box, which, g, L, null_cy=False):
    """Outputs: two device buffers, two pinned host buffers, all pre-filled
    with NaN bytes (0xFF); cy NULL when null_cy."""
    hip = _hip()
    n = 8 * L * L
    bufs = [C.c_void_p() for _ in range(4)]
    for i, b in enumerate(bufs):
        if i < 2:
            assert hip.hipMalloc(C.byref(b), n) == 0
            assert hip.hipMemset(b, 0xFF, n) == 0
        else:
            assert hip.hipHostMalloc(C.byref(b), n, 0) == 0
            C.memset(b, 0xFF, n)
    try:
        ptrs = [bufs[0], bufs[1], bufs[2], None if null_cy else bufs[3]]
        sim._check(sim.L.rc_box_field_noalloc(sim.h, box, which, g, *ptrs))
        out = []
        for i, b in enumerate(bufs):
            a = np.empty(L * L)
            if i < 2:
                assert hip.hipMemcpy(a.ctypes.data_as(C.c_void_p), b, n, 2) == 0   # hipMemcpyDeviceToHost
            else:
                C.memmove(a.ctypes.data_as(C.c_void_p), b, n)
            out.append(a.reshape(L, L))
        return out
    finally:
        for i, b in enumerate(bufs):
            (hip.hipFree if i < 2 else hip.hipHostFree)(b)


def _sim(L, nside, seed, fc):
    keys = dict(confgen.BASE_KEYS, seed=seed, friction_cell=fc, CGtolerance=1e-10)
    sim = Simulation([(confgen.lattice_conf(L, nside, seed=seed), keys)])
    sim.step(2)
    return sim


def test_noalloc_box_field_matches_rc_box_field_on_two_calls():
    big = _sim(160, 4, 21, 0)      # first call: a 160 x 160 dry box
    small = _sim(64, 2, 22, 1)     # second call: a smaller, wet box on the same static array
    for sim, L, g in ((big, 160, 5), (small, 64, 3)):
        for which in (0, 1):
            ref = sim.box_field(which, g)
            got = _noalloc(sim, 0, which, g, L)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (L, which, np.abs(a - b).max())
            assert np.abs(ref[0]).max() > 0
    # NULL outputs are skipped; the rest still agree
    ref = small.box_field(0, 4)
    got = _noalloc(small, 0, 0, 4, 64, null_cy=True)
    for a, b in zip(got[:3], ref[:3]):
        assert np.array_equal(a, b)
    assert np.all(np.isnan(got[3]))
    big.close()
    small.close()


def test_noalloc_box_field_refuses_a_box_above_the_static_limit():
    keys = dict(confgen.BASE_KEYS, seed=3)
    conf = confgen.lattice_conf(544, 1, seed=3)
    sim = Simulation([(conf, keys)])
    with pytest.raises(RCexception, match="RC_ERR_ARG.*static staging limit"):
        _noalloc(sim, 0, 0, 3, 544)
    sim.close()
