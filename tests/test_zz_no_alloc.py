"""scft_step_noalloc (one contour step on caller device arrays, scratch in a
static device array, no allocation) against the context path: slice s + 1
of a chain from scft_propagate equals the step applied to slice s, at 1e-10,
on two calls in a row, the second on a smaller grid."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libscft links (already loaded with it), for the test's
    own device buffers."""
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipDeviceSynchronize.argtypes = []
    for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


H2D, D2H = 1, 2


def _dev(h, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), a.nbytes) == 0
    assert h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
    return p


def _case(configs, which):
    if which == "c2":   # 256 x 1 x 128, MASK 1 (the C2 grid)
        p = configs.params_c2()
        w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
    else:               # 12 x 1 x 10 film
        p = configs.t_film_small()
        w, pf = configs.droplet_fields(p, radius=3.0, seed=1)
    return p, w, pf


def test_step_noalloc_matches_propagate(abi, configs, lib):
    h = _hip()
    for which in ("c2", "film"):
        p, w, pf = _case(configs, which)
        print("case", which, abi.grid_of(p), flush=True)
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        ctx.set_fields(w, pf)
        ctx.set_mask(configs.brush_mask(p))
        ctx.propagate(2)
        V = ctx.V
        wf, _ = ctx.get_fields()
        # homopolymer C (species 6, field C) and the dbk1 q (field A for the first block)
        for sp, field, s in ((6, 2, 0), (6, 2, 7), (0, 0, 3)):
            q_in = ctx.get_propagator(sp, 0, s)
            want = ctx.get_propagator(sp, 0, s + 1)
            dw, dq, do = _dev(h, wf[field]), _dev(h, q_in), _dev(h, np.zeros(V))
            rc = lib.scft_step_noalloc(C.byref(p), dw, dq, do, None)
            assert rc == 0, lib.scft_last_error()
            assert h.hipDeviceSynchronize() == 0
            got = np.empty(V)
            assert h.hipMemcpy(got.ctypes.data_as(C.c_void_p), do, got.nbytes, D2H) == 0
            for ptr in (dw, dq, do):
                assert h.hipFree(ptr) == 0
            err = np.abs(got - want).max() / np.abs(want).max()
            print(which, "species", sp, "slice", s, "rel %.2e" % err, flush=True)
            assert err <= 1e-10
        ctx.close()


def test_step_noalloc_rejects_oversized_grid(abi, configs, lib):
    """Past its voxel limit the entry point returns SCFT_EUNSUPPORTED with a
    message, before touching the device arrays."""
    p = configs.params_c3()   # 128 x 128 x 64 = 1 048 576 voxels
    one = C.c_void_p(1)
    rc = lib.scft_step_noalloc(C.byref(p), one, one, one, None)
    assert rc == -5
    assert b"exceed" in lib.scft_last_error()
