"""ctypes wrapper of liboracle.so, the CPU FP64 restatement of SCFT.c.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the reference results are compared with or as
the timed CPU baseline.
The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load(abi):
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    P = C.POINTER
    vp = C.c_void_p
    dp = P(C.c_double)
    cp = P(C.c_double)  # interleaved complex
    sig = {
        "orc_derive": ([P(abi.ScftParams), P(abi.ScftDerived)], C.c_int),
        "orc_fft3d": ([C.c_int, C.c_int, C.c_int, cp, C.c_int], None),
        "orc_dft_brute": ([C.c_int, cp, cp, C.c_int], None),
        "orc_create": ([P(abi.ScftParams)], vp),
        "orc_destroy": ([vp], None),
        "orc_set_fields": ([vp, dp, dp], None),
        "orc_get_fields": ([vp, dp, dp], None),
        "orc_set_mask": ([vp, dp], None),
        "orc_get_densities": ([vp, dp, dp], None),
        "orc_get_propagator": ([vp, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "orc_seed48": ([vp, P(C.c_ushort)], None),
        "orc_get_seed": ([vp, P(C.c_ushort)], None),
        "orc_drand48": ([vp], C.c_double),
        "orc_gauss": ([vp], C.c_double),
        "orc_gauss_n": ([vp, C.c_long, C.POINTER(C.c_double)], None),
        "orc_iterate": ([vp, C.c_ulong, P(abi.ScftIterOut)], C.c_int),
        "orc_pde_block": ([vp, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int, C.c_ulong], None),
        "orc_main": ([P(abi.ScftParams), C.c_ulong, C.c_int, C.c_int, C.c_char_p], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    def __init__(self, abi):
        self.abi = abi
        self.lib = _load(abi)

    def derive(self, params):
        d = self.abi.ScftDerived()
        rc = self.lib.orc_derive(C.byref(params), C.byref(d))
        return rc, d

    def fft3d(self, x, sign):
        a = np.ascontiguousarray(x, dtype=np.complex128).copy()
        n0, n1, n2 = a.shape
        self.lib.orc_fft3d(n0, n1, n2, a.ctypes.data_as(C.POINTER(C.c_double)), sign)
        return a

    def dft_brute(self, x, sign):
        a = np.ascontiguousarray(x, dtype=np.complex128)
        out = np.zeros_like(a)
        self.lib.orc_dft_brute(len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
                               out.ctypes.data_as(C.POINTER(C.c_double)), sign)
        return out

    def main(self, params, iter_max, d_wrt, d_wrt_energies, directory):
        return self.lib.orc_main(C.byref(params), iter_max, d_wrt, d_wrt_energies, directory.encode())

    def state(self, params):
        return OracleState(self, params)


class OracleState:
    def __init__(self, orc, params):
        self.orc = orc
        self.lib = orc.lib
        self.abi = orc.abi
        self.h = self.lib.orc_create(C.byref(params))
        if not self.h:
            raise ValueError("oracle rejected the parameters")
        rc, d = orc.derive(params)
        self.derived = d
        self.shape = (d.nx, d.ny, d.nz)
        self.V = d.nx * d.ny * d.nz

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, w, p):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(13, self.V)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(self.V)
        self.lib.orc_set_fields(self.h, _ptr(w), _ptr(p))

    def get_fields(self):
        w = np.zeros((13, self.V))
        p = np.zeros(self.V)
        self.lib.orc_get_fields(self.h, _ptr(w), _ptr(p))
        return w, p

    def set_mask(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(3, self.V)
        self.lib.orc_set_mask(self.h, _ptr(m))

    def get_densities(self):
        phi = np.zeros((13, self.V))
        tot = np.zeros(self.V)
        self.lib.orc_get_densities(self.h, _ptr(phi), _ptr(tot))
        return phi, tot

    def get_propagator(self, sp, which, s):
        out = np.zeros(self.V)
        rc = self.lib.orc_get_propagator(self.h, sp, which, s, _ptr(out))
        if rc != 0:
            raise ValueError("no such propagator slice")
        return out

    def seed48(self, seed):
        arr = (C.c_ushort * 3)(*seed)
        self.lib.orc_seed48(self.h, arr)

    def get_seed(self):
        arr = (C.c_ushort * 3)()
        self.lib.orc_get_seed(self.h, arr)
        return tuple(arr)

    def drand48(self):
        return self.lib.orc_drand48(self.h)

    def gauss(self):
        return self.lib.orc_gauss(self.h)

    def gauss_n(self, n):
        out = np.empty(n)
        self.lib.orc_gauss_n(self.h, n, out.ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def iterate(self, it):
        o = self.abi.ScftIterOut()
        self.lib.orc_iterate(self.h, it, C.byref(o))
        return o

    def pde_block(self, start, stop, ns_eff, q, w, seed_mode=0, it=2):
        q = np.ascontiguousarray(q, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        self.lib.orc_pde_block(self.h, start, stop, ns_eff, _ptr(q), _ptr(w), seed_mode, it)
        return q
