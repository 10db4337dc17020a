"""ctypes mirror of include/scft.h (the C ABI of libscft).

The structs here are byte-for-byte the C structs, so the same objects can be
handed to libscft and to any other library that takes an scft_params (the
CPU oracle under oracle/ uses the same header).
"""
import ctypes as C
import os

import numpy as np

NF = 13          # fields A..J, S1..S3
NSP = 17         # species: dbk1..4, homo A..J, S1..S3
FIELD_NAMES = list("ABCDEFGHIJ") + ["S1", "S2", "S3"]
SPECIES_NAMES = ["dbk1", "dbk2", "dbk3", "dbk4"] + ["homo" + c for c in "ABCDEFGHIJ"] + ["S1", "S2", "S3"]

SCFT_FP64 = 0
SCFT_FP32_SPLIT = 1      # FP32 slices, unbiased rounding
SCFT_FP32_NEAREST = 2    # FP32 slices, round to nearest

ERRORS = {0: "ok", -1: "invalid argument", -2: "HIP error", -3: "out of device memory",
          -4: "volumes do not sum to 1", -5: "unsupported size", -6: "non-finite value",
          -7: "device noise stream short of accepted pairs"}


class ScftParams(C.Structure):
    _fields_ = [
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("dr_target", C.c_double),
        ("ns", C.c_int), ("mask", C.c_int),
        ("g1", C.c_double), ("dt_noise", C.c_double),
        ("relaxation", C.c_ulong),
        ("enable_seed", C.c_int), ("seed_steps", C.c_int), ("seed_area_delta", C.c_int),
        ("chi", (C.c_double * NF) * NF),
        ("f_dbk", C.c_double * 4), ("len_dbk", C.c_double * 4), ("vol_dbk", C.c_double * 4),
        ("vol_homo", C.c_double * 10), ("len_homo", C.c_double * 10),
        ("vol_s", C.c_double * 3), ("len_s", C.c_double * 3),
    ]


class ScftDerived(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("nzext", C.c_int),
        ("ns_sp", C.c_int * NSP), ("ns_x", C.c_int * 4), ("ns_y", C.c_int * 4),
        ("alpha", C.c_double * NSP),
        ("phi_avg", C.c_double * NF),
        ("chi_eff", (C.c_double * NF) * NF),
        ("coating_vol", C.c_double),
        ("active", C.c_int * NF),
        ("n_active", C.c_int),
    ]


class ScftIterOut(C.Structure):
    _fields_ = [
        ("free_energy", C.c_double), ("rmsdev", C.c_double), ("dw_rms", C.c_double),
        ("qsum", C.c_double * NSP), ("chempot", C.c_double * NSP),
        ("mf", C.c_float),
    ]

    def as_dict(self):
        return {"F": self.free_energy, "rmsdev": self.rmsdev, "dw_rms": self.dw_rms,
                "qsum": list(self.qsum), "chempot": list(self.chempot), "mf": self.mf}


def field_index(name):
    return FIELD_NAMES.index(name)


def make_params(lx=8.0, ly=1.0, lz=8.0, dr_target=0.125, ns=101, mask=0, g1=0.1,
                dt_noise=10.0, relaxation=0, enable_seed=0, seed_steps=1, seed_area_delta=1,
                chi=None, f_dbk=(0.5, 0.5, 0.3, 0.5), len_dbk=(1.0,) * 4, vol_dbk=(0.0,) * 4,
                vol_homo=(0.0,) * 10, len_homo=(1.0,) * 10, vol_s=(0.0,) * 3, len_s=(0.1,) * 3):
    """Build an ScftParams.  chi is a dict like {"AB": 20.0, "S1A": 40.0}
    (names as in parameters.h, order of the two letters free)."""
    p = ScftParams()
    p.lx, p.ly, p.lz, p.dr_target = lx, ly, lz, dr_target
    p.ns, p.mask, p.g1, p.dt_noise = ns, mask, g1, dt_noise
    p.relaxation = relaxation
    p.enable_seed, p.seed_steps, p.seed_area_delta = enable_seed, seed_steps, seed_area_delta
    for i in range(NF):
        for j in range(NF):
            p.chi[i][j] = 0.0
    for key, val in (chi or {}).items():
        a, b = _split_pair(key)
        ia, ib = field_index(a), field_index(b)
        p.chi[ia][ib] = val
        p.chi[ib][ia] = val
    for i in range(4):
        p.f_dbk[i], p.len_dbk[i], p.vol_dbk[i] = f_dbk[i], len_dbk[i], vol_dbk[i]
    for i in range(10):
        p.vol_homo[i], p.len_homo[i] = vol_homo[i], len_homo[i]
    for i in range(3):
        p.vol_s[i], p.len_s[i] = vol_s[i], len_s[i]
    return p


def _split_pair(key):
    key = key.replace("_", "")
    toks = []
    i = 0
    while i < len(key):
        if key[i] == "S" and i + 1 < len(key) and key[i + 1].isdigit():
            toks.append(key[i:i + 2])
            i += 2
        else:
            toks.append(key[i])
            i += 1
    if len(toks) != 2:
        raise ValueError("bad chi key %r" % key)
    return toks


def grid_of(p):
    def n(L):
        return 1 if L == 1.0 else int(L / p.dr_target + 0.5)
    return n(p.lx), n(p.ly), n(p.lz)


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ScftError(RuntimeError):
    pass


PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libscft.so")


def load_library(path=None):
    """Load libscft.so (the HIP library).  Raises if it is not built."""
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise ScftError("libscft.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    dp = P(C.c_double)
    sig = {
        "scft_derive": ([P(ScftParams), P(ScftDerived)], C.c_int),
        "scft_create": ([P(ScftParams), C.c_int, P(vp)], C.c_int),
        "scft_destroy": ([vp], C.c_int),
        "scft_get_derived": ([vp, P(ScftDerived)], C.c_int),
        "scft_set_fields": ([vp, dp, dp], C.c_int),
        "scft_get_fields": ([vp, dp, dp], C.c_int),
        "scft_set_mask": ([vp, dp], C.c_int),
        "scft_get_densities": ([vp, dp, dp], C.c_int),
        "scft_get_propagator": ([vp, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "scft_propagate": ([vp, C.c_ulong], C.c_int),
        "scft_densities": ([vp, dp, dp], C.c_int),
        "scft_update_fields": ([vp, C.c_ulong, C.c_float, dp, P(ScftIterOut)], C.c_int),
        "scft_iterate": ([vp, C.c_ulong, C.c_float, dp, P(ScftIterOut)], C.c_int),
        "scft_rng_seed": ([vp, P(C.c_ushort)], C.c_int),
        "scft_rng_state": ([vp, P(C.c_ushort)], C.c_int),
        "scft_slot_rng_seed": ([vp, C.c_int, P(C.c_ushort)], C.c_int),
        "scft_slot_rng_state": ([vp, C.c_int, P(C.c_ushort)], C.c_int),
        "scft_sync": ([vp], C.c_int),
        "scft_last_error": ([], C.c_char_p),
        "scft_batch_create": ([P(ScftParams), C.c_int, C.c_int, C.c_int, P(vp)], C.c_int),
        "scft_batch_size": ([vp], C.c_int),
        "scft_slot_init": ([vp, C.c_int, P(ScftParams)], C.c_int),
        "scft_slot_activate": ([vp, C.c_int, C.c_int], C.c_int),
        "scft_slot_get_derived": ([vp, C.c_int, P(ScftDerived)], C.c_int),
        "scft_slot_set_fields": ([vp, C.c_int, dp, dp], C.c_int),
        "scft_slot_get_fields": ([vp, C.c_int, dp, dp], C.c_int),
        "scft_slot_copy_fields": ([vp, C.c_int, C.c_int], C.c_int),
        "scft_slot_set_mask": ([vp, C.c_int, dp], C.c_int),
        "scft_slot_get_densities": ([vp, C.c_int, dp, dp], C.c_int),
        "scft_slot_get_propagator": ([vp, C.c_int, C.c_int, C.c_int, C.c_int, dp], C.c_int),
        "scft_batch_iterate": ([vp, P(C.c_ulong), P(ScftIterOut)], C.c_int),
        "scft_batch_launch": ([vp, P(C.c_ulong)], C.c_int),
        "scft_batch_complete": ([vp, P(ScftIterOut)], C.c_int),
        "scft_batch_query": ([vp], C.c_int),
        "scft_get_schedule": ([vp, C.c_char_p, C.c_int], C.c_int),
        "scft_slot_bytes": ([P(ScftParams), C.c_int, C.c_int, P(C.c_size_t)], C.c_int),
        "scft_device_memory": ([P(C.c_size_t), P(C.c_size_t)], C.c_int),
        "scft_time_launches": ([vp, C.c_int, C.c_int, dp, dp, dp, dp, P(C.c_int)], C.c_int),
        "scft_step_noalloc": ([P(ScftParams), vp, vp, vp, vp], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def check(lib, rc, what):
    if rc != 0:
        msg = lib.scft_last_error()
        raise ScftError("%s failed: %s (%d): %s" % (what, ERRORS.get(rc, "?"), rc,
                                                   msg.decode() if msg else ""))


class Context:
    """One SCFT instance on the GPU, driven through the C ABI."""

    def __init__(self, lib, params, precision=SCFT_FP64):
        self.lib = lib
        self.params = params
        h = C.c_void_p()
        check(lib, lib.scft_create(C.byref(params), precision, C.byref(h)), "scft_create")
        self.h = h
        d = ScftDerived()
        check(lib, lib.scft_get_derived(h, C.byref(d)), "scft_get_derived")
        self.derived = d
        self.shape = (d.nx, d.ny, d.nz)
        self.V = d.nx * d.ny * d.nz

    def close(self):
        if self.h:
            self.lib.scft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_fields(self, w, p):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(NF, self.V)
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(self.V)
        check(self.lib, self.lib.scft_set_fields(self.h, _ptr(w), _ptr(p)), "scft_set_fields")

    def get_fields(self):
        w = np.zeros((NF, self.V))
        p = np.zeros(self.V)
        check(self.lib, self.lib.scft_get_fields(self.h, _ptr(w), _ptr(p)), "scft_get_fields")
        return w, p

    def set_mask(self, m):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(3, self.V)
        check(self.lib, self.lib.scft_set_mask(self.h, _ptr(m)), "scft_set_mask")

    def get_densities(self):
        phi = np.zeros((NF, self.V))
        tot = np.zeros(self.V)
        check(self.lib, self.lib.scft_get_densities(self.h, _ptr(phi), _ptr(tot)), "scft_get_densities")
        return phi, tot

    def get_propagator(self, sp, which, s):
        out = np.zeros(self.V)
        check(self.lib, self.lib.scft_get_propagator(self.h, sp, which, s, _ptr(out)), "scft_get_propagator")
        return out

    def propagate(self, it):
        check(self.lib, self.lib.scft_propagate(self.h, it), "scft_propagate")

    def densities(self):
        q = np.zeros(NSP)
        cp = np.zeros(NSP)
        check(self.lib, self.lib.scft_densities(self.h, _ptr(q), _ptr(cp)), "scft_densities")
        return q, cp

    def update_fields(self, it, mf=0.0, noise=None):
        o = ScftIterOut()
        nz = None if noise is None else _ptr(np.ascontiguousarray(noise, dtype=np.float64))
        check(self.lib, self.lib.scft_update_fields(self.h, it, mf, nz, C.byref(o)), "scft_update_fields")
        return o

    def iterate(self, it, mf=0.0, noise=None):
        o = ScftIterOut()
        nz = None if noise is None else _ptr(np.ascontiguousarray(noise, dtype=np.float64))
        check(self.lib, self.lib.scft_iterate(self.h, it, mf, nz, C.byref(o)), "scft_iterate")
        return o

    def rng_seed(self, xsubi):
        """Draw the noise stream on the device from this drand48 state
        (seed48 layout); iterate() then takes no noise array."""
        x = (C.c_ushort * 3)(*xsubi)
        check(self.lib, self.lib.scft_rng_seed(self.h, x), "scft_rng_seed")

    def rng_state(self):
        x = (C.c_ushort * 3)()
        check(self.lib, self.lib.scft_rng_state(self.h, x), "scft_rng_state")
        return tuple(x)

    def sync(self):
        check(self.lib, self.lib.scft_sync(self.h), "scft_sync")

    def schedule(self):
        buf = C.create_string_buffer(256)
        check(self.lib, self.lib.scft_get_schedule(self.h, buf, 256), "scft_get_schedule")
        return buf.value.decode()


class Batch:
    """A batch of SCFT instances (slots) that advance together on the GPU."""

    def __init__(self, lib, proto, nslots, max_nx=0, precision=SCFT_FP64):
        self.lib = lib
        self.proto = proto
        self.nslots = nslots
        h = C.c_void_p()
        check(lib, lib.scft_batch_create(C.byref(proto), nslots, max_nx, precision, C.byref(h)),
              "scft_batch_create")
        self.h = h
        self.max_nx = max(max_nx, grid_of(proto)[0])
        self.V = [0] * nslots
        self.params = [None] * nslots
        self._outs = (ScftIterOut * nslots)()
        self._iters = (C.c_ulong * nslots)()

    def close(self):
        if self.h:
            self.lib.scft_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_slot(self, s, params):
        check(self.lib, self.lib.scft_slot_init(self.h, s, C.byref(params)), "scft_slot_init")
        d = ScftDerived()
        check(self.lib, self.lib.scft_slot_get_derived(self.h, s, C.byref(d)), "scft_slot_get_derived")
        self.V[s] = d.nx * d.ny * d.nz
        self.params[s] = params
        return d

    def activate(self, s, on=True):
        check(self.lib, self.lib.scft_slot_activate(self.h, s, 1 if on else 0), "scft_slot_activate")

    def rng_seed(self, s, xsubi):
        """Give slot s its own noise stream from this drand48 state."""
        x = (C.c_ushort * 3)(*xsubi)
        check(self.lib, self.lib.scft_slot_rng_seed(self.h, s, x), "scft_slot_rng_seed")

    def rng_state(self, s):
        x = (C.c_ushort * 3)()
        check(self.lib, self.lib.scft_slot_rng_state(self.h, s, x), "scft_slot_rng_state")
        return tuple(x)

    def set_fields(self, s, w, p):
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(NF, self.V[s])
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(self.V[s])
        check(self.lib, self.lib.scft_slot_set_fields(self.h, s, _ptr(w), _ptr(p)), "scft_slot_set_fields")

    def get_fields(self, s):
        w = np.zeros((NF, self.V[s]))
        p = np.zeros(self.V[s])
        check(self.lib, self.lib.scft_slot_get_fields(self.h, s, _ptr(w), _ptr(p)), "scft_slot_get_fields")
        return w, p

    def copy_fields(self, dst, src):
        check(self.lib, self.lib.scft_slot_copy_fields(self.h, dst, src), "scft_slot_copy_fields")

    def set_mask(self, s, m):
        m = np.ascontiguousarray(m, dtype=np.float64).reshape(3, self.V[s])
        check(self.lib, self.lib.scft_slot_set_mask(self.h, s, _ptr(m)), "scft_slot_set_mask")

    def get_densities(self, s):
        phi = np.zeros((NF, self.V[s]))
        tot = np.zeros(self.V[s])
        check(self.lib, self.lib.scft_slot_get_densities(self.h, s, _ptr(phi), _ptr(tot)),
              "scft_slot_get_densities")
        return phi, tot

    def get_propagator(self, s, sp, which, idx):
        out = np.zeros(self.V[s])
        check(self.lib, self.lib.scft_slot_get_propagator(self.h, s, sp, which, idx, _ptr(out)),
              "scft_slot_get_propagator")
        return out

    @property
    def iteration_numbers(self):
        """The per-slot iteration-number array iterate() passes (fill in place)."""
        return self._iters

    def iterate(self, iters=None):
        """One iteration of every active slot; returns the out structs (by slot).
        iters: per-slot iteration numbers (seed box, MF), or None."""
        ip = None
        if iters is not None:
            if iters is not self._iters:
                for i, v in enumerate(iters):
                    self._iters[i] = v
            ip = self._iters
        check(self.lib, self.lib.scft_batch_iterate(self.h, ip, self._outs), "scft_batch_iterate")
        return self._outs

    def launch(self, iters=None):
        """Start one iteration of every active slot on the batch's stream and
        return at once; complete() waits for it and returns the outs."""
        ip = None
        if iters is not None:
            if iters is not self._iters:
                for i, v in enumerate(iters):
                    self._iters[i] = v
            ip = self._iters
        check(self.lib, self.lib.scft_batch_launch(self.h, ip), "scft_batch_launch")

    def complete(self):
        check(self.lib, self.lib.scft_batch_complete(self.h, self._outs), "scft_batch_complete")
        return self._outs

    def ready(self):
        """True once the launched iteration has finished (complete() will not wait)."""
        rc = self.lib.scft_batch_query(self.h)
        if rc < 0:
            check(self.lib, rc, "scft_batch_query")
        return rc == 1

    def schedule(self):
        buf = C.create_string_buffer(256)
        check(self.lib, self.lib.scft_get_schedule(self.h, buf, 256), "scft_get_schedule")
        return buf.value.decode()

    def time_launches(self, kind=0, reps=10):
        """Mean duration (us), kernel bytes, flops and voxel-steps of one launch
        of the z-operator (kind 0) or last-axis (kind 1) kernel of this schedule."""
        us, b, f, v, n = C.c_double(), C.c_double(), C.c_double(), C.c_double(), C.c_int()
        check(self.lib, self.lib.scft_time_launches(self.h, kind, reps, C.byref(us), C.byref(b), C.byref(f),
                                                    C.byref(v), C.byref(n)), "scft_time_launches")
        return {"avg_us": us.value, "bytes": b.value, "flops": f.value, "voxel_steps": v.value,
                "launches": n.value}
