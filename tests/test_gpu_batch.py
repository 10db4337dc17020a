"""Batched instances and the two z-operator schedules on the GPU.

- every slot of a batch (different chi, volumes, G1 and NX) matches its own
  CPU oracle run to 1e-10;
- a slot's trajectory is the same alone and in the mix;
- the FFT and the dense z operator both match the oracle (MASK 0 and 1).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _srel(a, b):
    return abs(a - b) / max(abs(b), 1.0)


def _film(configs, lx, chi_ab, chi_ac, vol_dbk=0.2, g1=0.01):
    return configs.chem_droplet(lx=lx, ly=1.0, lz=10.0, dr=1.0, ns=21, vol_dbk=vol_dbk, vol_s1=0.05,
                                len_s1=0.15, chi_ab=chi_ab, chi_ac=chi_ac, g1=g1)


def test_batch_slots_match_oracle(abi, configs, orc, lib):
    ps = [_film(configs, 12.0, 25.0, 40.0), _film(configs, 16.0, 18.0, 30.0, vol_dbk=0.25, g1=0.02),
          _film(configs, 10.0, 30.0, 45.0)]
    b = abi.Batch(lib, ps[0], 3, max_nx=16, precision=abi.SCFT_FP64)
    refs = []
    for s, p in enumerate(ps):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        m = configs.brush_mask(p)
        b.set_fields(s, w, pf)
        b.set_mask(s, m)
        b.activate(s)
        r = orc.state(p)
        r.set_fields(w, pf)
        r.set_mask(m)
        refs.append(r)
    print(b.schedule())
    for it in (2, 3, 4):
        outs = b.iterate()
        for s, r in enumerate(refs):
            o = r.iterate(it)
            assert _srel(outs[s].free_energy, o.free_energy) <= TOL, (s, outs[s].free_energy, o.free_energy)
            assert _srel(outs[s].dw_rms, o.dw_rms) <= TOL
            for sp in range(17):
                assert _srel(outs[s].chempot[sp], o.chempot[sp]) <= TOL
    for s, r in enumerate(refs):
        wa, pa = b.get_fields(s)
        wb, pb = r.get_fields()
        assert _rel(wa, wb) <= TOL
        assert _rel(pa, pb) <= TOL
    b.close()


def test_same_result_alone_and_in_mix(abi, configs, lib):
    ps = [_film(configs, 12.0, 25.0, 40.0), _film(configs, 16.0, 18.0, 30.0), _film(configs, 12.0, 30.0, 45.0)]
    b = abi.Batch(lib, ps[0], 3, max_nx=16, precision=abi.SCFT_FP32_SPLIT)

    def setup(slots):
        for s in slots:
            b.init_slot(s, ps[s])
            w, pf = configs.droplet_fields(ps[s], radius=3.0, seed=s)
            b.set_fields(s, w, pf)
            b.set_mask(s, configs.brush_mask(ps[s]))
            b.activate(s)

    setup([0])
    alone = [b.iterate()[0].free_energy for _ in range(5)]
    w_alone, _ = b.get_fields(0)
    setup([0, 1, 2])
    mixed = [b.iterate()[0].free_energy for _ in range(5)]
    w_mixed, _ = b.get_fields(0)
    for a, m in zip(alone, mixed):
        assert abs(a - m) <= 1e-12 * max(1.0, abs(a))
    assert _rel(w_mixed, w_alone) <= 1e-12
    b.close()


@pytest.mark.parametrize("zkind", ["fft", "dense"])
@pytest.mark.parametrize("case", ["film", "bulk", "film3d"])
def test_z_operator_kinds_match_oracle(abi, configs, orc, lib, zkind, case):
    os.environ["SCFT_ZKIND"] = zkind
    try:
        if case == "film":
            p = _film(configs, 12.0, 25.0, 40.0)
        elif case == "bulk":
            p = configs.params_c1()
        else:
            p = configs.t_3d_small()
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        assert ("z=" + zkind) in ctx.schedule()
        ref = orc.state(p)
        d = ctx.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=3, amp=0.4, fields=act)
        for s in (ctx, ref):
            s.set_fields(w, pf)
            if p.mask:
                s.set_mask(configs.brush_mask(p))
        for it in (2, 3):
            a, o = ctx.iterate(it), ref.iterate(it)
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy)
        wa, _ = ctx.get_fields()
        wb, _ = ref.get_fields()
        assert _rel(wa, wb) <= TOL
        ctx.close()
    finally:
        del os.environ["SCFT_ZKIND"]


@pytest.mark.parametrize("zkind", ["fft", "dense"])
@pytest.mark.parametrize("nz,mask", [(128, 1), (127, 1), (100, 1), (64, 1), (48, 1), (128, 0), (96, 0)])
def test_z_operator_sizes_match_oracle(abi, configs, orc, lib, zkind, nz, mask):
    """The z operator at the production sizes (NZ up to 128, hence transform
    length 258 = 2*3*43 and 64-wide dense halves) on a narrow NX = 4 grid."""
    os.environ["SCFT_ZKIND"] = zkind
    try:
        p = configs.chem_droplet(lx=0.5, ly=1.0, lz=nz * 0.125, dr=0.125, ns=21, vol_dbk=0.2, vol_s1=0.05,
                                 len_s1=0.15)
        p.mask = mask
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        assert ctx.shape == (4, 1, nz)
        assert ("z=" + zkind) in ctx.schedule() or zkind == "dense" and nz > 128
        ref = orc.state(p)
        d = ctx.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=5, amp=0.5, fields=act)
        for s in (ctx, ref):
            s.set_fields(w, pf)
            if p.mask:
                s.set_mask(configs.brush_mask(p))
        for it in (2, 3):
            a, o = ctx.iterate(it), ref.iterate(it)
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy)
        for s in range(d.ns_sp[6]):
            assert _rel(ctx.get_propagator(6, 0, s), ref.get_propagator(6, 0, s)) <= TOL, s
        ctx.close()
    finally:
        del os.environ["SCFT_ZKIND"]


@pytest.mark.parametrize("variant", ["reg", "wave", "lds"])
@pytest.mark.parametrize("nx,ny,nz", [(16, 1, 6), (48, 1, 6), (80, 1, 6), (96, 1, 6), (160, 1, 6), (192, 1, 6),
                                      (320, 1, 6), (384, 1, 6), (512, 1, 6), (12, 1, 6), (32, 16, 6), (16, 48, 6),
                                      (64, 64, 6), (256, 1, 6), (256, 8, 6), (32, 256, 6), (256, 3, 5),
                                      (8, 256, 5), (128, 1, 6), (128, 6, 5), (16, 128, 6), (128, 128, 5),
                                      (192, 3, 5), (8, 192, 6), (320, 3, 5), (6, 320, 6), (96, 3, 5),
                                      (8, 96, 6), (96, 96, 5), (64, 3, 5), (8, 64, 6), (64, 1, 6)])
def test_periodic_axis_sizes_match_oracle(abi, configs, orc, lib, variant, nx, ny, nz):
    """x and y axes through the four-step register FFTs (lengths 128, 256 and
    64, 96, 192, 320), the wave-per-line FFT (compile-time lengths) and the generic
    LDS FFT, against the oracle, on a thin film; odd NZ makes the strided
    lines odd-aligned."""
    os.environ["SCFT_XFFT"] = variant
    try:
        p = configs.chem_droplet(lx=nx * 0.25, ly=(ny * 0.25 if ny > 1 else 1.0), lz=nz * 0.25, dr=0.25, ns=11,
                                 vol_dbk=0.2, vol_s1=0.05, len_s1=0.3)
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        assert ctx.shape == (nx, ny, nz)
        sched = ctx.schedule()
        ref = orc.state(p)
        d = ctx.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=6, amp=0.5, fields=act)
        for s in (ctx, ref):
            s.set_fields(w, pf)
            s.set_mask(configs.brush_mask(p))
        for it in (2, 3):
            a, o = ctx.iterate(it), ref.iterate(it)
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy, sched)
        wa, _ = ctx.get_fields()
        wb, _ = ref.get_fields()
        assert _rel(wa, wb) <= TOL
        ctx.close()
    finally:
        del os.environ["SCFT_XFFT"]


@pytest.mark.parametrize("nx,ny,nz", [(96, 3, 5), (8, 96, 6), (96, 96, 5), (192, 3, 5), (320, 3, 5), (6, 320, 6),
                                      (128, 6, 5), (256, 3, 5), (64, 3, 5), (64, 64, 5)])
def test_periodic_register_ffts_fp32_split(abi, configs, orc, lib, nx, ny, nz):
    """The register FFTs with FP32 propagator slices (FP64 arithmetic)
    against the FP64 oracle: F within 1e-6 kT, the updated fields within
    1e-5 of their largest value, over two iterations."""
    os.environ["SCFT_XFFT"] = "reg"
    try:
        p = configs.chem_droplet(lx=nx * 0.25, ly=(ny * 0.25 if ny > 1 else 1.0), lz=nz * 0.25, dr=0.25, ns=11,
                                 vol_dbk=0.2, vol_s1=0.05, len_s1=0.3)
        ctx = abi.Context(lib, p, abi.SCFT_FP32_SPLIT)
        sched = ctx.schedule()
        ref = orc.state(p)
        d = ctx.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=6, amp=0.5, fields=act)
        for s in (ctx, ref):
            s.set_fields(w, pf)
            s.set_mask(configs.brush_mask(p))
        for it in (2, 3):
            a, o = ctx.iterate(it), ref.iterate(it)
            print((nx, ny, nz), it, "dF = %.3e" % abs(a.free_energy - o.free_energy), sched, flush=True)
            assert abs(a.free_energy - o.free_energy) <= 1e-6, (a.free_energy, o.free_energy, sched)
        wa, _ = ctx.get_fields()
        wb, _ = ref.get_fields()
        assert _rel(wa, wb) <= 1e-5
        n = dict(kv.split("=") for kv in ctx.schedule().split() if "=" in kv and kv[0] != "z")
        kind = "fftreg2" if {nx, ny} & {64, 96, 192, 320} else "fftreg"
        assert int(n[kind]) > 0, ctx.schedule()
        ctx.close()
    finally:
        del os.environ["SCFT_XFFT"]


@pytest.mark.parametrize("variant,nz,mask", [("valu", 128, 1), ("valu", 127, 1), ("valu", 48, 1), ("valu", 64, 0),
                                              ("valu", 20, 1), ("mfma1", 128, 1), ("mfma1", 96, 0),
                                              ("mfma1", 20, 1)])
def test_dense_variant_matches_oracle(abi, configs, orc, lib, variant, nz, mask):
    """The VALU z operator and the one-wave-per-SIMD MFMA kernel (the default
    for odd NZ, forced here for even NZ too) against the oracle."""
    os.environ["SCFT_ZKIND"] = "dense"
    os.environ["SCFT_DENSE"] = variant
    try:
        p = configs.chem_droplet(lx=0.5, ly=1.0, lz=nz * 0.125, dr=0.125, ns=21, vol_dbk=0.2, vol_s1=0.05,
                                 len_s1=0.15)
        p.mask = mask
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        ref = orc.state(p)
        d = ctx.derived
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=5, amp=0.5, fields=act)
        for s in (ctx, ref):
            s.set_fields(w, pf)
            if p.mask:
                s.set_mask(configs.brush_mask(p))
        for it in (2, 3):
            a, o = ctx.iterate(it), ref.iterate(it)
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy)
        ctx.close()
    finally:
        del os.environ["SCFT_ZKIND"]
        del os.environ["SCFT_DENSE"]
