"""The fused film step (k_film: the dense z operator and the x transform of
a 2D contour step in one launch, each workgroup a 32-column part of one
chain's plane, the scratch between the axes in LDS) against the split step
(k_zdense_mfma2 then the register x FFT) and against the oracle.

- equal to the split step up to rounding (the same operations; only the
  compiler's contraction of multiply-adds differs between the two kernels):
  a batch of C2 droplets (every factor mode: the start slice, E^2 inside a
  block, E E' at the diblock junction, the brush start from the mask), and
  a C4-style batch with NX 192, 256 and 320 in one batch (one film launch
  per x length), in FP64 and FP32-split, over three iterations: F, Q, the
  fields and stored slices of three chains;
- with the FP32 E^2 factor (SCFT_E2F=1) the same;
- against the oracle's FP64 fixtures: C2 in test_gpu_parity.py and the C4
  sweep batch in test_gpu_configs.py run with the film step forced too.
"""
import numpy as np
import pytest

from _cases import env as _env

pytestmark = pytest.mark.gpu


def _counts(sched):
    return dict((k, int(v)) for k, v in (kv.split("=") for kv in sched.split() if "=" in kv and kv[0] != "z"))


def _run(abi, configs, lib, protos, precision, film, iters=(2, 3, 4), e2f="0"):
    """Iterate a batch of the given slot parameters; return the outs per
    iteration, the fields and the stored propagator slices of slot 0."""
    # the register x FFT also for launches of few lines: the film step's x transform
    with _env(SCFT_ZKIND="dense", SCFT_FILM="1" if film else "0", SCFT_E2F=e2f, SCFT_XFFT="reg"):
        b = abi.Batch(lib, protos[0], len(protos), max_nx=max(abi.grid_of(p)[0] for p in protos), precision=precision)
    ds = []
    for s, p in enumerate(protos):
        ds.append(b.init_slot(s, p))
        w, pf = configs.droplet_fields(p, radius=6.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    outs = []
    for it in iters:
        o = b.iterate([it] * len(protos))
        outs.append([(x.free_energy, x.rmsdev, x.dw_rms, tuple(x.chempot)) for x in o])
    sched = b.schedule()
    fields = [b.get_fields(s) for s in range(len(protos))]
    d = ds[0]
    slices = {}
    for sp in (0, 4 + 2, 14):   # diblock AB (q and q-dagger), homopolymer C, brush S1
        ns = d.ns_sp[sp]
        if ns <= 0:
            continue
        for which in (0, 1):
            for idx in (1, ns // 2, ns - 1):
                slices[(sp, which, idx)] = b.get_propagator(0, sp, which, idx)
    b.close()
    return sched, outs, fields, slices


def _assert_same(a, b, what, fp32):
    """The film step against the split step: the same arithmetic up to how
    the compiler contracts multiply-adds in the two kernels, so FP64 values
    agree to a few ulps and an FP32-split slice value may land on the other
    FP32 neighbour."""
    sa, oa, fa, qa = a
    sb, ob, fb, qb = b
    assert _counts(sa)["film"] > 0, sa
    assert _counts(sb)["film"] == 0, sb
    qtol = 2.0 ** -23 if fp32 else 1e-14
    assert qa.keys() == qb.keys() and qa
    for k in sorted(qa):
        d = np.abs(qa[k] - qb[k]).max() / np.abs(qb[k]).max()
        print(what, "slice", k, "max rel diff %.3e" % d, flush=True)
        assert d <= qtol, (what, k, d)
    for it, (x, y) in enumerate(zip(oa, ob)):
        dF = max(abs(u[0] - v[0]) for u, v in zip(x, y))
        print(what, "iter", it, "max |dF| %.3e" % dF, flush=True)
        for u, v in zip(x, y):
            assert abs(u[0] - v[0]) <= 1e-12 * max(1.0, abs(v[0])), (what, it, u[0], v[0])
            assert abs(u[1] - v[1]) <= 1e-10 * abs(v[1]) and abs(u[2] - v[2]) <= 1e-10 * abs(v[2]), (what, it)
            for p, q in zip(u[3], v[3]):
                assert abs(p - q) <= 1e-12 * max(1.0, abs(q)), (what, it, p, q)
    for s, (x, y) in enumerate(zip(fa, fb)):
        for u, v in zip(x, y):
            assert np.abs(u - v).max() <= 1e-11 * max(np.abs(v).max(), 1.0), (what, s)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_film_step_equals_split_step_c2(abi, configs, lib, prec):
    precision = abi.SCFT_FP64 if prec == "fp64" else abi.SCFT_FP32_SPLIT
    protos = [configs.params_c2(chi_ab=25.0 + 2 * s, chi_ac=40.0 - 3 * s) for s in range(5)]
    a = _run(abi, configs, lib, protos, precision, film=True)
    b = _run(abi, configs, lib, protos, precision, film=False)
    print(a[0], "|", b[0], "| F", a[1][-1][0][0], flush=True)
    _assert_same(a, b, "c2 " + prec, prec == "fp32")


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_film_step_equals_split_step_mixed_nx(abi, configs, lib, prec):
    """NX 192, 256 and 320 in one batch: one film launch per x length."""
    precision = abi.SCFT_FP64 if prec == "fp64" else abi.SCFT_FP32_SPLIT
    protos = [configs.chem_droplet(lx=lx, chi_ab=20.0 + s, chi_ac=45.0 - s)
              for s, lx in enumerate((24.0, 32.0, 40.0, 40.0, 24.0, 32.0))]
    assert sorted({abi.grid_of(p)[0] for p in protos}) == [192, 256, 320]
    a = _run(abi, configs, lib, protos, precision, film=True, iters=(2, 3))
    b = _run(abi, configs, lib, protos, precision, film=False, iters=(2, 3))
    assert _counts(a[0])["film"] == 3 * _counts(b[0])["dense_mfma2"], (a[0], b[0])
    _assert_same(a, b, "mixed nx " + prec, prec == "fp32")


def test_film_step_equals_split_step_fp32_factor(abi, configs, lib):
    """The FP32 E^2 factor of the z operator (SCFT_E2F=1) in both forms."""
    protos = [configs.params_c2(chi_ab=25.0 + s) for s in range(3)]
    a = _run(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, film=True, iters=(2, 3), e2f="1")
    b = _run(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, film=False, iters=(2, 3), e2f="1")
    _assert_same(a, b, "e2f", True)


def test_film_timing_reports_the_fused_step(abi, configs, lib):
    """scft_time_launches kind 0 times the film launch: 16 bytes per
    voxel-step in FP32-split (slice 4 + factor 8 in, slice 4 out)."""
    p = configs.params_c2()
    with _env(SCFT_ZKIND="dense", SCFT_FILM="1"):
        b = abi.Batch(lib, p, 8, precision=abi.SCFT_FP32_SPLIT)
    for s in range(8):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=6.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    b.iterate([2] * 8)
    t = b.time_launches(0, reps=3)
    print(b.schedule(), t, flush=True)
    assert t["launches"] == _counts(b.schedule())["film"]
    assert t["avg_us"] > 0.0
    bpv = t["bytes"] / t["voxel_steps"]
    assert 16.0 <= bpv <= 16.0 * (1 + 1e-3) + 8.0 / 50   # a block junction step reads a second factor
    b.close()
