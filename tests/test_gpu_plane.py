"""The fused contour step on periodic 64 x 64 planes (k_plane64: both axis
transforms of a step in one workgroup per chain) against the oracle.

- the C1 bulk melt and a bulk blend with four chain types (diblock q and
  q-dagger across the A/B junction, homopolymers A and C), FP64 to 1e-10
  and FP32-split to 1e-6 kT in F (1e-5 in Q and the fields), with the z operator forced to the transform;
- a batch of three slots with different chi through the fused step against
  the same batch with the step split into its z and x launches;
- SCFT_PLANE=0 turns the fused step off; the split step then picks the
  wave-per-line FFT for a launch of few lines and the register FFT for a
  batch.
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


def _blend(abi, chi_ab=20.0, chi_ac=30.0):
    """Bulk 64 x 64 blend: AB diblock (f_A 0.4), homopolymers A and C."""
    vh = [0.0] * 10
    lh = [1.0] * 10
    vh[0], lh[0] = 0.2, 0.5
    vh[2] = 0.5
    return abi.make_params(lx=8.0, ly=1.0, lz=8.0, dr_target=0.125, ns=41, mask=0, g1=0.1,
                           chi={"AB": chi_ab, "AC": chi_ac, "BC": chi_ac},
                           f_dbk=(0.4, 0.5, 0.3, 0.5), vol_dbk=(0.3, 0, 0, 0),
                           vol_homo=tuple(vh), len_homo=tuple(lh))


def _plane_count(sched):
    n = dict(kv.split("=") for kv in sched.split() if "=" in kv and kv[0] != "z")
    return int(n["plane"])


@pytest.fixture
def z_fft():
    os.environ["SCFT_ZKIND"] = "fft"
    yield
    del os.environ["SCFT_ZKIND"]


@pytest.mark.parametrize("case", ["c1", "blend"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_plane_step_matches_oracle(abi, configs, orc, lib, z_fft, case, prec):
    p = configs.params_c1() if case == "c1" else _blend(abi)
    ctx = abi.Context(lib, p, abi.SCFT_FP64 if prec == "fp64" else abi.SCFT_FP32_SPLIT)
    assert ctx.shape == (64, 1, 64)
    ref = orc.state(p)
    d = ctx.derived
    act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
    w, pf = configs.random_fields(p, seed=3, amp=0.5, fields=act)
    for s in (ctx, ref):
        s.set_fields(w, pf)
    tol = TOL if prec == "fp64" else 1e-5
    for it in (2, 3, 4):
        a, o = ctx.iterate(it), ref.iterate(it)
        print(case, prec, it, "dF = %.3e" % abs(a.free_energy - o.free_energy), flush=True)
        if prec == "fp64":
            assert _srel(a.free_energy, o.free_energy) <= tol, (a.free_energy, o.free_energy)
        else:   # the FP32-split accuracy target: 1e-6 kT
            assert abs(a.free_energy - o.free_energy) <= 1e-6, (a.free_energy, o.free_energy)
        for sp in range(17):
            assert _srel(a.qsum[sp], o.qsum[sp]) <= tol, (sp, a.qsum[sp], o.qsum[sp])
    assert _plane_count(ctx.schedule()) > 0, ctx.schedule()
    wa, pa = ctx.get_fields()
    wb, pb = ref.get_fields()
    assert _rel(wa, wb) <= tol and _rel(pa, pb) <= tol
    ctx.close()


def _run_batch(abi, configs, lib, ps, fused):
    os.environ["SCFT_PLANE"] = "1" if fused else "0"
    try:
        b = abi.Batch(lib, ps[0], len(ps), precision=abi.SCFT_FP64)
    finally:
        del os.environ["SCFT_PLANE"]
    for s, p in enumerate(ps):
        d = b.init_slot(s, p)
        act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
        w, pf = configs.random_fields(p, seed=10 + s, amp=0.5, fields=act)
        b.set_fields(s, w, pf)
        b.activate(s)
    hist = [[o.free_energy for o in b.iterate()] for _ in range(3)]
    sched = b.schedule()
    fields = [b.get_fields(s)[0] for s in range(len(ps))]
    b.close()
    return hist, sched, fields


def test_plane_batch_matches_split_step(abi, configs, lib, z_fft):
    ps = [_blend(abi, 20.0, 30.0), _blend(abi, 15.0, 35.0), _blend(abi, 25.0, 25.0)]
    hf, sf, wf = _run_batch(abi, configs, lib, ps, True)
    hs, ss, ws = _run_batch(abi, configs, lib, ps, False)
    assert _plane_count(sf) > 0 and _plane_count(ss) == 0, (sf, ss)
    for a, b in zip(hf, hs):
        for fa, fb in zip(a, b):
            assert _srel(fa, fb) <= 1e-12, (fa, fb)
    for a, b in zip(wf, ws):
        assert _rel(a, b) <= 1e-12


def _kinds(sched):
    return {k: int(v) for k, v in (kv.split("=") for kv in sched.split() if "=" in kv and kv[0] != "z")}


def test_split_step_kind_follows_launch_size(abi, configs, lib, z_fft):
    """With the fused step off, a C1 step's 64-point transforms go to the
    wave-per-line FFT when the launch holds few lines (one solve) and to the
    8 x 8 register FFT when it holds a batch (192 solves)."""
    os.environ["SCFT_PLANE"] = "0"
    try:
        p = configs.params_c1()
        one = abi.Batch(lib, p, 1, precision=abi.SCFT_FP32_SPLIT)
        many = abi.Batch(lib, p, 192, precision=abi.SCFT_FP32_SPLIT)
    finally:
        del os.environ["SCFT_PLANE"]
    w, pf = configs.random_fields(p, seed=1, amp=0.3, fields=["A", "B"])
    F = []
    for b in (one, many):
        for s in range(b.nslots):
            b.init_slot(s, p)
            b.set_fields(s, w, pf)
            b.activate(s)
        F.append([o.free_energy for o in b.iterate()])
    k1, kb = _kinds(one.schedule()), _kinds(many.schedule())
    assert k1["plane"] == 0 and kb["plane"] == 0
    assert k1["fftwave"] > 0 and k1["fftreg2"] == 0, one.schedule()
    assert kb["fftreg2"] > 0 and kb["fftwave"] == 0, many.schedule()
    # the 192 slots hold the same solve: identical results, and the single
    # solve through the other transform within FP32-split rounding
    assert all(f == F[1][0] for f in F[1])
    assert _srel(F[0][0], F[1][0]) <= 1e-6, (F[0][0], F[1][0])
    one.close()
    many.close()
