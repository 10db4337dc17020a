"""Periodic last and middle axes through LDS in groups of 32 lines (k_fftx,
SCFT_XFFT=coal: coalesced row-segment loads and stores, the film step's
x transform in between) against the oracle.

- 2D films with x lengths 128, 192, 256 and 320 and a 3D grid whose y
  lines (length 128, a middle axis) go through it: FP64 to 1e-10 in F, Q
  and the fields; FP32-split to 1e-6 kT in F;
- C2 at full size (256x1x128) against the oracle's FP64 fixture.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _srel(a, b):
    return abs(a - b) / max(abs(b), 1.0)


def _rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _fftx(sched):
    return int(sched.split("fftx=")[1].split()[0])


@pytest.fixture
def coal():
    old = os.environ.get("SCFT_XFFT")
    os.environ["SCFT_XFFT"] = "coal"
    yield
    if old is None:
        os.environ.pop("SCFT_XFFT", None)
    else:
        os.environ["SCFT_XFFT"] = old


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("nx,ny,nz", [(256, 1, 32), (192, 1, 32), (320, 1, 32), (128, 1, 64), (8, 128, 32)])
def test_fftx_axes_match_oracle(abi, configs, orc, lib, coal, prec, nx, ny, nz):
    p = configs.chem_droplet(lx=nx * 0.25, ly=(ny * 0.25 if ny > 1 else 1.0), lz=nz * 0.25, dr=0.25, ns=11,
                             vol_dbk=0.2, vol_s1=0.05, len_s1=0.3)
    ctx = abi.Context(lib, p, abi.SCFT_FP64 if prec == "fp64" else abi.SCFT_FP32_SPLIT)
    assert ctx.shape == (nx, ny, nz)
    ref = orc.state(p)
    d = ctx.derived
    act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
    w, pf = configs.random_fields(p, seed=5, amp=0.5, fields=act)
    for s in (ctx, ref):
        s.set_fields(w, pf)
        s.set_mask(configs.brush_mask(p))
    for it in (2, 3):
        a, o = ctx.iterate(it), ref.iterate(it)
        print((nx, ny, nz), prec, it, "dF = %.3e" % abs(a.free_energy - o.free_energy), ctx.schedule(), flush=True)
        if prec == "fp64":
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy)
            for sp in range(17):
                assert _srel(a.qsum[sp], o.qsum[sp]) <= TOL, (sp, a.qsum[sp], o.qsum[sp])
        else:
            assert abs(a.free_energy - o.free_energy) <= 1e-6, (a.free_energy, o.free_energy)
    assert _fftx(ctx.schedule()) > 0, ctx.schedule()
    wa, _ = ctx.get_fields()
    wb, _ = ref.get_fields()
    assert _rel(wa, wb) <= (TOL if prec == "fp64" else 1e-5)
    ctx.close()


def test_fftx_c2_full_size_matches_golden(abi, configs, lib, coal):
    """C2 (256x1x128, MASK 1, NS 101), the dense z operator and k_fftx for
    x, against the oracle's FP64 fixture: three iterations, 1e-10."""
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "c2_oracle.json")) as f:
        g = json.load(f)["c2_droplet_seed0"]
    old = {k: os.environ.get(k) for k in ("SCFT_ZKIND", "SCFT_FILM")}
    os.environ["SCFT_ZKIND"] = "dense"
    os.environ["SCFT_FILM"] = "0"
    try:
        p = configs.params_c2()
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
        ctx.set_fields(w, pf)
        ctx.set_mask(configs.brush_mask(p))
        for row in g["iters"]:
            o = ctx.iterate(row["iter"])
            assert _srel(o.free_energy, row["F"]) <= TOL, (row["iter"], o.free_energy, row["F"])
            assert _srel(o.dw_rms, row["dw_rms"]) <= TOL
            for sp in range(17):
                assert _srel(o.chempot[sp], row["chempot"][sp]) <= TOL
        assert _fftx(ctx.schedule()) > 0, ctx.schedule()
        wf, _ = ctx.get_fields()
        idx = g["sample_idx"]
        for f, vals in g["w_samples"].items():
            assert _rel(wf[int(f)][idx], vals) <= TOL
        ctx.close()
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
