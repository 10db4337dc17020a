"""The MFMA z operator with the next tile's slice rows in flight
(k_zdense_mfma2p, the default for FP32 slices of NZ 48, 64 and 128) against
the form that loads each tile when it starts (SCFT_ZPF=0): the same
arithmetic, so results are bit-identical.

- C2 films (NZ 128) and 3D grids of NZ 64 and 48 (rows shorter than the
  operator's padded width) with uneven NX, FP32-split, three iterations:
  F of every slot and every slot's fields;
- FP64 slices keep the plain form (the schedule reports no prefetching
  launches).
The prefetching form also runs in every FP32-split parity test against the
oracle's fixtures (C2, C3, C4).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _counts(sched):
    return {k: int(v) for k, v in (kv.split("=") for kv in sched.split() if "=" in kv and kv[0] != "z")}


def _run(abi, configs, lib, protos, precision, zpf):
    old = {k: os.environ.get(k) for k in ("SCFT_ZPF", "SCFT_ZKIND", "SCFT_FILM")}
    os.environ.update({"SCFT_ZKIND": "dense", "SCFT_FILM": "0"})
    if zpf:
        os.environ.pop("SCFT_ZPF", None)
    else:
        os.environ["SCFT_ZPF"] = "0"
    try:
        b = abi.Batch(lib, protos[0], len(protos), max_nx=max(abi.grid_of(p)[0] for p in protos),
                      precision=precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    for s, p in enumerate(protos):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    F = [[o.free_energy for o in b.iterate([it] * len(protos))] for it in (2, 3, 4)]
    sched = b.schedule()
    fields = [b.get_fields(s)[0] for s in range(len(protos))]
    b.close()
    return sched, np.array(F), fields


CASES = {
    "c2": lambda configs: [configs.params_c2(chi_ab=25.0 + s) for s in range(3)],
    "nz64": lambda configs: [configs.chem_droplet(lx=lx, ly=4.0, lz=16.0, dr=0.25, ns=21, vol_dbk=0.2, vol_s1=0.05,
                                                  len_s1=0.3) for lx in (4.0, 6.0, 4.0)],
    "nz48": lambda configs: [configs.chem_droplet(lx=lx, ly=4.0, lz=12.0, dr=0.25, ns=21, vol_dbk=0.2, vol_s1=0.05,
                                                  len_s1=0.3) for lx in (4.0, 6.0, 4.0)],
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_prefetching_z_operator_is_bit_identical(abi, configs, lib, case):
    protos = CASES[case](configs)
    assert abi.grid_of(protos[0])[2] in (48, 64, 128)
    sa, Fa, wa = _run(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, True)
    sb, Fb, wb = _run(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, False)
    print(sa, "|", sb, flush=True)
    na, nb = _counts(sa), _counts(sb)
    assert na["mfma2_pf"] == na["dense_mfma2"] > 0 and nb["mfma2_pf"] == 0, (sa, sb)
    assert np.array_equal(Fa, Fb), np.abs(Fa - Fb).max()
    for x, y in zip(wa, wb):
        assert np.array_equal(x, y)


def test_fp64_slices_keep_the_plain_form(abi, configs, lib):
    protos = CASES["c2"](configs)[:1]
    sched, F, _ = _run(abi, configs, lib, protos, abi.SCFT_FP64, True)
    n = _counts(sched)
    assert n["dense_mfma2"] > 0 and n["mfma2_pf"] == 0, sched
    assert np.isfinite(F).all()
