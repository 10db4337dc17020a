"""Launches of one contour step and axis position side by side: a batch of
uneven NX has one x launch (or one fused film launch) per length, on
disjoint chains, and the iteration runs them on side streams forked from
and joined back into the batch's stream (SCFT_PARX=0: one after another).

- the same kernels either way, so results are bit-identical: a batch of
  NX 192, 256 and 320 in FP64 and FP32-split, split step and film step,
  three iterations (F of every slot, the fields of every slot);
- the FP64 results also match the oracle (the C4 sweep fixtures run in
  test_gpu_configs.py with the default, side by side).
"""
import numpy as np
import pytest

from _cases import env as _env

pytestmark = pytest.mark.gpu


def _run(abi, configs, lib, protos, precision, parx, film):
    # SCFT_XMIX=0: one x launch per length (FP32-split otherwise runs the
    # three lengths as one launch, tests/test_gpu_xmix.py)
    with _env(SCFT_PARX=parx, SCFT_ZKIND="dense", SCFT_FILM="1" if film else "0", SCFT_XFFT="reg",
              SCFT_XMIX="0"):
        b = abi.Batch(lib, protos[0], len(protos), max_nx=320, precision=precision)
    for s, p in enumerate(protos):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=6.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    F = []
    for it in (2, 3, 4):
        F.append([o.free_energy for o in b.iterate([it] * len(protos))])
    sched = b.schedule()
    fields = [b.get_fields(s)[0] for s in range(len(protos))]
    b.close()
    return sched, np.array(F), fields


@pytest.mark.parametrize("film", [False, True])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_side_by_side_launches_are_bit_identical(abi, configs, lib, prec, film):
    precision = abi.SCFT_FP64 if prec == "fp64" else abi.SCFT_FP32_SPLIT
    protos = [configs.chem_droplet(lx=lx, chi_ab=20.0 + s, chi_ac=45.0 - s)
              for s, lx in enumerate((24.0, 32.0, 40.0, 40.0, 24.0, 32.0))]
    assert sorted({abi.grid_of(p)[0] for p in protos}) == [192, 256, 320]
    sa, Fa, wa = _run(abi, configs, lib, protos, precision, None, film)
    sb, Fb, wb = _run(abi, configs, lib, protos, precision, "0", film)
    print(sa, "|", sb, flush=True)
    n = dict(kv.split("=") for kv in sa.split() if "=" in kv and kv[0] != "z")
    if film:   # one film launch per step and x length
        assert int(n["film"]) > 0 and int(n["film"]) % 3 == 0, sa
    else:      # three x launches per z launch
        assert int(n["fftreg"]) + int(n["fftreg2"]) == 3 * int(n["dense_mfma2"]) > 0, sa
    assert np.array_equal(Fa, Fb), np.abs(Fa - Fb).max()
    for x, y in zip(wa, wb):
        assert np.array_equal(x, y)
