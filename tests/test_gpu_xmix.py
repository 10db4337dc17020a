"""One x launch for the lengths 192, 256 and 320 of a batch of uneven NX
(k_fftmix, the default for FP32 slices; SCFT_XMIX=0 keeps one launch per
length).

- against one launch per length on C2-size films of NX 192/256/320: the
  same per-length bodies (bit-identical on the first run), asserted to
  1e-10 kT in F and 1e-9 of the fields' largest value over three
  iterations, since the compiler may contract a product into an FMA
  differently in the mixed kernel; the schedule has one x launch per z
  launch;
- against the oracle: every slot of a thin uneven batch (NX 192/256/320,
  NZ 16) within 1e-6 kT in F, the FP32-split bar, over two iterations;
- FP64 slices keep one launch per length.
The full 192-slot C4 batch runs the mixed launch in
test_gpu_configs.py::test_c4_sweep_batch_matches_golden (fp32, split).
"""
import numpy as np
import pytest

from _cases import env as _env

pytestmark = pytest.mark.gpu


def _counts(sched):
    return {k: int(v) for k, v in (kv.split("=") for kv in sched.split() if "=" in kv and kv[0] != "z")}


def _rel(a, b):
    return float(np.abs(a - b).max() / max(1e-300, np.abs(b).max()))


def _batch(abi, configs, lib, protos, precision, mix, iters):
    with _env(SCFT_XFFT="reg", SCFT_FILM="0", SCFT_ZKIND="dense", SCFT_XMIX=(None if mix else "0")):
        b = abi.Batch(lib, protos[0], len(protos), max_nx=max(abi.grid_of(p)[0] for p in protos),
                      precision=precision)
    for s, p in enumerate(protos):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    F = [[o.free_energy for o in b.iterate([it] * len(protos))] for it in iters]
    sched = b.schedule()
    fields = [b.get_fields(s)[0] for s in range(len(protos))]
    b.close()
    return sched, np.array(F), fields


def _uneven(configs):
    return [configs.chem_droplet(lx=lx, chi_ab=20.0 + s, chi_ac=45.0 - s)
            for s, lx in enumerate((24.0, 32.0, 40.0, 40.0, 24.0, 32.0))]


def test_mixed_x_launch_matches_per_length_launches(abi, configs, lib):
    protos = _uneven(configs)
    assert sorted({abi.grid_of(p)[0] for p in protos}) == [192, 256, 320]
    sa, Fa, wa = _batch(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, True, (2, 3, 4))
    sb, Fb, wb = _batch(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, False, (2, 3, 4))
    print(sa, "|", sb, flush=True)
    na, nb = _counts(sa), _counts(sb)
    assert na["fftmix"] == na["dense_mfma2"] > 0 and na["fftreg"] == na["fftreg2"] == 0, sa
    assert nb["fftmix"] == 0 and nb["fftreg"] + nb["fftreg2"] == 3 * nb["dense_mfma2"], sb
    dF = float(np.abs(Fa - Fb).max())
    dw = max(_rel(x, y) for x, y in zip(wa, wb))
    print("mixed vs per-length x launches: max |dF| %.3e kT, fields rel %.3e" % (dF, dw), flush=True)
    assert dF <= 1e-10 and dw <= 1e-9, (dF, dw)


def _thin(configs):
    return [configs.chem_droplet(lx=lx, lz=2.0, ns=11, vol_s1=0.05, len_s1=0.3, chi_ab=22.0 + s)
            for s, lx in enumerate((24.0, 32.0, 40.0))]


def test_mixed_x_launch_matches_oracle(abi, configs, orc, lib):
    protos = _thin(configs)
    assert [abi.grid_of(p) for p in protos] == [(192, 1, 16), (256, 1, 16), (320, 1, 16)]
    sched, F, fields = _batch(abi, configs, lib, protos, abi.SCFT_FP32_SPLIT, True, (2, 3))
    assert _counts(sched)["fftmix"] > 0, sched
    for s, p in enumerate(protos):
        ref = orc.state(p)
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        ref.set_fields(w, pf)
        ref.set_mask(configs.brush_mask(p))
        for k, it in enumerate((2, 3)):
            o = ref.iterate(it)
            print("slot %d NX %d iteration %d: |dF| %.3e kT" % (s, abi.grid_of(p)[0], it, abs(F[k][s] - o.free_energy)),
                  flush=True)
            assert abs(F[k][s] - o.free_energy) <= 1e-6, (s, it, F[k][s], o.free_energy)
        wr, _ = ref.get_fields()
        assert _rel(fields[s], wr) <= 1e-5


def test_fp64_keeps_per_length_launches(abi, configs, lib):
    sched, _, _ = _batch(abi, configs, lib, _thin(configs), abi.SCFT_FP64, True, (2,))
    n = _counts(sched)
    nx = n["fftreg"] + n["fftreg2"]
    assert n["fftmix"] == 0 and nx > 0 and nx % 3 == 0, sched
