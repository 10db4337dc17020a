"""The autotuner's choice of the periodic last-axis form (k_fftx against
the register or wave-per-line FFT, per size class of the active set).

- with SCFT_XFFT unset and a grid whose x length k_fftx takes, creation
  times both forms (SCFT_TUNE_LOG=1 prints every timing), and the tuned
  schedule matches the oracle to 1e-10 in FP64;
- SCFT_XFFT pins the form: no k_fftx timing, and the schedule follows it;
- a grid with no k_fftx length (NX 96) times no k_fftx form.

Each case uses its own contour length, so that no earlier context's
tuning result (cached per shape for the process) is reused.
"""
import numpy as np
import pytest

from _cases import env as _env

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _srel(a, b):
    return abs(a - b) / max(abs(b), 1.0)


def _tune_lines(err):
    return [ln for ln in err.splitlines() if ln.startswith("scft autotune:")]


def _params(configs, nx, ns):
    """A 2D film nx x 1 x 64 (64 x lines: whole 32-line groups)."""
    return configs.chem_droplet(lx=nx * 0.25, lz=64 * 0.25, dr=0.25, ns=ns, vol_dbk=0.2, vol_s1=0.05, len_s1=0.3)


def test_autotuner_times_the_last_axis_forms(abi, configs, orc, lib, capfd):
    p = _params(configs, 128, 13)
    with _env(SCFT_TUNE_LOG="1", SCFT_XFFT=None, SCFT_ZKIND=None, SCFT_FILM=None):
        b = abi.Batch(lib, p, 4, precision=abi.SCFT_FP64)
    lines = _tune_lines(capfd.readouterr().err)
    print("\n".join(lines), flush=True)
    assert any("k_fftx 1" in ln for ln in lines), lines
    assert any("k_fftx 0" in ln for ln in lines), lines
    ref = orc.state(p)
    w, pf = configs.droplet_fields(p, radius=3.0, seed=2)
    for s in range(4):
        b.init_slot(s, p)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    ref.set_fields(w, pf)
    ref.set_mask(configs.brush_mask(p))
    for it in (2, 3):
        outs, o = b.iterate([it] * 4), ref.iterate(it)
        print(it, b.schedule(), "F %.15f" % o.free_energy, flush=True)
        for a in outs:
            assert _srel(a.free_energy, o.free_energy) <= TOL, (a.free_energy, o.free_energy, b.schedule())
            for sp in range(17):
                assert _srel(a.qsum[sp], o.qsum[sp]) <= TOL
    wa, _ = b.get_fields(3)
    wb, _ = ref.get_fields()
    assert np.abs(wa - wb).max() <= TOL * max(np.abs(wb).max(), 1.0)
    b.close()


@pytest.mark.parametrize("form,kind", [("reg", "fftreg"), ("coal", "fftx")])
def test_pinned_last_axis_form_is_not_tuned(abi, configs, lib, capfd, form, kind):
    p = _params(configs, 128, 15 if form == "reg" else 17)
    with _env(SCFT_TUNE_LOG="1", SCFT_XFFT=form, SCFT_ZKIND="dense", SCFT_FILM="0"):
        b = abi.Batch(lib, p, 4, precision=abi.SCFT_FP32_SPLIT)
    lines = _tune_lines(capfd.readouterr().err)
    assert not any("k_fftx 1" in ln for ln in lines), lines
    for s in range(4):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        b.set_fields(s, w, pf)
        b.activate(s)
    b.iterate([2] * 4)
    n = dict(kv.split("=") for kv in b.schedule().split() if "=" in kv and kv[0] != "z")
    assert int(n[kind]) > 0, b.schedule()
    b.close()


def test_no_k_fftx_timing_without_its_lengths(abi, configs, lib, capfd):
    p = _params(configs, 96, 19)
    with _env(SCFT_TUNE_LOG="1", SCFT_XFFT=None, SCFT_ZKIND=None, SCFT_FILM=None):
        b = abi.Batch(lib, p, 4, precision=abi.SCFT_FP32_SPLIT)
    lines = _tune_lines(capfd.readouterr().err)
    print("\n".join(lines), flush=True)
    assert lines and not any("k_fftx 1" in ln for ln in lines), lines
    b.close()
