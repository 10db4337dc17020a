"""The noise stream drawn on the GPU (scft_rng_seed) against the oracle's
drand48 + polar-method replay of funGaussDev (SCFT.c:1769-1782), drawn once
per active field and voxel every iteration (SCFT.c:2103).

- whole iterations with MF = 1 and MF = 0 against the oracle: F, fields and
  the generator state after each iteration (bit-exact);
- the update alone at sizes where the candidate pairs span thousands of
  chunks (up to 4.2 M deviates per iteration): fields against the same
  update fed the oracle's deviates, state bit-exact;
- a shortfall of accepted pairs, forced from the host with a negative
  margin, is reported as an error and leaves the state where it was.
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


def test_device_noise_matches_oracle(abi, configs, orc, lib):
    p = configs.t_film_small()
    p.relaxation = 3        # MF = 1 on iterations 2 and 3, then 0
    p.dt_noise = 0.1
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    ref = orc.state(p)
    seed = (11, 22, 33)
    ref.seed48(seed)
    ctx.rng_seed(seed)
    assert ctx.rng_state() == seed
    w, pf = configs.droplet_fields(p, radius=3.0)
    for s in (ctx, ref):
        s.set_fields(w, pf)
        s.set_mask(configs.brush_mask(p))
    for it, mf in ((2, 1.0), (3, 1.0), (4, 0.0), (5, 0.0)):
        a = ctx.iterate(it, mf)
        b = ref.iterate(it)
        assert b.mf == mf
        assert abs(a.free_energy - b.free_energy) <= TOL * max(1.0, abs(b.free_energy))
        wa, pa = ctx.get_fields()
        wb, pb = ref.get_fields()
        assert _rel(wa, wb) <= TOL and _rel(pa, pb) <= TOL
        assert ctx.rng_state() == ref.get_seed(), it
    ctx.close()


@pytest.mark.parametrize("nx,ny,nz", [(6, 1, 6), (256, 1, 128), (128, 128, 64)])
def test_device_deviates_match_host_replay(abi, configs, orc, lib, nx, ny, nz):
    """Update-only calls: one context draws on the device, one is fed the
    oracle's deviates; the fields agree to rounding of the device log and
    the states bit for bit, over two iterations."""
    p = configs.chem_droplet(lx=nx * 0.25, ly=(ny * 0.25 if ny > 1 else 1.0), lz=nz * 0.25, dr=0.25, ns=11,
                             vol_dbk=0.2, vol_s1=0.05, len_s1=0.3)
    p.dt_noise = 1.0
    dev = abi.Context(lib, p, abi.SCFT_FP64)
    host = abi.Context(lib, p, abi.SCFT_FP64)
    assert dev.shape == (nx, ny, nz)
    gen = orc.state(configs.t_film_small())
    seed = (0x1234, 0xabcd, 0x0042)
    gen.seed48(seed)
    dev.rng_seed(seed)
    d = dev.derived
    act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
    w, pf = configs.random_fields(p, seed=4, amp=0.5, fields=act)
    for c in (dev, host):
        c.set_fields(w, pf)
    n = d.n_active * dev.V
    print("deviates per iteration", n, flush=True)
    for it in (2, 3):
        noise = gen.gauss_n(n)
        dev.update_fields(it, 1.0)
        host.update_fields(it, 1.0, noise)
        wd, pd = dev.get_fields()
        wh, ph = host.get_fields()
        assert np.abs(wd - wh).max() <= 1e-12, it
        assert np.abs(pd - ph).max() <= 1e-12, it
        assert dev.rng_state() == gen.get_seed(), it
    # the noise was applied: a third context without it differs by O(1)
    quiet = abi.Context(lib, p, abi.SCFT_FP64)
    quiet.set_fields(w, pf)
    quiet.update_fields(2, 0.0)
    quiet.update_fields(3, 0.0)
    assert np.abs(quiet.get_fields()[0] - wd).max() > 0.1
    for c in (dev, host, quiet):
        c.close()


def test_device_noise_shortfall_is_an_error(abi, configs, lib):
    p = configs.t_film_small()
    p.relaxation = 10
    os.environ["SCFT_RNG_MARGIN"] = "-20"
    try:
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        ctx.rng_seed((1, 2, 3))
    finally:
        del os.environ["SCFT_RNG_MARGIN"]
    w, pf = configs.droplet_fields(p, radius=3.0)
    ctx.set_fields(w, pf)
    ctx.set_mask(configs.brush_mask(p))
    with pytest.raises(abi.ScftError, match="short of accepted pairs"):
        ctx.iterate(2, 1.0)
    assert ctx.rng_state() == (1, 2, 3)
    ctx.close()


def test_host_noise_rejected_while_device_stream_on(abi, configs, lib):
    p = configs.t_film_small()
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    ctx.rng_seed((5, 6, 7))
    with pytest.raises(abi.ScftError, match="invalid argument"):
        ctx.iterate(2, 1.0, np.zeros(ctx.derived.n_active * ctx.V))
    ctx.close()
