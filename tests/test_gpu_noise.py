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


def test_batch_slot_streams_match_single_contexts(abi, configs, orc, lib):
    """Every slot of a batch with its own stream (different seeds and
    RELAXATION) follows the single-instance context with the same seed, bit
    for bit, and ends at the oracle generator's state."""
    ps = [configs.t_film_small(), configs.t_film_small(), configs.t_film_small()]
    chis = [(25.0, 40.0), (18.0, 30.0), (30.0, 45.0)]
    relax = [2, 3, 0]
    seeds = [(1, 2, 3), (7, 8, 9), (0x330E, 0xABCD, 0x1234)]
    for p, (a, c), r in zip(ps, chis, relax):
        p.chi[0][1] = p.chi[1][0] = a
        p.chi[0][2] = p.chi[2][0] = c
        p.relaxation = r
        p.dt_noise = 0.1
    b = abi.Batch(lib, ps[0], 3, max_nx=12, precision=abi.SCFT_FP64)
    singles, n = [], []
    for s, p in enumerate(ps):
        w, pf = configs.droplet_fields(p, radius=3.0, seed=s)
        m = configs.brush_mask(p)
        d = b.init_slot(s, p)
        n.append(d.n_active * b.V[s])
        b.set_fields(s, w, pf)
        b.set_mask(s, m)
        b.rng_seed(s, seeds[s])
        b.activate(s)
        c = abi.Context(lib, p, abi.SCFT_FP64)
        c.set_fields(w, pf)
        c.set_mask(m)
        c.rng_seed(seeds[s])
        singles.append(c)
    gens = []
    for s in range(3):
        g = orc.state(configs.t_film_small())
        g.seed48(seeds[s])
        gens.append(g)
    for it in range(1, 6):
        outs = b.iterate([it, it, it])
        for s, c in enumerate(singles):
            o = c.iterate(it, 1.0 if it <= relax[s] else 0.0)
            assert outs[s].free_energy == o.free_energy, (it, s, outs[s].free_energy, o.free_energy)
            assert outs[s].mf == (1.0 if it <= relax[s] else 0.0)
            gens[s].gauss_n(n[s])
            assert b.rng_state(s) == gens[s].get_seed() == c.rng_state(), (it, s)
    for s, c in enumerate(singles):
        wa, _ = b.get_fields(s)
        wb, _ = c.get_fields()
        assert np.array_equal(wa, wb), s
        c.close()
    b.close()


def test_batch_slot_within_relaxation_needs_a_stream(abi, configs, lib):
    p = configs.t_film_small()
    p.relaxation = 2
    b = abi.Batch(lib, p, 1, precision=abi.SCFT_FP64)
    b.init_slot(0, p)
    w, pf = configs.droplet_fields(p, radius=3.0)
    b.set_fields(0, w, pf)
    b.set_mask(0, configs.brush_mask(p))
    b.activate(0)
    with pytest.raises(abi.ScftError, match="RELAXATION"):
        b.iterate([1])
    b.iterate([3])            # past RELAXATION: MF = 0, no stream needed
    b.close()


def test_served_request_with_relaxation_matches_single_context(abi, configs, lib):
    """serve.Pool gives a request with RELAXATION > 0 its own stream (drand48's
    initial state by default) and numbers its iterations from 1: its
    free-energy trajectory equals a single context run the same way."""
    import serve
    p = configs.t_film_small()
    p.relaxation = 3
    p.dt_noise = 0.1
    w, pf = configs.droplet_fields(p, radius=3.0, seed=4)
    pool = serve.Pool(lib, p, precision=abi.SCFT_FP64, max_slots=2, tol=1e-30, max_iters=6, record_every=1)
    req = serve.Request("r", p, (serve.SEED_FIELDS, (w, pf)))
    pool.submit(req)
    pool.run()
    assert req.status == "max_iters"
    c = abi.Context(lib, p, abi.SCFT_FP64)
    c.set_fields(w, pf)
    c.set_mask(configs.brush_mask(p))
    c.rng_seed(serve.DRAND48_INITIAL)
    ref = [c.iterate(it, 1.0 if it <= 3 else 0.0).free_energy for it in range(1, 7)]
    assert [h[1] for h in req.history] == ref
    c.close()
    pool.close()
