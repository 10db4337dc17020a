"""libscft on the GPU against the CPU oracle and against closed forms.

FP64 path: every per-iteration scalar (F, Q_p, -ln Q_p, RMSDev, ||dw||),
the updated fields and the densities agree with the oracle to 1e-10
relative.  FP32-split path: |dF| against the FP64 oracle is reported and
bounded.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _init(configs, abi, p, kind, seed=0):
    d = abi.ScftDerived()
    import ctypes as C
    lib = abi.load_library()
    assert lib.scft_derive(C.byref(p), C.byref(d)) == 0
    act = [abi.FIELD_NAMES[f] for f in range(13) if d.active[f]]
    if kind == "drop":
        return configs.droplet_fields(p, radius=3.0, seed=seed)
    return configs.random_fields(p, seed=seed, amp=0.4, fields=act)


CASES = {
    "bulk8": ("t_bulk_small", {}, "rand"),
    "film12x10": ("t_film_small", {}, "drop"),
    "film3d": ("t_3d_small", {}, "rand"),
    "all_species_mask1": ("t_all_species", {"mask": 1}, "rand"),
    "all_species_mask0": ("t_all_species", {"mask": 0}, "rand"),
    "c1_64x64": ("params_c1", {}, "rand"),
}


def _rel(a, b):
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def _scalar_rel(a, b):
    return abs(a - b) / max(abs(b), 1.0)


def _run_pair(abi, configs, orc, lib, name, iters=(2, 3, 4), precision=None):
    fn, kw, kind = CASES[name]
    p = getattr(configs, fn)(**kw)
    print("case", name, flush=True)
    ctx = abi.Context(lib, p, abi.SCFT_FP64 if precision is None else precision)
    ref = orc.state(p)
    w, pf = _init(configs, abi, p, kind)
    for s in (ctx, ref):
        s.set_fields(w, pf)
        if p.mask:
            s.set_mask(configs.brush_mask(p))
    outs = []
    for it in iters:
        outs.append((ctx.iterate(it), ref.iterate(it)))
    return ctx, ref, outs, p


@pytest.mark.parametrize("name", list(CASES))
def test_fp64_iterations_match_oracle(abi, configs, orc, lib, name):
    ctx, ref, outs, p = _run_pair(abi, configs, orc, lib, name)
    for a, b in outs:
        assert _scalar_rel(a.free_energy, b.free_energy) <= TOL, (a.free_energy, b.free_energy)
        assert _scalar_rel(a.rmsdev, b.rmsdev) <= TOL
        assert _scalar_rel(a.dw_rms, b.dw_rms) <= TOL
        for sp in range(17):
            assert _scalar_rel(a.qsum[sp], b.qsum[sp]) <= TOL, (sp, a.qsum[sp], b.qsum[sp])
            assert _scalar_rel(a.chempot[sp], b.chempot[sp]) <= TOL, (sp, a.chempot[sp], b.chempot[sp])
    wa, pa = ctx.get_fields()
    wb, pb = ref.get_fields()
    assert _rel(wa, wb) <= TOL
    assert _rel(pa, pb) <= TOL
    fa, ta = ctx.get_densities()
    fb, tb = ref.get_densities()
    assert _rel(fa, fb) <= TOL
    assert _rel(ta, tb) <= TOL
    # every propagator slice of every species
    d = ctx.derived
    for sp in range(17):
        for s in range(d.ns_sp[sp]):
            for which in ((0,) if 4 <= sp < 14 else (0, 1)):
                assert _rel(ctx.get_propagator(sp, which, s), ref.get_propagator(sp, which, s)) <= TOL, (sp, which, s)
    ctx.close()


def test_seed_box_start_matches_oracle(abi, configs, orc, lib):
    """dbk1 forward A block from the seed box on iterations <= SEED_STEPS
    (SCFT.c:2346-2362), then from ones."""
    p = configs.t_film_small()
    p.enable_seed = 1
    p.seed_steps = 2
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    ref = orc.state(p)
    w, pf = configs.droplet_fields(p, radius=3.0)
    for s in (ctx, ref):
        s.set_fields(w, pf)
        s.set_mask(configs.brush_mask(p))
    for it in (1, 2, 3):
        a, b = ctx.iterate(it), ref.iterate(it)
        assert _scalar_rel(a.free_energy, b.free_energy) <= TOL
        assert _rel(ctx.get_propagator(0, 0, 0), ref.get_propagator(0, 0, 0)) == 0.0
    ctx.close()


def test_noise_matches_oracle(abi, configs, orc, lib):
    """MF = 1 (iIter <= RELAXATION): the host replays the reference's
    drand48/polar stream and uploads it in draw order (SCFT.c:2103)."""
    p = configs.t_film_small()
    p.relaxation = 10
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    ref = orc.state(p)
    gen = orc.state(p)
    seed = (11, 22, 33)
    ref.seed48(seed)
    gen.seed48(seed)
    w, pf = configs.droplet_fields(p, radius=3.0)
    for s in (ctx, ref):
        s.set_fields(w, pf)
        s.set_mask(configs.brush_mask(p))
    n = ctx.derived.n_active * ctx.V
    for it in (2, 3):
        noise = np.array([gen.gauss() for _ in range(n)])
        a = ctx.iterate(it, 1.0, noise)
        b = ref.iterate(it)
        assert b.mf == 1.0
        assert _scalar_rel(a.free_energy, b.free_energy) <= TOL
        wa, _ = ctx.get_fields()
        wb, _ = ref.get_fields()
        assert _rel(wa, wb) <= TOL
    assert gen.get_seed() == ref.get_seed()
    ctx.close()


def test_uniform_field_closed_form(abi, configs, lib):
    """Periodic, uniform w: K(0) = 1 keeps q uniform, q(s) = exp(-w s DS)."""
    p = abi.make_params(lx=8.0, ly=1.0, lz=6.0, dr_target=0.5, ns=21, mask=0,
                        vol_homo=(0, 0, 1.0, 0, 0, 0, 0, 0, 0, 0))
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    w = np.zeros((13, ctx.V))
    w[2] = 0.7
    ctx.set_fields(w, np.zeros(ctx.V))
    o = ctx.iterate(2)
    ds = 1.0 / 20
    for s in range(21):
        q = ctx.get_propagator(6, 0, s)
        assert np.abs(q - np.exp(-0.7 * s * ds)).max() <= 1e-13
    assert abs(o.chempot[6] - 0.7 * 1.0) <= 1e-12          # -ln Q = w * alpha
    phi, tot = ctx.get_densities()
    assert np.abs(phi[2] - 1.0).max() <= 1e-12              # Simpson of a constant
    ctx.close()


def test_dirichlet_sine_mode_decay(abi, configs, lib):
    """MASK 1, w = 0, brush q started from a sine x Fourier mode: each step
    multiplies it by Kz(m) Kx(mx) with the reference's extended z grid
    N = 2(NZ+1) (SCFT.c:1000-1004): q_s = (Kz Kx)^s q_0."""
    p = configs.chem_droplet(lx=8.0, ly=1.0, lz=5.0, dr=0.5, ns=21, vol_dbk=0.1, vol_s1=0.05, len_s1=0.5)
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    nx, ny, nz = ctx.shape
    ctx.set_fields(np.zeros((13, ctx.V)), np.zeros(ctx.V))
    m, mx = 2, 3
    i = np.arange(nx)[:, None]
    k = np.arange(nz)[None, :]
    q0 = np.sin(np.pi * m * (k + 1) / (nz + 1)) * np.cos(2 * np.pi * mx * i / nx)
    mask = np.zeros((3, ctx.V))
    mask[0] = q0.reshape(-1)
    ctx.set_mask(mask)
    ctx.propagate(2)
    ctx.sync()
    ds = 1.0 / 20
    N = 2 * (nz + 1)
    dz = p.lz / nz
    dx = p.lx / nx
    kz = 2 * np.pi * m / N
    kx = 2 * np.pi * mx / nx
    fac = np.exp(-ds / dz ** 2 * kz ** 2) * np.exp(-ds / dx ** 2 * kx ** 2)
    ns = ctx.derived.ns_sp[14]
    for s in range(ns):
        q = ctx.get_propagator(14, 0, s)
        assert np.abs(q - fac ** s * q0.reshape(-1)).max() <= 1e-13, s
    ctx.close()


@pytest.mark.parametrize("name", ["film12x10", "film3d", "c1_64x64"])
def test_fp32_split_close_to_fp64_oracle(abi, configs, orc, lib, name):
    ctx, ref, outs, p = _run_pair(abi, configs, orc, lib, name, precision=abi.SCFT_FP32_SPLIT)
    for a, b in outs:
        dF = abs(a.free_energy - b.free_energy)
        print(name, "|dF| fp32-split vs fp64 oracle = %.3e" % dF)
        assert dF <= 1e-6
    ctx.close()


def _golden():
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "c2_oracle.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("zkind", ["auto", "fft", "dense", "film"])
def test_c2_full_size_matches_golden(abi, configs, lib, zkind):
    """C2 (256x1x128, MASK 1, NS 101) against the oracle's FP64 fixture
    (tests/golden/make_golden.py): three iterations, 1e-10.  "dense" runs
    the dense z operator and the x transform as two launches per step,
    "film" as one fused launch (k_film)."""
    import os
    g = _golden()["c2_droplet_seed0"]
    if zkind != "auto":
        os.environ["SCFT_ZKIND"] = "dense" if zkind == "film" else zkind
    if zkind in ("dense", "film"):
        os.environ["SCFT_FILM"] = "1" if zkind == "film" else "0"
    try:
        p = configs.params_c2()
        ctx = abi.Context(lib, p, abi.SCFT_FP64)
        print(ctx.schedule())
        w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
        ctx.set_fields(w, pf)
        ctx.set_mask(configs.brush_mask(p))
        for row in g["iters"]:
            o = ctx.iterate(row["iter"])
            assert _scalar_rel(o.free_energy, row["F"]) <= TOL, (row["iter"], o.free_energy, row["F"])
            assert _scalar_rel(o.rmsdev, row["rmsdev"]) <= TOL
            assert _scalar_rel(o.dw_rms, row["dw_rms"]) <= TOL
            for sp in range(17):
                assert _scalar_rel(o.chempot[sp], row["chempot"][sp]) <= TOL
        if zkind in ("dense", "film"):   # the schedule exists once an iteration ran
            nfilm = int(ctx.schedule().split("film=")[1].split()[0])
            assert (nfilm > 0) == (zkind == "film"), ctx.schedule()
        wf, pfin = ctx.get_fields()
        phi, _ = ctx.get_densities()
        idx = g["sample_idx"]
        for f, vals in g["w_samples"].items():
            assert _rel(wf[int(f)][idx], vals) <= TOL
            assert abs(np.sqrt((wf[int(f)] ** 2).sum()) - g["w_l2"][f]) <= TOL * g["w_l2"][f]
        for f, vals in g["phi_samples"].items():
            assert _rel(phi[int(f)][idx], vals) <= TOL
        assert _rel(pfin[idx], g["p_samples"]) <= TOL
        ctx.close()
    finally:
        os.environ.pop("SCFT_ZKIND", None)
        os.environ.pop("SCFT_FILM", None)


def test_c2_fp32_split_batch_matches_golden(abi, configs, lib):
    """The served path (batch of 3 C2 slots, FP32-split propagators, one
    hipGraph per iteration) against the oracle's FP64 fixture: F, RMSDev and
    ||dw|| within 1e-7 relative, the fields within 1e-6 of their norm."""
    g = _golden()["c2_droplet_seed0"]
    p = configs.params_c2()
    b = abi.Batch(lib, p, 3, precision=abi.SCFT_FP32_SPLIT)
    print(b.schedule())
    w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
    for s in range(3):
        b.init_slot(s, p)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    for row in g["iters"]:
        outs = b.iterate([row["iter"]] * 3)
        for s in range(3):
            o = outs[s]
            assert _scalar_rel(o.free_energy, row["F"]) <= 1e-7, (s, row["iter"], o.free_energy, row["F"])
            assert _scalar_rel(o.rmsdev, row["rmsdev"]) <= 1e-6
            assert _scalar_rel(o.dw_rms, row["dw_rms"]) <= 1e-5
    idx = g["sample_idx"]
    for s in range(3):
        wf, pfin = b.get_fields(s)
        for f, vals in g["w_samples"].items():
            assert np.abs(wf[int(f)][idx] - np.asarray(vals)).max() <= 1e-6 * g["w_l2"][f] / np.sqrt(len(wf[0]))
    b.close()


def test_c1_full_run_matches_golden(abi, configs, lib):
    """C1 (64x1x64 AB melt, chi N 20, G1 0.1) for all of its 2000 iterations
    against the oracle's FP64 trajectory (tests/golden/make_golden.py c1):
    F, RMSDev and -ln Q every 10 iterations and the final fields, 1e-10."""
    import json
    import os
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "golden", "c1_oracle.json")) as f:
        g = json.load(f)["c1_random_seed0"]
    p = configs.params_c1()
    ctx = abi.Context(lib, p, abi.SCFT_FP64)
    w, pf = configs.random_fields(p, seed=0, amp=0.5)
    ctx.set_fields(w, pf)
    rows = {r["iter"]: r for r in g["iters"]}
    worst = 0.0
    for it in range(2, g["n"] + 2):
        o = ctx.iterate(it)
        r = rows.get(it)
        if r is None:
            continue
        worst = max(worst, _scalar_rel(o.free_energy, r["F"]))
        assert _scalar_rel(o.free_energy, r["F"]) <= TOL, (it, o.free_energy, r["F"])
        assert _scalar_rel(o.rmsdev, r["rmsdev"]) <= TOL, it
        assert _scalar_rel(o.chempot[0], r["chempot"][0]) <= TOL, it
    print("C1 2000 iterations: worst relative F deviation %.2e" % worst)
    wf, pfin = ctx.get_fields()
    idx = g["sample_idx"]
    for f, vals in g["w_samples"].items():
        assert _rel(wf[int(f)][idx], vals) <= TOL
    ctx.close()
