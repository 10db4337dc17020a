"""Timing runs of the reference's workloads on one GPU, outside the C2
bench: per batched iteration and per launch of the z operator and the
last-axis transform (HIP events, scft_time_launches).

    python workloads.py c1 [--slots B] [--iters N] [--precision fp32|fp64]
        2D bulk AB diblock melt, 64x1x64, NS 101, MASK 0; B resident (1)
    python workloads.py c2 [--slots B] [--iters N] [--precision fp32|fp64]
        2D thin-film droplet (the bench workload), 256x1x128, NS 101,
        MASK 1; B resident (192)
    python workloads.py c3 [--slots B] [--iters N] [--precision fp32|fp64]
        3D thin-film droplet, 128x128x64, NS 201 (C3 chemistry); B resident (1)
    python workloads.py c5 [--slots B] [--iters N] [--precision ...]
        3D served requests, 96x96x48 (C3 chemistry), B resident (8)
    python workloads.py c4 [--slots B] [--iters N] [--precision ...]
        phase-diagram sweep batch: B slots of the C2 chemistry with
        chi_AB N over 15..36 and chi_AC N over 20..55 (8 x 8 grid), LX drawn
        per slot from {24, 32, 40} R_g (seeded), so grids are 192/256/320 x
        1 x 128 of uneven size in one batch; B resident (192)

    python workloads.py c5stream [--rates R,R,..] [--duration S] [--scale X]
        the config-5 serving stream: Poisson arrivals at each rate (requests
        per second), 80% 2D trial solves (C2 chemistry, chi and LX from the
        C4 grid) and 20% 3D 96x96x48 solves (C3 chemistry), each seeded with
        its own droplet; a 2D pool of up to 192 slots and a 3D pool of up to
        32, sized against free HBM (serve.Server).  Each request runs a
        budget of iterations drawn from the measured C2
        iterations-to-converge distribution (profiles/c2_convergence_fp32.json)
        times `scale`, so a run finishes in minutes; with every service time
        scaled by the same factor at the same load, time-to-converge scales
        by 1/scale and throughput in solves by scale, which the line also
        prints.

Prints one line per workload.  Fields start from the droplet seed; the
timed iterations follow two warm-up iterations.
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import abi  # noqa: E402
import configs  # noqa: E402


def c1_params(slots):
    return [configs.params_c1() for _ in range(slots)]


def c2_params(slots):
    return [configs.params_c2() for _ in range(slots)]


def c3_params(slots):
    return [configs.params_c3() for _ in range(slots)]


def c5_params(slots):
    return [configs.params_c5_3d() for _ in range(slots)]


def c4_params(slots, seed=0):
    rng = np.random.default_rng(seed)
    lx = rng.choice([24.0, 32.0, 40.0], size=slots)
    chi_ab = np.linspace(15.0, 36.0, 8)
    chi_ac = np.linspace(20.0, 55.0, 8)
    out = []
    for s in range(slots):
        i, j = (s // 8) % 8, s % 8
        out.append(configs.chem_droplet(lx=float(lx[s]), chi_ab=float(chi_ab[i]), chi_ac=float(chi_ac[j])))
    return out


def c5_stream(lib, rate, duration, scale, precision, seed=0, drain=None):
    """One Poisson stream of config-5 requests through serve.Server."""
    import json
    import serve
    with open(os.path.join(os.path.dirname(HERE), "profiles", "c2_convergence_fp32.json")) as f:
        dist = json.load(f)["iters"]
    rng = np.random.default_rng(seed)
    chi_ab = np.linspace(15.0, 36.0, 8)
    chi_ac = np.linspace(20.0, 55.0, 8)
    reqs, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        a, c = float(rng.choice(chi_ab)), float(rng.choice(chi_ac))
        if rng.random() < 0.8:
            p = configs.chem_droplet(lx=float(rng.choice([24.0, 32.0, 40.0])), chi_ab=a, chi_ac=c)
        else:
            p = configs.params_c5_3d(chi_ab=a, chi_ac=c)
        n = max(1, int(round(float(rng.choice(dist)) * scale)))
        reqs.append(serve.Request(k, p, (serve.SEED_DROPLET, {"radius": 6.0, "seed": k}), arrival=t, max_iters=n))
        k += 1
    key3 = serve.structure_key(configs.params_c5_3d())
    server = serve.Server(lib, precision=precision, default_slots=192, slots={key3: 32}, max_nx=320, tol=0.0)
    # create both pools before the clock starts (the 2D one at the largest NX)
    server.pool_for(serve.Request(-1, configs.chem_droplet(lx=40.0), None))
    server.pool_for(serve.Request(-1, configs.params_c5_3d(), None))
    t0 = time.perf_counter()
    server.run_stream(reqs, deadline=duration + (drain if drain is not None else 3.0 * duration))
    wall = time.perf_counter() - t0
    done = [r for r in server.done() if r.rid >= 0]
    lat = [r.time_to_converge for r in done]
    lat2 = [r.time_to_converge for r in done if r.params.ly == 1.0]
    lat3 = [r.time_to_converge for r in done if r.params.ly != 1.0]
    pc, pc2, pc3 = serve.percentiles(lat), serve.percentiles(lat2), serve.percentiles(lat3)
    iters = sum(r.iters for r in done)
    pools = {("3D" if kk == key3 else "2D"): pl for kk, pl in server.pools.items()}
    occ = {n: (pl.slot_iterations / max(pl.iterations, 1)) for n, pl in pools.items()}
    res = {"rate": rate, "duration": duration, "requests": len(reqs), "done": len(done), "done_3d": len(lat3),
           "wall_s": wall, "req_per_s": len(done) / wall, "solve_iterations_per_s": iters / wall,
           "p50_s": pc[50], "p99_s": pc[99], "p50_2d_s": pc2[50], "p99_2d_s": pc2[99], "p50_3d_s": pc3[50],
           "p99_3d_s": pc3[99], "mean_active_2d": occ.get("2D", 0), "mean_active_3d": occ.get("3D", 0),
           "slots_2d": pools["2D"].nslots if "2D" in pools else 0, "slots_3d": pools["3D"].nslots if "3D" in pools else 0,
           "mean_budget_iterations": iters / max(len(done), 1), "scale": scale,
           "full_length_solves_per_hour": len(done) / wall * scale * 3600.0,
           "full_length_p50_s": (pc[50] or 0) / scale, "full_length_p99_s": (pc[99] or 0) / scale}
    print("C5 stream rate %.2f/s x %.0f s (%d requests, %d done, %d 3D): throughput %.3f req/s, %.0f solve-iterations/s; "
          "time-to-converge p50 %.2f s p99 %.2f s (2D p50 %.2f p99 %.2f, 3D p50 %.2f p99 %.2f); mean active 2D %.1f 3D %.1f "
          "of %d / %d slots; mean budget %.0f iterations (scale %g): at full length %.2f solves/hour, p50 %.0f s, p99 %.0f s"
          % (rate, duration, len(reqs), len(done), len(lat3), len(done) / wall, iters / wall, pc[50] or 0, pc[99] or 0,
             pc2[50] or 0, pc2[99] or 0, pc3[50] or 0, pc3[99] or 0, occ.get("2D", 0), occ.get("3D", 0),
             pools["2D"].nslots if "2D" in pools else 0, pools["3D"].nslots if "3D" in pools else 0,
             iters / max(len(done), 1), scale, len(done) / wall * scale * 3600.0, (pc[50] or 0) / scale,
             (pc[99] or 0) / scale), flush=True)
    server.close()
    return res


def run(lib, name, protos, precision, iters):
    maxnx = max(abi.grid_of(p)[0] for p in protos)
    b = abi.Batch(lib, protos[0], len(protos), max_nx=maxnx, precision=precision)
    for s, p in enumerate(protos):
        b.init_slot(s, p)
        w, pf = configs.droplet_fields(p, radius=6.0, seed=s)
        b.set_fields(s, w, pf)
        b.set_mask(s, configs.brush_mask(p))
        b.activate(s)
    b.iterate()
    b.iterate()
    t0 = time.perf_counter()
    for _ in range(iters):
        outs = b.iterate()
    dt = (time.perf_counter() - t0) / iters
    z = b.time_launches(0, 5)
    x = b.time_launches(1, 5)
    print("%s slots=%d %s: %.3f ms per batched iteration, %.1f us per instance-iteration; "
          "z launch %.1f us x%d (%.0f GB/s), last-axis launch %.1f us x%d (%.0f GB/s); F[0]=%.10f; %s"
          % (name, len(protos), "fp32-split" if precision == abi.SCFT_FP32_SPLIT else "fp64", dt * 1e3,
             dt / len(protos) * 1e6, z["avg_us"], z["launches"], z["bytes"] / max(z["avg_us"], 1e-9) * 1e-3,
             x["avg_us"], x["launches"], x["bytes"] / max(x["avg_us"], 1e-9) * 1e-3, outs[0].free_energy,
             b.schedule()), flush=True)
    b.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", choices=["c1", "c2", "c3", "c4", "c5", "c5stream"])
    ap.add_argument("--rates", default="2,4,6", help="c5stream: arrival rates (requests/s)")
    ap.add_argument("--duration", type=float, default=40.0, help="c5stream: seconds of arrivals per rate")
    ap.add_argument("--scale", type=float, default=1e-3, help="c5stream: iteration budget scale")
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--slots", type=int, default=None)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    a = ap.parse_args(argv)
    lib = abi.load_library()
    prec = abi.SCFT_FP32_SPLIT if a.precision == "fp32" else abi.SCFT_FP64
    if a.workload == "c5stream":
        for r in a.rates.split(","):
            c5_stream(lib, float(r), a.duration, a.scale, prec)
        return
    slots = a.slots or {"c1": 1, "c2": 192, "c3": 1, "c4": 192, "c5": 8}[a.workload]
    if a.workload == "c1":
        run(lib, "C1 64x1x64", c1_params(slots), prec, a.iters)
    elif a.workload == "c2":
        run(lib, "C2 256x1x128", c2_params(slots), prec, a.iters)
    elif a.workload == "c3":
        run(lib, "C3 128x128x64", c3_params(slots), prec, a.iters)
    elif a.workload == "c5":
        run(lib, "C5 96x96x48", c5_params(slots), prec, a.iters)
    else:
        run(lib, "C4 sweep (NX 192/256/320)", c4_params(slots), prec, a.iters)


if __name__ == "__main__":
    main()
