#!/usr/bin/env python3
"""Benchmark: converged SCFT solves/hour (with p50/p99 time-to-converge) of
the 2D thin-film nanodroplet workload on one MI355X.

Workload (256x1x128, NS 101, MASK 1): symmetric AB diblock (f 0.5, vol 0.15)
in homopolymer C (vol 0.85) over an S1 brush (vol 0.05) grafted at z = 0,
chi_AB N = 25, chi_AC N = chi_BC N = 40, G1 = 0.01; each solve starts from
a hemispherical droplet of radius 6 R_g with seeded +-0.1 noise and runs to
||dw||_rms < 1e-5.  The served path keeps `--slots` solves resident and
advances them together (FP32-split propagators, FP64 reductions/update);
converged slots are refilled from the queue.

A step is one batched iteration of every resident solve.  A C2 solve needs
~1e5 iterations (the reference runs ITER = 300000), far more than fit in the
timed window, so solves/hour = measured instance-iterations per hour divided
by the mean iterations-to-converge, and p50/p99 time-to-converge =
iterations-to-converge percentiles x measured time per batched iteration.
The iterations-to-converge distribution is measured by running solves of
this workload to convergence (profiles/c2_convergence_<precision>.json,
produced by `python bench.py --converge N --precision P`).

`roofline` is for the dominant kernel (the z operator): 2 b bytes per
voxel-step over its launch time.  `iteration_roofline` is for the whole
iteration: 2 b V S + 2 b V P + 8 V (3 n + 2) bytes per instance-iteration
(S propagator steps, P Simpson points, n active fields) times the measured
instance-iterations per second.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "stimuli-responsive-nanodroplets_amd")
for _p in (ROOT, os.path.join(ROOT, "oracle"), PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "converged SCFT solves/hour + p50/p99 time-to-converge at |\u0394F|\u22641e-6 kT vs FP64"
PEAK_HBM_GBS = 8000.0
CONV_PROFILE = os.path.join(ROOT, "profiles", "c2_convergence_%s.json")   # per precision (fp32, fp32_nearest, fp64)
# HBM bytes per z-operator dispatch from the FETCH_SIZE / WRITE_SIZE runs
# (calibrated with a copy probe of known size); see profiles/c2_served_b192_hbm_traffic.md
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "c2_served_b192_hbm_traffic.json")


def measured_traffic(schedule, slots, precision):
    """Counter-measured bytes per launch of the dominant kernel, when the
    committed measurement is of the same kernel and workload (192 slots,
    FP32-split); None otherwise."""
    if not os.path.exists(TRAFFIC_PROFILE) or slots != 192 or precision != "fp32":
        return None
    with open(TRAFFIC_PROFILE) as f:
        t = json.load(f)
    if t.get("kernel") != dominant_kernel(schedule).split()[0] or "dense_mfma2=100" not in schedule:
        return None
    return float(t["bytes_per_launch"])


def dominant_kernel(schedule):
    """Name of the kernel that scft_time_launches kind 0 times in this schedule."""
    n = dict((k, int(v)) for k, v in (kv.split("=") for kv in schedule.split() if "=" in kv and kv[0] != "z"))
    if n.get("film", 0):
        return "k_film (dense z operator and x transform fused, one launch per contour step and x length)"
    if n.get("mfma2_pf", 0):
        return "k_zdense_mfma2p (z operator of every chain, one launch per contour step, next tile's slice rows in flight)"
    if n.get("dense_mfma2", 0):
        return "k_zdense_mfma2 (z operator of every chain, one launch per contour step)"
    return "first axis of every chain, one launch per contour step"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_pool(abi, configs, serve, lib, slots, precision):
    p = configs.params_c2()
    pool = serve.Pool(lib, p, precision=precision, max_slots=slots)
    for i in range(pool.nslots):
        pool.submit(serve.Request(i, configs.params_c2(), (serve.SEED_DROPLET, {"radius": 6.0, "seed": i})))
    return pool


def time_iterations(pool, warmup, steps):
    for _ in range(warmup):
        pool.step()
    t0 = time.perf_counter()
    it0 = pool.slot_iterations
    for _ in range(steps):
        pool.step()
    dt = time.perf_counter() - t0
    return dt, pool.slot_iterations - it0


def conv_stats(prec_name):
    path = CONV_PROFILE % prec_name.replace("-", "_")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        runs = json.load(f)
    if not runs.get("iters"):
        return None
    import numpy as np
    # solves still running when the profile run stopped enter at the
    # iteration they had reached: a lower bound (right-censored)
    cens = [u["iters"] for u in runs.get("unconverged", [])]
    it = np.asarray(list(runs["iters"]) + cens, float)
    return {"mean": float(it.mean()), "p50": float(np.percentile(it, 50)), "p99": float(np.percentile(it, 99)),
            "n": int(len(it)), "n_censored": len(cens), "censored_at": (min(cens) if cens else None),
            "source": "profiles/c2_convergence_%s.json" % prec_name.replace("-", "_")}


def iteration_bytes(abi, lib, params, b_el):
    """Algorithmic bytes of one instance-iteration, 2 b V S + 2 b V P +
    8 V (3 n + 2): S propagator steps (diblocks and brushes two chains,
    homopolymers one), P Simpson points (a diblock's two blocks share their
    junction point, counted in both), n active fields."""
    import ctypes as C
    d = abi.ScftDerived()
    abi.check(lib, lib.scft_derive(C.byref(params), C.byref(d)), "scft_derive")
    V = d.nx * d.ny * d.nz
    S = P = 0
    for sp in range(abi.NSP):
        ns = d.ns_sp[sp]
        if ns <= 0:
            continue
        chains = 1 if 4 <= sp < 14 else 2
        S += chains * (ns - 1)
        P += ns + 1 if sp < 4 else ns
    n = d.n_active
    return 2.0 * b_el * V * S + 2.0 * b_el * V * P + 8.0 * V * (3 * n + 2), {"V": V, "S": S, "P": P, "n": n}


def _cpu_worker(args):
    """One oracle process: C2 iterations from the droplet seed until `seconds`
    have passed (at least one); returns seconds per iteration."""
    seconds, max_iters = args
    import abi
    import configs
    import oracle
    p = configs.params_c2()
    st = oracle.Oracle(abi).state(p)
    w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
    st.set_fields(w, pf)
    st.set_mask(configs.brush_mask(p))
    n = 0
    t0 = time.perf_counter()
    while n < max_iters:
        st.iterate(2 + n)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    return (time.perf_counter() - t0) / n, n


def cpu_baseline(abi, configs, seconds, procs, conv64):
    """The FP64 oracle (the plain C restatement of SCFT.c's iteration,
    single-threaded like the reference) as one process per host core, each
    on its own C2 solve: throughput = procs x 3600 / (s per iteration x
    mean FP64 iterations-to-converge)."""
    import multiprocessing as mp
    procs = procs or min(16, os.cpu_count() or 1)
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_cpu_worker, [(seconds, 5)] * procs, chunksize=1)
    t_iter = sum(r[0] for r in res) / len(res)
    n_iter = sum(r[1] for r in res)
    return {"value": (procs * 3600.0 / (conv64["mean"] * t_iter)) if conv64 else None,
            "unit": "solves/hour", "cores": procs, "kind": "port",
            "sample": "%d processes x C2 iterations of the FP64 oracle (oracle/scft_oracle.c, single-threaded, "
                      "the project's own mixed-radix FFT with O(p^2) butterflies for the prime 43 of the 258-point "
                      "Dirichlet extension, so slower than FFTW would be: FFTW3 is not installed), %d iterations in all, "
                      "%.2f s per iteration per process; solves/hour = processes x 3600 / (s per iteration x "
                      "mean FP64 iterations-to-converge)" % (procs, n_iter, t_iter),
            "nproc": os.cpu_count(), "s_per_iteration": t_iter,
            "iters_to_converge_source": conv64["source"] if conv64 else None}


ACCURACY_PROFILE = os.path.join(ROOT, "profiles", "c2_convergence_pairs.json")


def accuracy_summary(prec_name):
    """The FP32-split vs FP64 comparison at convergence of the same C2
    seeds (profiles/c2_convergence_pairs.json), when it exists."""
    if prec_name == "fp64":
        return {"note": "FP64 throughout: matches the oracle to 1e-10 per iteration (tests/test_gpu_*.py)"}
    if not os.path.exists(ACCURACY_PROFILE):
        return None
    with open(ACCURACY_PROFILE) as f:
        a = json.load(f)
    return {k: a[k] for k in ("summary", "source") if k in a}


def converge_seeds(spec):
    """'0-15' or '0,3,7' -> list of ints."""
    out = []
    for part in spec.split(","):
        if "-" in part:
            a, b = part.split("-")
            out.extend(range(int(a), int(b) + 1))
        elif part:
            out.append(int(part))
    return out


PERTURBED_RID = 1000   # request id offset of the perturbed copies of --converge-perturb


def converge_run(abi, configs, serve, lib, seeds, prec_name, tlimit, out, fields_dir=None, state_in=None,
                 state_out=None, perturb=0.0, perturb_seeds=()):
    """Run C2 solves (droplet seeds `seeds`, all resident) to convergence;
    write their iterations-to-converge, converged F and ||dw|| history to
    `out` (JSON).  With `perturb`, each seed in `perturb_seeds` also runs a
    copy (request id PERTURBED_RID + seed) whose w_A and w_B start from the
    droplet seed plus `perturb` x U(-1, 1) (PCG64(PERTURBED_RID + seed)): it
    shows how far a perturbation of that size moves the end state.  With
    `state_in`, the solves continue from saved states (fields, iteration
    count, history, and the records of solves already converged) instead;
    with `state_out`, the solves still running at the time limit are saved
    there (the iteration is deterministic, so a resumed solve continues
    exactly where it stopped)."""
    import numpy as np
    precision = {"fp32": abi.SCFT_FP32_SPLIT, "fp32-nearest": abi.SCFT_FP32_NEAREST}.get(prec_name, abi.SCFT_FP64)
    p = configs.params_c2()
    prior = []
    reqs = []
    if state_in:
        for name in sorted(os.listdir(state_in)):
            if not name.endswith(".npz"):
                continue
            z = np.load(os.path.join(state_in, name))
            r = serve.Request(int(z["seed"]), configs.params_c2(), (serve.SEED_FIELDS, (z["w"], z["p"])),
                              keep_fields=fields_dir is not None)
            r.iters = int(z["iters"])
            r.start_iter = 1
            r.history = [tuple(h) for h in z["history"].tolist()]
            reqs.append(r)
        done_path = os.path.join(state_in, "done.json")
        if os.path.exists(done_path):
            with open(done_path) as f:
                prior = json.load(f)
    else:
        for i in seeds:
            reqs.append(serve.Request(i, configs.params_c2(), (serve.SEED_DROPLET, {"radius": 6.0, "seed": i}),
                                      keep_fields=fields_dir is not None))
        for i in perturb_seeds:
            w, pf = configs.droplet_fields(p, radius=6.0, seed=i)
            rng = np.random.Generator(np.random.PCG64(PERTURBED_RID + i))
            w = w.copy()
            w[0] += perturb * rng.uniform(-1.0, 1.0, w.shape[1])
            w[1] += perturb * rng.uniform(-1.0, 1.0, w.shape[1])
            reqs.append(serve.Request(PERTURBED_RID + i, configs.params_c2(), (serve.SEED_FIELDS, (w, pf)),
                                      keep_fields=fields_dir is not None))
    pool = serve.Pool(lib, p, precision=precision, max_slots=len(reqs), record_every=1000,
                      fp64_check=precision != abi.SCFT_FP64)
    for r in reqs:
        pool.submit(r)
    t0 = time.perf_counter()
    last = t0
    while pool.busy() and time.perf_counter() - t0 < tlimit:
        pool.step()
        if time.perf_counter() - last > 30:
            last = time.perf_counter()
            dws = ["%.1e" % r.history[-1][2] if r.history else "-" for r in pool.running.values()]
            log("converge %s: %d iterations, %d converged, %.0f s, dw %s" % (
                prec_name, pool.iterations, len(pool.done), last - t0, " ".join(dws)))
    wall = time.perf_counter() - t0
    res = {"precision": prec_name, "workload": "C2 droplet seeds %s" % ",".join(str(s) for s in seeds),
           "perturb": perturb, "perturb_seeds": list(perturb_seeds), "tol": pool.tol,
           "wall_s": wall, "batched_iterations": pool.iterations, "s_per_batched_iteration": wall / max(1, pool.iterations),
           "iters": [], "F": [], "seed": [], "F64": [], "dF": [], "dw64": [], "unconverged": [], "history": {}}
    records = list(prior)
    for r in pool.done:
        if r.status == "converged":
            records.append({"seed": r.rid, "iters": r.iters, "F": r.F, "history": r.history,
                            "F64": r.F64, "dF": r.dF, "dw64": r.dw64})
            if fields_dir:
                os.makedirs(fields_dir, exist_ok=True)
                w, pf = r.fields
                np.savez_compressed(os.path.join(fields_dir, "%s_seed%d.npz" % (prec_name, r.rid)), w=w, p=pf)
    for rec in records:
        res["iters"].append(rec["iters"])
        res["F"].append(rec["F"])
        res["seed"].append(rec["seed"])
        for k in ("F64", "dF", "dw64"):
            res[k].append(rec.get(k))
        res["history"][str(rec["seed"])] = rec["history"]
    for s, r in pool.running.items():
        res["unconverged"].append({"seed": r.rid, "iters": r.iters, "last": r.history[-1] if r.history else None})
        res["history"][str(r.rid)] = r.history
        if state_out:
            os.makedirs(state_out, exist_ok=True)
            w, pf = pool.batch.get_fields(s)
            np.savez_compressed(os.path.join(state_out, "%s_seed%d.npz" % (prec_name, r.rid)), w=w, p=pf, seed=r.rid,
                                iters=r.iters, history=np.asarray(r.history, float).reshape(-1, 3))
    if state_out:
        os.makedirs(state_out, exist_ok=True)
        with open(os.path.join(state_out, "done.json"), "w") as f:
            json.dump(records, f)
    pool.close()
    with open(out, "w") as f:
        json.dump(res, f)
    log("converged %d of %d in %.0f s (%d still running)" % (len(res["iters"]), len(records) + len(res["unconverged"]),
                                                             wall, len(res["unconverged"])))


def convergence_pairs():
    """profiles/c2_convergence_pairs.json from the FP32-split and FP64
    convergence profiles of the same C2 seeds: per seed the end state of
    each run (iterations, F), the FP64 evaluation of the FP32-split end
    state (F64, |dF|, FP64 ||dw|| there), and, where the FP64 run holds a
    perturbed copy of the seed, that copy's end state."""
    import numpy as np
    with open(CONV_PROFILE % "fp32") as f:
        a = json.load(f)
    with open(CONV_PROFILE % "fp64") as f:
        b = json.load(f)
    r32 = {s: (i, F, F64, dF, dw) for s, i, F, F64, dF, dw in
           zip(a["seed"], a["iters"], a["F"], a.get("F64", [None] * len(a["seed"])),
               a.get("dF", [None] * len(a["seed"])), a.get("dw64", [None] * len(a["seed"])))}
    r64 = {s: (i, F) for s, i, F in zip(b["seed"], b["iters"], b["F"])}
    rows = []
    for s in sorted(k for k in r32 if k < PERTURBED_RID):
        i32, F32, F64e, dF, dw64 = r32[s]
        row = {"seed": s, "fp32_iters": i32, "fp32_F": F32, "fp64_eval_F": F64e, "fp64_eval_dF": dF,
               "fp64_eval_dw": dw64}
        if s in r64:
            row.update({"fp64_iters": r64[s][0], "fp64_F": r64[s][1], "dF_end_states": abs(F32 - r64[s][1])})
        if PERTURBED_RID + s in r64:
            row.update({"fp64_perturbed_iters": r64[PERTURBED_RID + s][0], "fp64_perturbed_F": r64[PERTURBED_RID + s][1],
                        "dF_fp64_vs_perturbed": abs(r64[s][1] - r64[PERTURBED_RID + s][1]) if s in r64 else None})
        rows.append(row)
    dfe = [r["fp64_eval_dF"] for r in rows if r["fp64_eval_dF"] is not None]
    same32 = [r for r in rows if "dF_end_states" in r]
    pert = [r for r in rows if r.get("dF_fp64_vs_perturbed") is not None]
    summary = {
        "seeds": len(rows),
        "fp32_results_within_1e-6_of_fp64_eval": sum(1 for x in dfe if x <= 1e-6),
        "fp32_results_checked": len(dfe),
        "max_fp64_eval_dF": max(dfe) if dfe else None,
        "max_fp64_eval_dw": max((r["fp64_eval_dw"] for r in rows if r["fp64_eval_dw"] is not None), default=None),
        "fp32_vs_fp64_same_end_state": sum(1 for r in same32 if r["dF_end_states"] <= 1e-6),
        "fp32_vs_fp64_pairs": len(same32),
        "fp64_vs_perturbed_same_end_state": sum(1 for r in pert if r["dF_fp64_vs_perturbed"] <= 1e-6),
        "fp64_vs_perturbed_pairs": len(pert),
        "perturbation": b.get("perturb"),
        "mean_iters_fp32": float(np.mean([r["fp32_iters"] for r in rows])),
        "mean_iters_fp64": float(np.mean([r["fp64_iters"] for r in same32])) if same32 else None,
    }
    out = {"summary": summary, "rows": rows,
           "source": "profiles/c2_convergence_fp32.json, profiles/c2_convergence_fp64.json"}
    path = os.path.join(ROOT, "profiles", "c2_convergence_pairs.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(summary, indent=1))


def c5_line(abi, configs, args):
    """The config-5 stream measured: Poisson arrivals at --rate for
    --duration seconds, 80% 2D 192/256/320 x 1 x 128 (C2 chemistry, chi and
    LX from the C4 grid) and 20% 3D 96 x 96 x 48 (C3 chemistry) requests,
    FP32-split, through serve.Server.  Each request runs a budget of
    iterations drawn from the measured C2 iterations-to-converge
    distribution times --scale (the convergence test off), so the window
    takes minutes; p50/p99 are measured at that scale, and the full-length
    figures divide times by the scale and multiply throughput by it."""
    import workloads
    lib = abi.load_library()
    r = workloads.c5_stream(lib, args.rate, args.duration, args.scale, abi.SCFT_FP32_SPLIT)
    line = {"metric": METRIC, "value": r["full_length_solves_per_hour"], "unit": "solves/hour", "n_gpus": 1,
            "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64 arithmetic, f32 propagator slices", "data": "synthetic",
            "config": {"workload": "config-5 served stream: Poisson arrivals, 80% 2D 256x1x128-class trial solves "
                                   "(C2 chemistry, LX 24/32/40) and 20% 3D 96x96x48 (C3 chemistry), FP32-split",
                       "measured": r,
                       "basis": "value = measured completed requests/s x 3600 x scale (every service time scaled "
                                "by the same factor at the same load); measured p50/p99 are at the scaled budgets"}}
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed batched iterations")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--slots", type=int, default=192, help="resident solves (capped by HBM footprint)")
    ap.add_argument("--precision", choices=["fp32", "fp32-nearest", "fp64"], default="fp32",
                    help="fp32: FP32-split (slices rounded without bias); fp32-nearest: slices rounded to nearest")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU baseline: seconds per oracle process")
    ap.add_argument("--cpu-procs", type=int, default=0, help="CPU baseline processes (default: host cores, <= 16)")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the FP64 one-solve path (profiling runs: one context per process)")
    ap.add_argument("--config", choices=["c2", "c5"], default="c2",
                    help="c2: the served C2 batch (default); c5: the 80/20 2D/3D Poisson stream, measured")
    ap.add_argument("--rate", type=float, default=6.0, help="c5: arrivals per second")
    ap.add_argument("--duration", type=float, default=40.0, help="c5: seconds of arrivals")
    ap.add_argument("--scale", type=float, default=1e-3, help="c5: iteration budget scale")
    ap.add_argument("--pairs", action="store_true",
                    help="write profiles/c2_convergence_pairs.json from the two convergence profiles and exit")
    ap.add_argument("--converge", type=int, default=0, help="run N solves to convergence (profile mode)")
    ap.add_argument("--converge-seeds", default=None, help="droplet seeds of --converge (default 0..N-1)")
    ap.add_argument("--converge-perturb", type=float, default=0.0,
                    help="also run perturbed copies of --converge-perturb-seeds with this amplitude")
    ap.add_argument("--converge-perturb-seeds", default="", help="seeds of the perturbed copies")
    ap.add_argument("--converge-seconds", type=float, default=1000.0)
    ap.add_argument("--converge-out", default=None, help="output JSON of --converge (default: the profile)")
    ap.add_argument("--converge-fields", default=None, help="directory for the converged fields (npz)")
    ap.add_argument("--converge-state-in", default=None, help="continue the solves saved in this directory")
    ap.add_argument("--converge-state-out", default=None, help="save solves still running at the limit here")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (one process, one MI355X); --gpus %d is not supported" % args.gpus)
        sys.exit(2)
    if args.pairs:
        return convergence_pairs()

    import abi
    import configs
    import serve

    if args.config == "c5":
        return c5_line(abi, configs, args)

    # the CPU baseline runs first, in forked processes, before anything in
    # this process touches the GPU
    cpu = None
    if not args.converge and not args.no_cpu:
        cpu = cpu_baseline(abi, configs, args.cpu_seconds, args.cpu_procs, conv_stats("fp64"))
        log("cpu baseline: %d processes, %.2f s per C2 iteration" % (cpu["cores"], cpu["s_per_iteration"]))

    lib = abi.load_library()

    if args.converge:
        seeds = converge_seeds(args.converge_seeds) if args.converge_seeds else list(range(args.converge))
        converge_run(abi, configs, serve, lib, seeds, args.precision, args.converge_seconds,
                     args.converge_out or CONV_PROFILE % args.precision.replace("-", "_"), args.converge_fields,
                     args.converge_state_in, args.converge_state_out, args.converge_perturb,
                     converge_seeds(args.converge_perturb_seeds) if args.converge_perturb else ())
        return

    precision = {"fp32": abi.SCFT_FP32_SPLIT, "fp32-nearest": abi.SCFT_FP32_NEAREST}.get(args.precision, abi.SCFT_FP64)
    log("served pool: C2, %s, up to %d slots" % (args.precision, args.slots))
    pool = make_pool(abi, configs, serve, lib, args.slots, precision)
    B = pool.nslots
    dt, inst_iters = time_iterations(pool, args.warmup, args.steps)
    ms_per_step = 1000.0 * dt / args.steps
    inst_iter_per_s = inst_iters / dt
    log("timed: %d steps x %d slots in %.3f s (%.3f ms/step)" % (args.steps, B, dt, ms_per_step))
    sched = pool.batch.schedule()
    log(sched)

    # dominant kernel (the z operator) timed with HIP events on the pool's stream
    zt = pool.batch.time_launches(kind=0, reps=5)
    xt = pool.batch.time_launches(kind=1, reps=5)
    pool.close()

    # plain one-solve-at-a-time FP64 GPU path (the denominator of the 5x target)
    t_single = None
    if not args.no_single:
        single = make_pool(abi, configs, serve, lib, 1, abi.SCFT_FP64)
        dts, _ = time_iterations(single, 3, 30)
        single.close()
        t_single = dts / 30

    # iterations-to-converge distributions measured on the GPU, per precision;
    # no distribution stands in for another
    conv = conv_stats(args.precision)
    conv64 = conv_stats("fp64")
    value = p50 = p99 = None
    single_sph = None
    if conv:
        value = 3600.0 * inst_iter_per_s / conv["mean"]
        p50 = conv["p50"] * dt / args.steps
        p99 = conv["p99"] * dt / args.steps
    if conv64 and t_single:
        single_sph = 3600.0 / (conv64["mean"] * t_single)


    # achieved: the path's algorithmic bytes per unit (a propagator voxel-step
    # reads q_s and writes q_{s+1}: 2 b bytes, b = 4 in FP32-split, 8 in FP64;
    # E and K are per-block operands and not counted) times the voxel-steps one
    # z launch processes, over the z launch's duration.  The kernel itself also
    # reads E (8 B per voxel-step); its bytes are reported beside.
    b_el = 8.0 if args.precision == "fp64" else 4.0
    units = zt["voxel_steps"]
    alg_bytes = 2.0 * b_el * units
    ach = alg_bytes / (zt["avg_us"] * 1e-6) / 1e9 if zt["avg_us"] > 0 else None
    kern_gbs = zt["bytes"] / (zt["avg_us"] * 1e-6) / 1e9 if zt["avg_us"] > 0 else None
    # the whole iteration against the same peak: algorithmic bytes of every
    # instance-iteration over the measured time per instance-iteration
    it_bytes, it_terms = iteration_bytes(abi, lib, configs.params_c2(), b_el)
    it_gbs = it_bytes * inst_iter_per_s / 1e9
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 arithmetic, f32 propagator slices",
        "data": "synthetic",
        "config": {"workload": "2D thin-film nanodroplet served solves (C2): 256x1x128, NS 101, MASK 1, "
                               "dbk AB f=0.5 vol 0.15 + homo C 0.85 + brush S1 0.05, chiN AB 25 AC=BC 40, G1 0.01",
                   "slots": B, "precision": {"fp32": "fp32-split (FP32 slices, unbiased rounding)",
                                             "fp32-nearest": "fp32-split, slices rounded to nearest"}.get(args.precision,
                                                                                                           "fp64"),
                   "measured": {"instance_iterations_per_s": inst_iter_per_s,
                                "ms_per_batched_iteration": ms_per_step,
                                "schedule": sched},
                   "projection": {
                       "basis": "value = 3600 x measured instance-iterations/s / mean iterations-to-converge; "
                                "p50/p99 = iterations-to-converge percentiles x measured time per batched "
                                "iteration with every slot resident.  The distribution is measured on the GPU "
                                "by running C2 solves of this precision to ||dw||_rms < 1e-5 (" +
                                (conv["source"] if conv else "no profile: value null") + ")",
                       "iters_to_converge": conv,
                       "p50_time_to_converge_s": p50, "p99_time_to_converge_s": p99},
                   "accuracy": accuracy_summary(args.precision),
                   "fp64_single_solve": {"ms_per_iteration": 1000.0 * t_single if t_single else None,
                                         "solves_per_hour": single_sph,
                                         "iters_to_converge_source": conv64["source"] if conv64 else None},
                   "speedup_vs_fp64_single": (value / single_sph) if (value and single_sph) else None},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": (ach / PEAK_HBM_GBS) if ach else None,
                     "traffic": measured_traffic(sched, B, args.precision),
                     "kernel": dominant_kernel(sched),
                     "launch_us": zt["avg_us"], "voxel_steps_per_launch": units,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "kernel_bytes_per_launch": zt["bytes"], "kernel_GBs": kern_gbs,
                     "fp64_tflops": zt["flops"] / (zt["avg_us"] * 1e-6) / 1e12 if zt["avg_us"] > 0 else None,
                     "x_axis_kernel": {"launch_us": xt["avg_us"], "bytes": xt["bytes"],
                                       "GBs": xt["bytes"] / (xt["avg_us"] * 1e-6) / 1e9 if xt["avg_us"] > 0 else None}},
        "iteration_roofline": {"bytes_per_instance_iteration": it_bytes, "terms": it_terms,
                               "achieved": it_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": it_gbs / PEAK_HBM_GBS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
