"""Timing runs of the reference's workloads on one GPU, outside the C2
bench: per batched iteration and per launch of the z operator and the
last-axis transform (HIP events, scft_time_launches).

    python workloads.py c1 [--slots B] [--iters N] [--precision fp32|fp64]
        2D bulk AB diblock melt, 64x1x64, NS 101, MASK 0; B resident (1)
    python workloads.py c3 [--slots B] [--iters N] [--precision fp32|fp64]
        3D thin-film droplet, 128x128x64, NS 201 (C3 chemistry); B resident (1)
    python workloads.py c5 [--slots B] [--iters N] [--precision ...]
        3D served requests, 96x96x48 (C3 chemistry), B resident (8)
    python workloads.py c4 [--slots B] [--iters N] [--precision ...]
        phase-diagram sweep batch: B slots of the C2 chemistry with
        chi_AB N over 15..36 and chi_AC N over 20..55 (8 x 8 grid), LX drawn
        per slot from {24, 32, 40} R_g (seeded), so grids are 192/256/320 x
        1 x 128 of uneven size in one batch; B resident (192)

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
          "z launch %.1f us x%d (%.0f GB/s), last-axis launch %.1f us x%d; F[0]=%.10f; %s"
          % (name, len(protos), "fp32-split" if precision == abi.SCFT_FP32_SPLIT else "fp64", dt * 1e3,
             dt / len(protos) * 1e6, z["avg_us"], z["launches"], z["bytes"] / max(z["avg_us"], 1e-9) * 1e-3,
             x["avg_us"], x["launches"], outs[0].free_energy, b.schedule()), flush=True)
    b.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("workload", choices=["c1", "c3", "c4", "c5"])
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--slots", type=int, default=None)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    a = ap.parse_args(argv)
    lib = abi.load_library()
    prec = abi.SCFT_FP32_SPLIT if a.precision == "fp32" else abi.SCFT_FP64
    slots = a.slots or {"c1": 1, "c3": 1, "c4": 192, "c5": 8}[a.workload]
    if a.workload == "c1":
        run(lib, "C1 64x1x64", c1_params(slots), prec, a.iters)
    elif a.workload == "c3":
        run(lib, "C3 128x128x64", c3_params(slots), prec, a.iters)
    elif a.workload == "c5":
        run(lib, "C5 96x96x48", c5_params(slots), prec, a.iters)
    else:
        run(lib, "C4 sweep (NX 192/256/320)", c4_params(slots), prec, a.iters)


if __name__ == "__main__":
    main()
