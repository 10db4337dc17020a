"""Writes the golden fixtures of this directory from the CPU oracle.

The reference cannot be built here (it needs FFTW3), so the fixtures are
the oracle's FP64 results on workloads too large to recompute inside a
GPU test (one C2 iteration takes ~15 s in the oracle).  Each fixture holds
the inputs' recipe (configs.py generator + seed), the per-iteration
scalars and a compact summary of the final fields.

    python tests/golden/make_golden.py        # c2_oracle.json
    python tests/golden/make_golden.py c1     # c1_oracle.json (C1, 2000 iterations)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _paths  # noqa: E402,F401
import numpy as np  # noqa: E402
import abi  # noqa: E402
import configs  # noqa: E402
import oracle  # noqa: E402


def sample_idx(V, n=97):
    return [int(i) for i in np.linspace(0, V - 1, n).round()]


def run(name, p, w, pf, mask, iters):
    st = oracle.Oracle(abi).state(p)
    st.set_fields(w, pf)
    if mask is not None:
        st.set_mask(mask)
    rows = []
    for it in iters:
        o = st.iterate(it)
        rows.append({"iter": it, "F": o.free_energy, "rmsdev": o.rmsdev, "dw_rms": o.dw_rms,
                     "qsum": list(o.qsum), "chempot": list(o.chempot)})
        print(name, it, o.free_energy, flush=True)
    wf, pfin = st.get_fields()
    phi, tot = st.get_densities()
    idx = sample_idx(st.V)
    act = [f for f in range(13) if st.derived.active[f]]
    return {"name": name, "iters": rows,
            "sample_idx": idx,
            "w_samples": {str(f): [wf[f][i] for i in idx] for f in act},
            "p_samples": [pfin[i] for i in idx],
            "phi_samples": {str(f): [phi[f][i] for i in idx] for f in act},
            "w_l2": {str(f): float(np.sqrt((wf[f] ** 2).sum())) for f in act},
            "phi_sum": {str(f): float(phi[f].sum()) for f in act}}


def run_trajectory(name, p, w, pf, mask, n, every):
    """n iterations from counter 1 (iterations 2..n+1), scalars every `every`."""
    st = oracle.Oracle(abi).state(p)
    st.set_fields(w, pf)
    if mask is not None:
        st.set_mask(mask)
    rows = []
    for it in range(2, n + 2):
        o = st.iterate(it)
        if (it - 1) % every == 0 or it == 2:
            rows.append({"iter": it, "F": o.free_energy, "rmsdev": o.rmsdev, "dw_rms": o.dw_rms,
                         "chempot": list(o.chempot)})
            if (it - 1) % (10 * every) == 0:
                print(name, it, o.free_energy, o.dw_rms, flush=True)
    wf, pfin = st.get_fields()
    phi, tot = st.get_densities()
    idx = sample_idx(st.V)
    act = [f for f in range(13) if st.derived.active[f]]
    return {"name": name, "iters": rows, "n": n,
            "sample_idx": idx,
            "w_samples": {str(f): [wf[f][i] for i in idx] for f in act},
            "p_samples": [pfin[i] for i in idx],
            "phi_samples": {str(f): [phi[f][i] for i in idx] for f in act},
            "w_l2": {str(f): float(np.sqrt((wf[f] ** 2).sum())) for f in act}}


def main_c1():
    """C1 (64x1x64 AB melt, chi N 20, G1 0.1) for its full 2000 iterations."""
    p = configs.params_c1()
    w, pf = configs.random_fields(p, seed=0, amp=0.5)
    out = {"c1_random_seed0": run_trajectory("c1_random_seed0", p, w, pf, None, 2000, 10)}
    out["c1_random_seed0"]["recipe"] = "configs.params_c1(); configs.random_fields(p, seed=0, amp=0.5)"
    with open(os.path.join(HERE, "c1_oracle.json"), "w") as f:
        json.dump(out, f)


def main():
    out = {}
    p = configs.params_c2()
    w, pf = configs.droplet_fields(p, radius=6.0, seed=0)
    out["c2_droplet_seed0"] = run("c2_droplet_seed0", p, w, pf, configs.brush_mask(p), [2, 3, 4])
    out["c2_droplet_seed0"]["recipe"] = "configs.params_c2(); configs.droplet_fields(p, radius=6.0, seed=0); configs.brush_mask(p)"
    with open(os.path.join(HERE, "c2_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "c1":
        main_c1()
    else:
        main()
