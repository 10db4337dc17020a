"""ICA phase-diagram workflow (README.md:59-128) over the serving pool.

The phase diagram is a grid over (chi_AB, chi_AC), duplicated across three
planes of substrate affinity.  Every grid point (plane, i, j) is one SCFT
solve and owns a JSON record with its current converged free energy, the
file holding the field that produced it, and its history.  A first pass
solves every point from a droplet seed (README.md:86-95); later passes try
every neighbour's converged field as the seed of a trial solve at the
point's own parameters and keep the trial when F_trial < F (README.md:99-118,
compared in FP64: F comes from the solver's FP64 reductions), until a pass
replaces nothing.  Replaced entries move into the record's history
(README.md:120-128).

Batching.  The README's loop is sequential; here every trial of a pass is
submitted to one `serve.Pool` and runs concurrently (continuous batching),
seeded from the state at the start of the pass.  Decisions are then applied
in the README's loop order (plane, i, j, then the neighbour order below):
for each point, each trial in turn replaces the state if its F is lower
than the current one by more than `eps` (1e-6 kT by default).  The margin
is this module's: a converged solve's F is only defined to the convergence
tolerance (a re-converged copy of the same state differs by ~1e-8), and a
bare F_trial < F then replaces states by copies of themselves and never
ends; 1e-6 is the accuracy every result is held to against FP64.  A
(point, neighbour) pair counts as tried for the neighbour's state version
it was seeded from; when the neighbour's state is later replaced, the pair
is tried again in the next pass.  The loop ends on a pass with no
replacement.

Trial seeds.  A trial starts from the neighbour's converged field plus a
seeded perturbation of amplitude `kick` (1e-3 by default, uniform, PCG64
seeded from the point, the neighbour and its version) on the active fields.
A neighbour's field relaxed at other parameters can sit near a saddle of
the point's free energy (a symmetric state); without a perturbation the
trial may settle on the saddle or escape from it depending on rounding-level
asymmetries alone, so FP64 and FP32 runs could decide differently.  With
the perturbation the escape is set by the seed, the same in both.

FP64 polishing (`polish`, FP32-split runs only).  A converged trial whose
F lies within `polish` of the current state's F is a close call.  Its
state and the current one are re-converged in FP64 (a second pool, seeded
with their converged fields), and the decision uses those two F values.
A replacing trial then enters with its FP64 F and fields.  The FP32-split
converged states are FP64 fixed points at the tolerance, so polishing
takes few iterations.

Precision.  In FP32-split runs the rounding of the FP32 slices matters at
saddles: FP64 carries the unstable mode of a symmetric state at its own
rounding level (2^-53), slices rounded to nearest (abi.SCFT_FP32_NEAREST)
do not carry it at all, and unbiased rounding (abi.SCFT_FP32_SPLIT) at
2^-24.  On small films, where FP64 stops some solves at saddles before it
leaves them, SCFT_FP32_NEAREST reproduces FP64's decisions and
SCFT_FP32_SPLIT leaves those saddles and decides differently.

Neighbours: in-plane (i-1, j), (i+1, j), (i, j-1), (i, j+1) and cross-plane
the same (i, j) on the previous and next plane of the cycle P0 -> P1 -> P2
-> P0 (README.md:76-79).

Planes.  The README names the planes (P0 near-90-degree contact angle, P1
near-dewetted, P2 near-wetted, README.md:69-74) but not their parameters;
this module sets them through the brush-droplet interaction:
chi_S1A = chi_S1B = chi_AC + delta with delta 0 (P0), +DELTA (P1: the
substrate repels the droplet) and -DELTA (P2: it attracts the droplet).
"""
import json
import os

import numpy as np

import abi
import configs
import serve

PLANES = ("P0", "P1", "P2")


class Point:
    __slots__ = ("plane", "i", "j", "params", "F", "field", "iteration", "version", "history", "tried",
                 "status", "iters")

    def __init__(self, plane, i, j, params):
        self.plane, self.i, self.j, self.params = plane, i, j, params
        self.F = None
        self.field = None          # path of the npz holding (w, p)
        self.iteration = 0         # ICA pass that produced the current state
        self.version = 0           # bumped on every replacement
        self.history = []          # replaced states: dicts (iteration, F, field)
        self.tried = set()         # (plane, i, j, version) of tried neighbour states
        self.status = None
        self.iters = 0             # SCFT iterations of the current state's solve

    @property
    def key(self):
        return (self.plane, self.i, self.j)

    def record(self, chi_ab, chi_ac):
        return {"plane": PLANES[self.plane], "i": self.i, "j": self.j, "chi_AB_N": chi_ab, "chi_AC_N": chi_ac,
                "F": self.F, "field": self.field, "iteration": self.iteration, "scft_iterations": self.iters,
                "status": self.status, "history": self.history,
                "tried": sorted("%s-%d-%d@%d" % (PLANES[t[0]], t[1], t[2], t[3]) for t in self.tried)}


class ICA:
    """The phase-diagram workflow over one serving pool.

    make_params(chi_ab, chi_ac, plane) -> scft_params of one point (every
    point must share the chain structure: same grid, NS and species);
    seed(point) -> the first-pass seed, a serve.Request seed tuple."""

    def __init__(self, lib, chi_ab, chi_ac, outdir, make_params=None, seed=None, precision=abi.SCFT_FP32_SPLIT,
                 max_slots=256, tol=1e-5, max_iters=400000, delta=15.0, base=None, eps=1e-6, kick=1e-3,
                 polish=None, pool_factory=None):
        self.lib = lib
        self.pool_factory = pool_factory or serve.Pool
        self.chi_ab = list(chi_ab)
        self.chi_ac = list(chi_ac)
        self.outdir = outdir
        self.delta = delta
        self.eps = eps
        self.kick = kick
        self.base = base or {}
        self.make_params = make_params or self._default_params
        self.seed = seed or (lambda pt: (serve.SEED_DROPLET, {"radius": 6.0, "seed": 0}))
        self.precision = precision
        os.makedirs(os.path.join(outdir, "fields"), exist_ok=True)
        self.points = {}
        for pl in range(len(PLANES)):
            for i, a in enumerate(self.chi_ab):
                for j, c in enumerate(self.chi_ac):
                    self.points[(pl, i, j)] = Point(pl, i, j, self.make_params(a, c, pl))
        self._pool_args = dict(precision=precision, max_slots=max_slots, tol=tol, max_iters=max_iters)
        self.pool = None           # created on the first solve
        self.decisions = []        # (pass, point key, neighbour key, neighbour version, F_trial, F_before, replaced)
        self.passes = 0
        self.solves = 0
        # FP64 polishing of close calls in the FP32-split mode: a converged
        # trial whose F is within `polish` of the current state's is decided
        # on both states re-converged in FP64 (seeded with their fields)
        self.polish = polish if (polish and precision != abi.SCFT_FP64) else None
        self.pool64 = None
        self.f64 = {}              # (point key..., version) -> FP64 F of that state, once polished
        self.polished = []         # (pass, point key, neighbour tag, F_trial, F_before, F64_trial, F64_current, replaced)

    def _default_params(self, chi_ab, chi_ac, plane):
        d = (0.0, self.delta, -self.delta)[plane]
        return configs.chem_droplet(chi_ab=chi_ab, chi_ac=chi_ac, chi_s1a=chi_ac + d, chi_s1b=chi_ac + d,
                                    **self.base)

    def order(self):
        return [self.points[k] for k in sorted(self.points)]

    def neighbours(self, pt):
        out = []
        for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            k = (pt.plane, pt.i + di, pt.j + dj)
            if k in self.points:
                out.append(self.points[k])
        n = len(PLANES)
        for dp in (-1, 1):
            k = ((pt.plane + dp) % n, pt.i, pt.j)
            if k in self.points and self.points[k] is not pt and self.points[k] not in out:
                out.append(self.points[k])
        return out

    # ------------------------------------------------------------ solves
    def _save_field(self, tag, fields):
        path = os.path.join(self.outdir, "fields", tag + ".npz")
        np.savez_compressed(path, w=fields[0], p=fields[1])
        return os.path.relpath(path, self.outdir)

    def _load_field(self, rel):
        z = np.load(os.path.join(self.outdir, rel))
        return z["w"], z["p"]

    def _trial_seed(self, pt, nb):
        w, p = self._load_field(nb.field)
        if self.kick > 0.0:
            rng = np.random.Generator(np.random.PCG64(
                [pt.plane, pt.i, pt.j, nb.plane, nb.i, nb.j, nb.version]))
            w = w.copy()
            for f in range(w.shape[0]):
                if np.any(w[f] != 0.0):      # active fields only
                    w[f] += self.kick * rng.uniform(-1.0, 1.0, w.shape[1])
        return w, p

    def _run(self, requests):
        if self.pool is None:
            proto = next(iter(self.points.values())).params
            self.pool = self.pool_factory(self.lib, proto, **self._pool_args)
        for r in requests:
            self.pool.submit(r)
        self.pool.run()
        self.pool.take_done()    # the caller holds the requests; the pool keeps none
        self.solves += len(requests)

    def _polish(self, items):
        """Re-converge states in FP64 from their FP32-split fields; items are
        (key, params, (w, p)); returns key -> (F, fields), (None, None) when
        the FP64 solve does not converge."""
        if self.pool64 is None:
            proto = next(iter(self.points.values())).params
            self.pool64 = self.pool_factory(self.lib, proto, **dict(self._pool_args, precision=abi.SCFT_FP64))
        reqs = [serve.Request(("polish",) + tuple(k), p, (serve.SEED_FIELDS, f), keep_fields=True)
                for k, p, f in items]
        for r in reqs:
            self.pool64.submit(r)
        self.pool64.run()
        self.pool64.take_done()
        return {k: ((r.F, r.fields) if r.status == "converged" else (None, None)) for (k, _, _), r in zip(items, reqs)}

    def _current_f64(self, pt):
        """FP64 F of the point's current state (polished on demand)."""
        key = pt.key + (pt.version,)
        if key not in self.f64:
            self.f64[key] = self._polish([(("c",) + key, pt.params, self._load_field(pt.field))])[("c",) + key][0]
        return self.f64[key]

    def initial(self):
        """First pass: every point from its seed (README.md:86-95)."""
        reqs = []
        for pt in self.order():
            reqs.append(serve.Request(("init",) + pt.key, pt.params, self.seed(pt), keep_fields=True))
        self._run(reqs)
        for r, pt in zip(reqs, self.order()):
            pt.status = r.status
            pt.iters = r.iters
            if r.status == "converged":
                pt.F = r.F
                pt.field = self._save_field("%s_%d_%d_v0" % (PLANES[pt.plane], pt.i, pt.j), r.fields)
        self.write_records()

    def update_pass(self):
        """One pass of trials over every untried neighbour state; returns
        whether any point was replaced (README.md:99-118)."""
        self.passes += 1
        trials = []
        for pt in self.order():
            if pt.F is None:
                continue
            for nb in self.neighbours(pt):
                tag = nb.key + (nb.version,)
                if nb.F is None or tag in pt.tried:
                    continue
                r = serve.Request(("trial", self.passes) + pt.key + tag, pt.params,
                                  (serve.SEED_FIELDS, self._trial_seed(pt, nb)), keep_fields=True)
                trials.append((pt, nb, tag, r))
        if not trials:
            return False
        self._run([t[3] for t in trials])
        pol = {}
        if self.polish:
            # the close calls against the states at the start of the pass: the
            # trial states and the current states they compete with,
            # re-converged in FP64 together in one batch
            items, seen = [], set()
            for pt, nb, tag, r in trials:
                if r.status != "converged" or abs(r.F - pt.F) >= self.polish:
                    continue
                items.append((("t",) + pt.key + tag, pt.params, r.fields))
                ck = pt.key + (pt.version,)
                if ck not in self.f64 and ck not in seen:
                    seen.add(ck)
                    items.append((("c",) + ck, pt.params, self._load_field(pt.field)))
            if items:
                for k, v in self._polish(items).items():
                    if k[0] == "c":
                        self.f64[k[1:]] = v[0]
                    else:
                        pol[k] = v
        updated = False
        for pt, nb, tag, r in trials:
            pt.tried.add(tag)
            before = pt.F
            ok = r.status == "converged"
            tk = ("t",) + pt.key + tag
            F_new, fields_new = r.F, r.fields
            # closeness is judged against the point's state now: an earlier
            # trial of this pass may have replaced it
            close = bool(self.polish) and ok and abs(r.F - pt.F) < self.polish
            if close:
                if tk not in pol:   # a close call only against a state replaced in this pass
                    pol.update(self._polish([(tk, pt.params, r.fields)]))
                F_new, fields_new = pol[tk]
                F_cur = self._current_f64(pt)
                replace = F_new is not None and F_cur is not None and F_new < F_cur - self.eps
                self.polished.append((self.passes, pt.key, tag, r.F, before, F_new, F_cur, replace))
            else:
                replace = ok and r.F < pt.F - self.eps   # FP64 comparison
            self.decisions.append((self.passes, pt.key, nb.key, tag[3], r.F if ok else None, before, replace))
            if replace:
                pt.history.append({"iteration": pt.iteration, "F": pt.F, "field": pt.field})
                pt.version += 1
                pt.F = F_new
                pt.iters = r.iters
                pt.iteration = self.passes
                pt.field = self._save_field("%s_%d_%d_v%d" % (PLANES[pt.plane], pt.i, pt.j, pt.version),
                                            fields_new)
                if close:   # the new state's F is an FP64 re-converged value
                    self.f64[pt.key + (pt.version,)] = F_new
                updated = True
        self.write_records()
        return updated

    def run(self, max_passes=50):
        self.initial()
        while self.passes < max_passes and self.update_pass():
            pass
        return self

    def write_records(self):
        for pt in self.order():
            path = os.path.join(self.outdir, "%s_%d_%d.json" % (PLANES[pt.plane], pt.i, pt.j))
            with open(path, "w") as f:
                json.dump(pt.record(self.chi_ab[pt.i], self.chi_ac[pt.j]), f, indent=1)

    def close(self):
        for name in ("pool", "pool64"):
            pool = getattr(self, name)
            if pool is not None:
                pool.close()
                setattr(self, name, None)
