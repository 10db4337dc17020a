"""Serving loop for SCFT trial solves (the phase-diagram workflow of
README.md:59-128), on top of the libscft batch ABI.

A Pool owns one batch context: slots sized from the device's free HBM and
each slot's byte footprint (scft_slot_bytes), all slots sharing one chain
structure.  Requests (parameters + a seed field) queue up; after every
batched iteration, converged or failed slots are retired and waiting
requests are admitted into the freed slots (continuous batching), so a
request that arrives while others run starts on the next iteration.  Every
active slot advances in the same launches; the iteration is one hipGraph
replay while the active set is unchanged.

Convergence follows the project's criterion (the reference itself runs a
fixed ITER): ||w_new - w_old||_rms < tol over the active fields.

Iterations of a request are numbered from 1 as in a fresh reference run
(SCFT.c:1088-1090): the number selects the dbk1 seed box (ENABLE_SEED) and
MF.  A request whose parameters have RELAXATION > 0 gets its slot's own
device noise stream (scft_slot_rng_seed), seeded from `Request.rng` or,
by default, drand48's initial state; its first RELAXATION iterations add
the reference's noise.
"""
import collections
import ctypes as C
import math
import os
import time

import numpy as np

import abi
from abi import grid_of

SEED_DROPLET = "droplet"
SEED_LAMELLA = "lamella"
SEED_FIELDS = "fields"
SEED_FROM = "from"        # another request's converged field
DRAND48_INITIAL = (0x330E, 0xABCD, 0x1234)   # POSIX drand48 state before any seeding


class Request:
    __slots__ = ("rid", "params", "seed", "mask", "arrival", "slot", "admit_t", "done_t", "iters",
                 "F", "status", "fields", "dw", "history", "keep_fields", "start_iter", "rng", "max_iters",
                 "F64", "dF", "dw64")

    def __init__(self, rid, params, seed, mask=None, arrival=None, keep_fields=False, rng=None, max_iters=None):
        self.rid = rid
        self.params = params
        self.seed = seed              # (kind, arg)
        self.mask = mask
        self.arrival = arrival
        self.slot = None
        self.admit_t = None
        self.done_t = None
        self.iters = 0
        self.F = None
        self.dw = None
        self.status = "queued"
        self.fields = None
        self.history = []
        self.keep_fields = keep_fields
        self.start_iter = 1           # number of the request's first iteration
        self.rng = rng                # drand48 state (3 x u16) of its noise stream, or None
        self.max_iters = max_iters    # iteration budget of this request (None: the pool's)
        self.F64 = None               # FP64 free energy of the converged state (Pool fp64_check)
        self.dF = None                # |F - F64|
        self.dw64 = None              # FP64 ||dw||_rms from the converged state

    @property
    def time_to_converge(self):
        if self.done_t is None or self.arrival is None:
            return None
        return self.done_t - self.arrival


def slot_bytes(lib, params, max_nx, precision):
    b = C.c_size_t()
    abi.check(lib, lib.scft_slot_bytes(C.byref(params), max_nx, precision, C.byref(b)), "scft_slot_bytes")
    return b.value


def device_memory(lib):
    f, t = C.c_size_t(), C.c_size_t()
    abi.check(lib, lib.scft_device_memory(C.byref(f), C.byref(t)), "scft_device_memory")
    return f.value, t.value


def _jsonable(x):
    """Request ids may be tuples (the ICA workflow's): JSON keeps them as
    tagged lists."""
    if isinstance(x, tuple):
        return {"t": [_jsonable(v) for v in x]}
    return x


def _from_json(x):
    if isinstance(x, dict) and "t" in x:
        return tuple(_from_json(v) for v in x["t"])
    return x


def _write_rst(path, iteration, w, pf):
    w = np.ascontiguousarray(w, dtype=np.float64)
    pf = np.ascontiguousarray(pf, dtype=np.float64)
    with open(path, "wb") as f:
        f.write(b"SCFTRST1")
        f.write(np.array([iteration, pf.size], dtype=np.uint64).tobytes())
        f.write(w.tobytes())
        f.write(pf.tobytes())


def _read_rst(path):
    raw = open(path, "rb").read()
    if raw[:8] != b"SCFTRST1":
        raise ValueError("%s is not a full-precision restart file" % path)
    it, V = np.frombuffer(raw[8:24], dtype=np.uint64)
    data = np.frombuffer(raw[24:], dtype=np.float64)
    if data.size != 14 * int(V):
        raise ValueError("%s: %d values, %d expected" % (path, data.size, 14 * int(V)))
    return int(it), data[:13 * int(V)].reshape(13, int(V)).copy(), data[13 * int(V):].copy()


class Pool:
    """One batch of slots sharing a chain structure."""

    def __init__(self, lib, proto, precision=abi.SCFT_FP32_SPLIT, max_slots=256, max_nx=0,
                 hbm_fraction=0.85, tol=1e-5, max_iters=10 ** 6, record_every=0, diverge_dw=1e3,
                 fp64_check=False):
        self.lib = lib
        # FP32-split pools: every converged state is also iterated once in
        # FP64 (a one-slot FP64 batch), which gives each result its |dF|
        # against FP64 and FP64's own ||dw|| at that state
        self.fp64_check = bool(fp64_check) and precision != abi.SCFT_FP64
        self._check = None
        self.proto = proto
        self.precision = precision
        self.tol = tol
        self.max_iters = max_iters
        self.record_every = record_every
        self.diverge_dw = diverge_dw
        self.per_slot = slot_bytes(lib, proto, max_nx, precision)
        free, total = device_memory(lib)
        if self.fp64_check:
            free -= slot_bytes(lib, proto, max_nx, abi.SCFT_FP64)
        fit = int(free * hbm_fraction // self.per_slot)
        self.nslots = max(1, min(max_slots, fit))
        self.batch = abi.Batch(lib, proto, self.nslots, max_nx=max_nx, precision=precision)
        self.free_slots = collections.deque(range(self.nslots))
        self.running = {}             # slot -> Request
        self.queue = collections.deque()
        self.done = []
        self.by_id = {}
        self.iterations = 0           # batched iterations run
        self.slot_iterations = 0      # instance-iterations run
        self.t_iterate = 0.0
        self._inflight = False        # an iteration launched by step_launch, not yet completed
        self._t_launch = 0.0

    # ------------------------------------------------------------ requests
    def submit(self, req):
        if req.arrival is None:
            req.arrival = time.perf_counter()
        self.queue.append(req)
        self.by_id[req.rid] = req

    def _seed_fields(self, req, s):
        kind, arg = req.seed
        import configs
        p = req.params
        if kind == SEED_DROPLET:
            w, pf = configs.droplet_fields(p, **(arg or {}))
            self.batch.set_fields(s, w, pf)
        elif kind == SEED_LAMELLA:
            w, pf = configs.lamella_fields(p, **(arg or {}))
            self.batch.set_fields(s, w, pf)
        elif kind == SEED_FIELDS:
            w, pf = arg
            self.batch.set_fields(s, w, pf)
        elif kind == SEED_FROM:
            src = self.by_id[arg]
            if src.fields is None:
                raise ValueError("seed request %r kept no converged fields" % (arg,))
            self.batch.set_fields(s, *src.fields)
        else:
            raise ValueError("unknown seed kind %r" % (kind,))
        if p.mask:
            import configs as cf
            self.batch.set_mask(s, req.mask if req.mask is not None else cf.brush_mask(p))

    def _admit(self):
        while self.queue and self.free_slots:
            req = self.queue[0]
            if req.seed[0] == SEED_FROM:
                src = self.by_id.get(req.seed[1])
                if src is not None and src.status in ("queued", "running"):
                    # its seed has not converged yet: keep order, try later
                    break
                if src is None or src.fields is None:
                    # the seed request ended without converged fields kept
                    # (failed, out of budget, keep_fields off, or unknown):
                    # this request cannot start; it fails without a slot
                    self.queue.popleft()
                    req.status = "seed_failed"
                    req.done_t = time.perf_counter()
                    self.done.append(req)
                    continue
            self.queue.popleft()
            s = self.free_slots.popleft()
            self.batch.init_slot(s, req.params)
            self._seed_fields(req, s)
            if req.seed[0] == SEED_FIELDS:
                req.seed = (SEED_FIELDS, None)   # the slot holds the fields now
            if req.params.relaxation > 0 or req.rng is not None:
                self.batch.rng_seed(s, req.rng if req.rng is not None else DRAND48_INITIAL)
            self.batch.activate(s, True)
            req.slot = s
            req.admit_t = time.perf_counter()
            req.status = "running"
            self.running[s] = req

    def _retire(self, s, status, out, now):
        req = self.running.pop(s)
        req.status = status
        req.done_t = now
        req.F = out.free_energy
        req.dw = out.dw_rms
        if status == "converged" and (req.keep_fields or self.fp64_check):
            fields = self.batch.get_fields(s)
            if req.keep_fields:
                req.fields = fields
            if self.fp64_check:
                self._fp64_eval(req, fields)
        self.batch.activate(s, False)
        self.free_slots.append(s)
        req.slot = None
        self.done.append(req)

    def _fp64_eval(self, req, fields):
        """One FP64 iteration from the converged fields: F64 (F of that state
        in FP64 arithmetic throughout, the oracle's to 1e-10), |F - F64| and
        the FP64 ||dw|| there."""
        import configs
        if self._check is None:
            self._check = abi.Batch(self.lib, self.proto, 1, max_nx=self.batch.max_nx, precision=abi.SCFT_FP64)
        c = self._check
        c.init_slot(0, req.params)
        c.set_fields(0, *fields)
        if req.params.mask:
            c.set_mask(0, req.mask if req.mask is not None else configs.brush_mask(req.params))
        c.activate(0, True)
        out = c.iterate([req.start_iter + req.iters])[0]
        c.activate(0, False)
        req.F64 = out.free_energy
        req.dF = abs(req.F - out.free_energy)
        req.dw64 = out.dw_rms

    # ---------------------------------------------------------- iteration
    def step(self):
        """Admit, run one batched iteration, retire. Returns #active slots."""
        self.step_launch()
        return self.step_complete()

    def step_launch(self):
        """Admit and start one batched iteration on the pool's stream; returns
        at once (step_complete waits for it).  Returns #slots launched."""
        if self._inflight:
            raise RuntimeError("step_launch before the previous iteration's step_complete")
        self._admit()
        if not self.running:
            return 0
        its = self.batch.iteration_numbers
        for s, req in self.running.items():
            its[s] = req.start_iter + req.iters
        self._t_launch = time.perf_counter()
        self.batch.launch(its)
        self._inflight = True
        return len(self.running)

    def step_complete(self):
        """Wait for the iteration step_launch started and retire converged,
        failed and finished slots.  Returns #active slots."""
        if not self._inflight:
            return 0
        self._inflight = False
        outs = self.batch.complete()
        now = time.perf_counter()
        self.t_iterate += now - self._t_launch
        self.iterations += 1
        self.slot_iterations += len(self.running)
        for s in list(self.running):
            req = self.running[s]
            req.iters += 1
            o = outs[s]
            if self.record_every and req.iters % self.record_every == 0:
                req.history.append((req.iters, o.free_energy, o.dw_rms))
            if not math.isfinite(o.free_energy) or not math.isfinite(o.dw_rms) or o.dw_rms > self.diverge_dw:
                self._retire(s, "failed", o, now)
            elif o.dw_rms < self.tol:
                self._retire(s, "converged", o, now)
            elif req.iters >= (req.max_iters or self.max_iters):
                self._retire(s, "max_iters", o, now)
        return len(self.running)

    # ------------------------------------------------------- checkpoint
    def checkpoint(self, directory):
        """Save every running and queued request so that resume() continues
        each exactly where it stopped (the iteration is deterministic and its
        state is the fields, plus the noise stream's position): parameters,
        iteration count and numbering, history, budget, noise-stream state,
        and the fields at full precision in the drop-in driver's
        restart_fields.bin layout ("SCFTRST1", u64 iteration, u64 V, then
        w[13][V] and p[V] as FP64).  Call it between iterations."""
        import json
        import os
        if self._inflight:
            raise RuntimeError("checkpoint with an iteration in flight: step_complete first")
        os.makedirs(directory, exist_ok=True)
        entries = []

        def put(k, req, fields):
            e = {"rid": _jsonable(req.rid), "params": bytes(req.params).hex(), "iters": req.iters,
                 "start_iter": req.start_iter, "max_iters": req.max_iters, "history": req.history,
                 "keep_fields": req.keep_fields, "rng": list(req.rng) if req.rng is not None else None,
                 "seed": None, "mask": None}
            if fields is not None:
                w, pf = fields
                e["fields"] = "req%d.bin" % k
                _write_rst(os.path.join(directory, e["fields"]), req.start_iter + req.iters - 1, w, pf)
            else:
                kind, arg = req.seed
                if kind == SEED_FIELDS:
                    e["fields"] = "req%d.bin" % k
                    _write_rst(os.path.join(directory, e["fields"]), 0, arg[0], arg[1])
                elif kind == SEED_FROM:
                    e["seed"] = [kind, _jsonable(arg)]
                else:
                    e["seed"] = [kind, arg]
            if req.mask is not None:
                e["mask"] = "req%d_mask.npy" % k
                np.save(os.path.join(directory, e["mask"]), np.asarray(req.mask))
            entries.append(e)

        k = 0
        for s in sorted(self.running):
            req = self.running[s]
            if req.params.relaxation > 0 or req.rng is not None:
                req.rng = self.batch.rng_state(s)
            put(k, req, self.batch.get_fields(s))
            k += 1
        for req in self.queue:
            put(k, req, None)
            k += 1
        with open(os.path.join(directory, "pool.json"), "w") as f:
            json.dump({"format": 1, "requests": entries}, f)

    def resume(self, directory):
        """Queue the requests of a checkpoint() directory, running ones first
        (each continues from its saved fields, iteration count and noise
        state), in their saved order."""
        import json
        import os
        with open(os.path.join(directory, "pool.json")) as f:
            data = json.load(f)
        out = []
        for e in data["requests"]:
            params = abi.ScftParams.from_buffer_copy(bytes.fromhex(e["params"]))
            mask = np.load(os.path.join(directory, e["mask"])) if e.get("mask") else None
            if e.get("fields"):
                _, w, pf = _read_rst(os.path.join(directory, e["fields"]))
                seed = (SEED_FIELDS, (w, pf))
            else:
                kind, arg = e["seed"]
                seed = (kind, _from_json(arg) if kind == SEED_FROM else arg)
            r = Request(_from_json(e["rid"]), params, seed, mask=mask, keep_fields=e["keep_fields"],
                        rng=tuple(e["rng"]) if e["rng"] is not None else None, max_iters=e["max_iters"])
            r.iters = e["iters"]
            r.start_iter = e["start_iter"]
            r.history = [tuple(h) for h in e["history"]]
            self.submit(r)
            out.append(r)
        return out

    def take_done(self):
        """Return the finished requests and forget them (the pool keeps no
        reference, so their fields can be freed), except those a queued
        request still waits on as its seed (SEED_FROM)."""
        out, self.done = self.done, []
        needed = {r.seed[1] for r in self.queue if r.seed[0] == SEED_FROM}
        for r in out:
            if r.rid not in needed:
                self.by_id.pop(r.rid, None)
        return out

    @property
    def inflight(self):
        return self._inflight

    def ready(self):
        """True when the launched iteration has finished on the GPU."""
        return self._inflight and self.batch.ready()

    def busy(self):
        return bool(self.running or self.queue)

    def run(self, until=None, deadline=None):
        """Step until no work is left (or `until()` is true, or the deadline)."""
        while self.busy():
            if until is not None and until():
                break
            if deadline is not None and time.perf_counter() > deadline:
                break
            if self.step() == 0 and self.queue:
                # queue head waits on a seed that cannot run: nothing to do
                break

    def close(self):
        self.batch.close()
        if self._check is not None:
            self._check.close()
            self._check = None


def request_from_run_dir(rid, directory, header="parameters.h", **kw):
    """A request for the run a reference user prepared in `directory`: its
    parameters.h (read at run time, paramsfile.parse), counter.dat,
    restart_fields.dat and, for MASK 1, wMask.dat.  Iterations are numbered
    on from counter.dat, and the budget is the reference's
    (counter + 1 ... ITER + 1, SCFT.c:1088-1090) unless max_iters is given."""
    import os
    import paramsfile
    p, run = paramsfile.load(os.path.join(directory, header))
    counter = int(open(os.path.join(directory, "counter.dat")).read().split()[0])
    if counter == 0:
        raise ValueError("counter.dat = 0: the reference starts such a run from clock-seeded fields; "
                         "give the request a seed instead")
    w, pf, mask = paramsfile.read_inputs(directory, p)
    kw.setdefault("max_iters", max(1, run["iter"] - counter + 1))
    r = Request(rid, p, (SEED_FIELDS, (w, pf)), mask=mask, **kw)
    r.start_iter = counter + 1
    return r


def percentiles(values, qs=(50, 99)):
    if not values:
        return {q: None for q in qs}
    v = np.sort(np.asarray(values, float))
    return {q: float(np.percentile(v, q)) for q in qs}


def structure_key(p):
    """Requests whose parameters share this key share one chain structure
    and can be slots of one batch (the library's same_structure rule: NS,
    MASK, LY, LZ, DR, whether there is an x axis, and per species whether
    it is present, its length and composition; LX and chi may differ)."""
    k = [p.ns, p.mask, p.ly, p.lz, p.dr_target, p.lx == 1.0]
    for i in range(4):
        k.append((p.vol_dbk[i] > 0, p.len_dbk[i] if p.vol_dbk[i] > 0 else 0.0,
                  p.f_dbk[i] if p.vol_dbk[i] > 0 else 0.0))
    for i in range(10):
        k.append((p.vol_homo[i] > 0, p.len_homo[i] if p.vol_homo[i] > 0 else 0.0))
    for i in range(3):
        k.append((p.vol_s[i] > 0, p.len_s[i] if p.vol_s[i] > 0 else 0.0))
    return tuple(k)


class Server:
    """Continuous batching over requests of several chain structures (e.g.
    2D and 3D trial solves in one stream): one Pool per structure, created
    on the structure's first request and sized against the HBM still free
    then (capped by `slots`, a dict structure-key -> slots, or
    `default_slots`).  Requests with an arrival time in the future wait in
    `pending` until the server's clock reaches it (run_stream)."""

    def __init__(self, lib, precision=abi.SCFT_FP32_SPLIT, default_slots=192, slots=None, max_nx=0,
                 hbm_fraction=0.85, tol=1e-5, max_iters=10 ** 6):
        self.lib = lib
        self.precision = precision
        self.default_slots = default_slots
        self.slots = slots or {}
        self.max_nx = max_nx
        self.hbm_fraction = hbm_fraction
        self.tol = tol
        self.max_iters = max_iters
        self.pools = {}
        self.pending = []         # (arrival, seq, request) not yet arrived
        self.progressed = False   # whether the last advance() completed or launched anything
        self.rejected = []
        self._seq = 0
        self.t0 = None

    def pool_for(self, req):
        key = structure_key(req.params)
        pool = self.pools.get(key)
        if pool is None:
            # a new pool times its z-operator candidates when it is created:
            # finish the iterations in flight first, so they do not compete
            for other in self.pools.values():
                if other.inflight:
                    other.step_complete()
            nx = grid_of(req.params)[0]
            # Pools of different structures run concurrently here, each on
            # its own stream, so the server's pools are created with two
            # library defaults changed (unless the variables are set; the
            # library reads them when a batch is created):
            # - SCFT_FILM=0: the fused film step wins on an idle GPU where few
            #   chains share a launch, but next to another pool's kernels it
            #   lost: at 4 and 6 requests/s of the 80/20 2D/3D stream, 2D p50
            #   1.58 and 2.68 s with it against 1.28 and 2.20 s with the z and
            #   x launches apart;
            # - SCFT_PARX=0: a pool of uneven NX would run its x launches side
            #   by side on extra streams, which share the process's few
            #   hardware queues with the other pools' streams: p50 3.37 /
            #   8.39 s at 4 / 6 requests/s against 1.60 / 2.64 s without.
            pins = {k: "0" for k in ("SCFT_FILM", "SCFT_PARX") if k not in os.environ}
            os.environ.update(pins)
            try:
                pool = Pool(self.lib, req.params, precision=self.precision,
                            max_slots=self.slots.get(key, self.default_slots), max_nx=max(self.max_nx, nx),
                            hbm_fraction=self.hbm_fraction, tol=self.tol, max_iters=self.max_iters)
            finally:
                for k in pins:
                    os.environ.pop(k, None)
            self.pools[key] = pool
        return pool

    def submit(self, req):
        pool = self.pool_for(req)
        if grid_of(req.params)[0] > pool.batch.max_nx:
            req.status = "rejected"
            self.rejected.append(req)
            return
        pool.submit(req)

    def busy(self):
        return bool(self.pending) or any(p.busy() for p in self.pools.values())

    def step(self):
        """One batched iteration of every busy pool: all are launched before
        any is waited for, so the pools' iterations run on the GPU
        concurrently (each batch has its own stream)."""
        launched = [pool for pool in self.pools.values() if pool.busy() and pool.step_launch() > 0]
        return sum(pool.step_complete() for pool in launched)

    def advance(self):
        """Without waiting: retire the iteration of every pool whose launch has
        finished, and launch the next iteration of every idle pool with work,
        so that each pool advances at its own pace (a fast 2D pool does not
        wait for a slow 3D one).  Returns the number of running slots."""
        n = 0
        self.progressed = False
        for pool in self.pools.values():
            if pool.inflight and pool.ready():
                pool.step_complete()
                self.progressed = True
            if not pool.inflight and pool.busy():
                if pool.step_launch() > 0:
                    self.progressed = True
            n += len(pool.running)
        return n

    def stuck(self):
        """True when requests wait but none can ever start: nothing in
        flight, nothing running, and every queue head waits on a seed
        request that cannot finish (SEED_FROM on a request not yet run)."""
        if self.pending:
            return False
        return any(p.queue for p in self.pools.values()) and \
            not any(p.inflight or p.running for p in self.pools.values())

    def done(self):
        return [r for p in self.pools.values() for r in p.done]

    def run_stream(self, requests, deadline=None):
        """Serve requests whose `arrival` is a time offset (s) from now: each
        is submitted when the clock passes it, while earlier ones run."""
        import heapq
        self.t0 = time.perf_counter()
        for r in requests:
            heapq.heappush(self.pending, (r.arrival, self._seq, r))
            self._seq += 1
        while self.busy():
            now = time.perf_counter() - self.t0
            while self.pending and self.pending[0][0] <= now:
                _, _, r = heapq.heappop(self.pending)
                r.arrival = self.t0 + r.arrival      # absolute, as Pool.submit expects
                self.submit(r)
            if deadline is not None and now > deadline:
                break
            n = self.advance()
            if self.stuck():
                raise RuntimeError("serving stream stuck: queued requests wait on seeds that cannot run")
            if n == 0 and self.pending:
                time.sleep(min(0.001, max(0.0, self.pending[0][0] - now)))
            elif not self.progressed:
                time.sleep(50e-6)   # iterations take milliseconds: poll, do not spin

    def close(self):
        for p in self.pools.values():
            p.close()
        self.pools = {}
