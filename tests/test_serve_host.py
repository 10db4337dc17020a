"""Host logic of the serving front (no GPU)."""


def test_structure_key_groups_by_chain_structure(configs):
    import serve
    a = configs.chem_droplet(lx=24.0, chi_ab=15.0)
    b = configs.chem_droplet(lx=40.0, chi_ab=30.0, chi_ac=50.0)
    c = configs.params_c5_3d()
    assert serve.structure_key(a) == serve.structure_key(b)      # LX and chi may differ
    assert serve.structure_key(a) != serve.structure_key(c)


def test_iteration_bytes_match_the_accounting_table(abi, configs, lib):
    """bench.iteration_bytes: 2 b V S + 2 b V P + 8 V (3 n + 2) with S, P, n
    from the derived chain structure; C2 and C3 in FP32-split give the
    per-instance-iteration bytes of the hot-path accounting (142.1 MB with
    S = 316, P = 212, n = 4; 10.65 GB with S = 734, P = 521)."""
    import bench
    b, t = bench.iteration_bytes(abi, lib, configs.params_c2(), 4.0)
    assert t == {"V": 32768, "S": 316, "P": 212, "n": 4}
    assert b == 142082048.0
    b3, t3 = bench.iteration_bytes(abi, lib, configs.params_c3(), 4.0)
    assert (t3["S"], t3["P"], t3["n"]) == (734, 521, 4)
    assert abs(b3 - 10.645e9) < 1e6


class _FakePool:
    """Stands in for serve.Pool in Server.advance: an iteration launched by
    step_launch is ready after `polls` calls of ready()."""

    def __init__(self, polls, work):
        self.polls, self.work = polls, work
        self.inflight, self.left = False, 0
        self.running = {0: "r"}
        self.launched = self.completed = 0

    def busy(self):
        return self.work > 0

    def ready(self):
        assert self.inflight
        self.left -= 1
        return self.left <= 0

    def step_launch(self):
        assert not self.inflight              # never two launches in flight
        self.inflight, self.left = True, self.polls
        self.launched += 1
        return 1

    def step_complete(self):
        assert self.inflight
        self.inflight = False
        self.completed += 1
        self.work -= 1
        if self.work == 0:
            self.running = {}
        return len(self.running)


def test_advance_lets_each_pool_run_at_its_own_pace():
    """Server.advance completes whichever pool has finished and relaunches
    it at once: a pool whose iterations take one poll runs five times as
    many iterations as one whose iterations take five, and both finish all
    their work."""
    import serve
    srv = serve.Server.__new__(serve.Server)
    fast, slow = _FakePool(1, 50), _FakePool(5, 10)
    srv.pools = {"2D": fast, "3D": slow}
    srv.pending = []
    for _ in range(25):
        srv.advance()
    assert fast.completed >= 4 * slow.completed > 0, (fast.completed, slow.completed)
    while srv.busy():
        srv.advance()
    assert (fast.completed, slow.completed) == (50, 10)
    assert not fast.inflight and not slow.inflight


def test_seed_from_a_request_without_fields_fails_without_a_slot():
    """A SEED_FROM request whose seed request ended without kept fields
    (failed, or keep_fields off) is marked seed_failed and takes no slot;
    one whose seed is still running waits at the queue head."""
    import collections
    import serve
    pool = serve.Pool.__new__(serve.Pool)
    pool.queue = collections.deque()
    pool.free_slots = collections.deque([0, 1])
    pool.by_id = {}
    pool.done = []
    pool.running = {}
    src_failed = serve.Request("a", None, (serve.SEED_DROPLET, None))
    src_failed.status = "failed"
    src_running = serve.Request("b", None, (serve.SEED_DROPLET, None))
    src_running.status = "running"
    pool.by_id = {"a": src_failed, "b": src_running}
    r1 = serve.Request("t1", None, (serve.SEED_FROM, "a"))
    r2 = serve.Request("t2", None, (serve.SEED_FROM, "b"))
    pool.queue.extend([r1, r2])
    pool.by_id.update({"t1": r1, "t2": r2})
    pool._admit()
    assert r1.status == "seed_failed" and r1 in pool.done and r1.slot is None
    assert list(pool.queue) == [r2] and r2.status == "queued"
    assert list(pool.free_slots) == [0, 1] and not pool.running
