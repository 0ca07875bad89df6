"""Host logic of the serving loop (admission by room, first come first served,
finishing, retirement) against a stand-in batch that counts steps."""
import numpy as np

from cylag_amd.serve import Request, ServeLoop, latency_stats


class _PSet:
    def __init__(self, n):
        self.npart = n
        self.dim = 2


class _Ctx:
    def device_info(self):
        return "stand-in", 10_000 * 480, 0


class FakeBatch:
    """Scenarios with step counters; every run(k) advances each by k (at most its steps)."""
    log = []

    def __init__(self, field_set, psets, params, streams, nsteps, device=0, sort_interval=8):
        self.ctx = _Ctx()
        self.sc = []
        self.next = 0
        self._add(psets, nsteps)

    def _add(self, psets, nsteps):
        ids = []
        for p, k in zip(psets, nsteps):
            self.sc.append(dict(id=self.next, pset=p, left=int(k)))
            ids.append(self.next)
            self.next += 1
        return ids

    nscen = property(lambda self: len(self.sc))
    ids = property(lambda self: [c["id"] for c in self.sc])
    particle_sets = property(lambda self: [c["pset"] for c in self.sc])
    npart = property(lambda self: sum(c["pset"].npart for c in self.sc))

    def steps_left(self, s):
        return self.sc[s]["left"]

    def run(self, k):
        assert k >= 1
        FakeBatch.log.append((k, [c["id"] for c in self.sc if c["left"] > 0]))
        for c in self.sc:
            c["left"] = max(0, c["left"] - k)

    def collect(self, s):
        assert self.sc[s]["left"] == 0

    def rng_state(self, s):
        return (s, 1)

    def retire(self, pos):
        self.sc = [c for i, c in enumerate(self.sc) if i not in set(pos)]

    def reshape(self, keep, psets, params, streams, nsteps):
        self.sc = [self.sc[s] for s in keep]
        return self._add(psets, nsteps)

    def close(self):
        pass


class FakeClock:
    def __init__(self):
        self.t = 0.

    def __call__(self):
        self.t += 0.001
        return self.t


def make(n, steps, arrival):
    return Request(_PSet(n), {}, (1, 1), steps, arrival)


def test_serve_admits_by_room_and_overlaps():
    FakeBatch.log = []
    reqs = [make(3000, 20, 0.), make(3000, 5, 0.), make(5000, 10, 0.), make(1000, 3, 0.002)]
    loop = ServeLoop(None, capacity=8000, tick_steps=4, clock=FakeClock(), batch_factory=FakeBatch)
    done = loop.serve(reqs)
    assert len(done) == 4 and not loop.unfinished
    # the first two fit together; the 5000-row request waits for room (first come, first served)
    first = FakeBatch.log[0][1]
    assert 0 in first and 1 in first and 2 not in first
    assert loop.max_running >= 2
    # every request ran exactly its steps while resident
    for r in reqs:
        ran = sum(k for k, ids in FakeBatch.log if r.id in ids)
        assert ran == r.nsteps, (r.id, ran, r.nsteps)
        assert r.finished >= r.admitted >= r.arrival
    st = latency_stats(done)
    assert st["completed"] == 4 and st["p99_s"] >= st["p50_s"] > 0


def test_serve_runs_oversized_request_alone_and_stops_at_horizon():
    FakeBatch.log = []
    reqs = [make(20000, 6, 0.), make(100, 1000, 0.)]
    # reserve_after_s=0: nothing passes the waiting head, which runs alone as it never fits
    loop = ServeLoop(None, capacity=5000, tick_steps=10, clock=FakeClock(), batch_factory=FakeBatch,
                     reserve_after_s=0.)
    loop.serve(reqs, horizon=0.05)
    assert reqs[0].finished is not None
    assert reqs[1].finished is None and reqs[1] in loop.unfinished
    assert np.isfinite(loop.elapsed)


def test_serve_room_from_free_memory():
    """Without a given capacity the first batch fixes it from the free HBM;
    the other waiting requests join before the first step."""
    FakeBatch.log = []
    reqs = [make(2000, 8, 0.), make(2000, 8, 0.), make(2000, 8, 0.)]
    loop = ServeLoop(None, hbm_fraction=0.5, tick_steps=8, clock=FakeClock(), batch_factory=FakeBatch)
    loop.serve(reqs)
    # 0.5 * (free + resident rows) / row bytes = 0.5 * (10000 + 2000) rows = 6000 rows
    assert loop.capacity == 6000
    assert FakeBatch.log[0][1] == [0, 1, 2]


def test_small_requests_pass_a_blocked_large_one():
    """A 5000-row request that does not fit waits; a 2000-row request behind
    it that fits is admitted (bypass); with bypass off, or once the large
    request has waited reserve_after_s, the small one waits behind it."""
    def run(**kw):
        FakeBatch.log = []
        reqs = [make(5000, 200, 0.), make(5000, 10, 0.001), make(2000, 10, 0.002)]
        loop = ServeLoop(None, capacity=8000, tick_steps=4, clock=FakeClock(), batch_factory=FakeBatch, **kw)
        loop.serve(reqs)
        assert all(r.finished is not None for r in reqs)
        return reqs
    a, b, c = run()
    assert c.admitted < b.admitted and c.finished < b.admitted
    a, b, c = run(bypass=False)
    assert c.admitted >= b.admitted
    a, b, c = run(reserve_after_s=0.)
    assert c.admitted >= b.admitted


def test_latency_stats_censored_over_all_arrivals():
    done = [make(1, 1, a) for a in (0., 1., 2.)]
    for r, f in zip(done, (3., 5., 4.)):
        r.finished = f
    late = [make(1, 1, a) for a in (4., 8.)]
    st = latency_stats(done, late, horizon=10.)
    assert st["completed"] == 3 and st["arrived"] == 5 and st["unfinished"] == 2
    assert st["completed_frac"] == 3 / 5
    lat = np.array([3., 4., 2., 6., 2.])   # completed latencies, then horizon - arrival
    assert st["p50_censored_s"] == np.percentile(lat, 50)
    assert st["p99_censored_s"] == np.percentile(lat, 99)
    assert st["p50_s"] == 3.
