"""Serving: a request loop over one device-resident BatchSolver.

Requests (a particle set, its parameters, its random stream and its number of
steps) arrive over time.  The loop admits waiting requests while the device
has room for their rows (admission by free HBM), merges them into the running
batch, advances the batch until the next request finishes (or a tick of
steps), writes finished requests' rows back into their particle sets and drops
them from the device.  Rows never leave the device while a request runs.

Admission goes through the queue in arrival order and takes every request
that fits, so small requests pass a large one that waits for room; once the
oldest waiting request has waited reserve_after_s, nothing passes it any more
and the room freed by finishing requests goes to it (no starvation).

Each request's result is the result of a plain Solver run of that request
alone with the same stream and steps (with frozen fields, at any admission
time; with time-dependent fields, for a Solver built at the batch time at
which it was admitted).
"""
import time
from collections import deque

import numpy as np

from .batch import BatchSolver

# device bytes one particle row holds (both state copies, slot tables, scans,
# sort spares, normals and event scratch), with margin
ROW_BYTES = {2: 480, 3: 640}


class Request:
    """particle_set may be None with `make` a callable that builds it at
    admission (`npart` and `dim` its size), so that queued requests hold no rows."""

    def __init__(self, particle_set, parameters, rng_state, nsteps, arrival=0., make=None, npart=None, dim=None):
        if particle_set is None and (make is None or npart is None or dim is None):
            raise ValueError("a request without a particle set needs make, npart and dim")
        self.particle_set = particle_set
        self.make = make
        self._npart = npart
        self._dim = dim
        self.parameters = parameters
        self.rng_state = rng_state
        self.nsteps = int(nsteps)
        self.arrival = float(arrival)     # seconds after the loop starts
        self.id = None                    # batch scenario id once admitted
        self.admitted = None              # seconds after the loop starts
        self.finished = None
        self.final_rng_state = None

    @property
    def npart(self):
        return self.particle_set.npart if self.particle_set is not None else int(self._npart)

    @property
    def dim(self):
        return self.particle_set.dim if self.particle_set is not None else int(self._dim)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class ServeLoop:
    """Continuous batching over a field set (one context on one GPU).

    capacity: rows the device may hold (default: hbm_fraction of the free HBM
    after the mesh and fields are resident, at ROW_BYTES per row; a reshape
    holds the old and the new pool at once, hence the default fraction),
    capped at max_rows if given (a bound on the step time).
    tick_steps: most steps between two admission points.
    keep_results: False drops a finished request's rows after collecting them.
    bypass: False admits strictly first come, first served (a request that does
    not fit blocks the ones behind it); True lets later requests that fit pass
    it until it has waited reserve_after_s seconds."""

    def __init__(self, field_set, device=0, capacity=None, hbm_fraction=0.4, tick_steps=50, sort_interval=8,
                 clock=time.perf_counter, log=None, batch_factory=BatchSolver, keep_results=True, max_rows=None,
                 bypass=True, reserve_after_s=10.):
        self.field_set = field_set
        self.keep_results = keep_results
        self.max_rows = max_rows
        self.bypass = bypass
        self.reserve_after_s = reserve_after_s
        self.batch_factory = batch_factory
        self.device = device
        self.capacity = capacity
        self.hbm_fraction = hbm_fraction
        self.tick_steps = int(tick_steps)
        self.sort_interval = sort_interval
        self.clock = clock
        self.log = log
        self.batch = None
        self.done = []
        self.steps = 0
        self.particle_steps = 0
        self.max_running = 0

    # -- device side --------------------------------------------------------------
    def _resident(self):
        return 0 if self.batch is None else self.batch.npart

    def _room(self, dim):
        if self.capacity is None:
            free = self.batch.ctx.device_info()[1] if self.batch is not None else None
            if free is None:
                return None  # decided once the first batch exists
            self.capacity = int(self.hbm_fraction * (free + self._resident() * ROW_BYTES[dim]) / ROW_BYTES[dim])
            if self.max_rows is not None:
                self.capacity = min(self.capacity, int(self.max_rows))
        return self.capacity - self._resident()

    def _admit(self, reqs, now):
        for r in reqs:
            if r.particle_set is None:
                r.particle_set = r.make()
        psets = [r.particle_set for r in reqs]
        params = [r.parameters for r in reqs]
        streams = [r.rng_state for r in reqs]
        nsteps = [r.nsteps for r in reqs]
        if self.batch is None:
            self.batch = self.batch_factory(self.field_set, psets, params, streams, nsteps=nsteps, device=self.device,
                                            sort_interval=self.sort_interval)
            ids = self.batch.ids
        else:
            finished = [s for s in range(self.batch.nscen) if self.batch.steps_left(s) == 0]
            keep = [s for s in range(self.batch.nscen) if s not in set(finished)]
            ids = self.batch.reshape(keep, psets, params, streams, nsteps)
        for r, sid in zip(reqs, ids):
            r.id = sid
            r.admitted = now

    def _finish(self, running, now):
        """Collect the requests whose steps are done; drop them from the device
        unless nothing else would remain there."""
        b = self.batch
        fin = [s for s in range(b.nscen) if b.steps_left(s) == 0 and b.ids[s] in running]
        for s in fin:
            r = running.pop(b.ids[s])
            b.collect(s)
            r.final_rng_state = b.rng_state(s)
            r.finished = now
            if not self.keep_results:
                r.particle_set = None
            self.done.append(r)
        gone = [s for s in range(b.nscen) if b.steps_left(s) == 0]
        if gone and len(gone) < b.nscen:
            b.retire(gone)
        return fin

    # -- loop -----------------------------------------------------------------------
    def serve(self, requests, horizon=None):
        """Run the loop until every request is done, or `horizon` seconds after
        the start (requests not done then are left unfinished)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        running = {}
        t0 = self.clock()
        while True:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            if self.batch is not None:
                self._finish(running, now)
            # admission in arrival order while the device has room
            adm = []
            dim = waiting[0].dim if waiting else 2
            room = self._room(dim)
            if room is None:
                if waiting:
                    adm.append(waiting.popleft())  # the first batch fixes the capacity
            else:
                adm = self._pick(waiting, room, now)
            if waiting and not adm and not running:
                adm.append(waiting.popleft())  # larger than the device's room: it runs alone
            if adm:
                self._admit(adm, now)
                for r in adm:
                    running[r.id] = r
                self.max_running = max(self.max_running, len(running))
                if self.log:
                    self.log("serve: t={:.2f} s admitted {} ({} rows resident, {} running, {} waiting)".format(
                        now, len(adm), self._resident(), len(running), len(waiting)))
                if room is None:
                    continue  # the device's room is known now: admit the others before stepping
            if horizon is not None and now >= horizon:
                break
            if running:
                b = self.batch
                left = [b.steps_left(s) for s in range(b.nscen) if b.ids[s] in running]
                k = min(self.tick_steps, min(left))
                active = sum(b.particle_sets[s].npart for s in range(b.nscen) if b.ids[s] in running)
                b.run(k)
                self.steps += k
                self.particle_steps += k * active
            elif pending:
                time.sleep(max(0., min(pending[0].arrival - (self.clock() - t0), 0.01)))
            elif not waiting:
                break
        self.elapsed = self.clock() - t0
        self.unfinished = list(running.values()) + list(waiting) + list(pending)
        return self.done

    def _pick(self, waiting, room, now):
        """Requests admitted now (removed from `waiting`): in arrival order, each
        one that fits the room left; one that does not fit blocks the rest
        unless bypassing is on and it has waited less than reserve_after_s."""
        adm, free = [], room
        for r in list(waiting):
            if r.npart <= free:
                adm.append(r)
                free -= r.npart
            elif not self.bypass or now - r.arrival >= self.reserve_after_s:
                break
        for r in adm:
            waiting.remove(r)
        return adm

    def close(self):
        if self.batch is not None:
            self.batch.close()
            self.batch = None


def latency_stats(done, unfinished=(), horizon=None):
    """Latency percentiles of the completed requests, and over every arrival:
    a request unfinished at `horizon` (seconds after the start) counts with
    horizon - arrival, a lower bound of its latency, so the all-arrival
    percentiles are lower bounds of the true ones (censored at the horizon)."""
    lat = np.array([r.latency for r in done if r.latency is not None])
    out = dict(completed=int(lat.size), p50_s=None, p99_s=None)
    if lat.size:
        out.update(p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
                   mean_s=float(lat.mean()), max_s=float(lat.max()))
    if horizon is not None:
        cens = np.array([horizon - r.arrival for r in unfinished if r.arrival <= horizon])
        allv = np.concatenate([lat, cens])
        out.update(arrived=int(allv.size), unfinished=int(cens.size),
                   completed_frac=float(lat.size / allv.size) if allv.size else None)
        if allv.size:
            out.update(p50_censored_s=float(np.percentile(allv, 50)), p99_censored_s=float(np.percentile(allv, 99)))
    return out
