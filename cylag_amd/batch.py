"""BatchSolver: many release scenarios on one mesh and field set, advanced
together by one device context.

A scenario is a LagrangianParticleSet with its own parameters (LSM1 or LSM2,
diffusion, boundary conditions...), its own random stream (state, inc) and its
own number of steps.  The scenarios' pools are concatenated into one device
pool; every step advances all scenarios that have steps left in a few kernel
launches (one per kernel class: the fused 2D LSM-1 and LSM-2 kernels, the
generic kernel per model), and the element re-sort, the event tables and the
normal generator are shared.  Each scenario's rows, random stream, deleted-row
tags and boundary statistics are the ones a plain Solver run of that scenario
alone gives (the reference's nearest analogue is SolverParallel's split of one
release into independently seeded ranks, solver_parallel.pyx:99-130).

Scenarios can leave and join a running batch (continuous batching): admit()
adds scenarios, collect() writes a scenario's rows back into its particle set,
retire() drops scenarios from the device.  Rows stay on the device throughout;
a scenario that joins later runs on the batch clock from its current time.

All scenarios share the time step, the frozen-fields flag, the particle
constants (density, drag, buoyancy...) and one rng_method.
"""
import ctypes as C

import numpy as np

from . import _hip
from .parameters import Parameters
from .simutime import SimuTime
from .solver import _params_struct, _pconst_struct, check_setup, upload_fields, upload_mesh


def _steps_of(simutime):
    """forward() calls Solver.solve() makes (solver.pyx:399-401, simutime.pyx:84)."""
    t, it, fin, n = simutime.time, simutime.iteration, simutime.is_finished, 0
    for _ in range(simutime.nt):
        if fin:
            break
        t += simutime.time_step
        it += 1
        n += 1
        fin = t >= simutime.final_time or it >= simutime.maxiteration
    return n


class _Scenario:
    __slots__ = ("id", "pset", "params", "nsteps", "done", "stats")

    def __init__(self, sid, pset, params, nsteps):
        self.id = sid
        self.pset = pset
        self.params = params
        self.nsteps = int(nsteps)   # steps the scenario runs in all
        self.done = 0               # steps run so far
        self.stats = []             # boundary statistics, one (k, nopen) block per run call


class BatchSolver:
    def __init__(self, field_set, particle_sets, parameters, rng_states, nsteps=None, device=0,
                 sort_interval=8, use_graph=True):
        if not (len(particle_sets) == len(parameters) == len(rng_states)) or not particle_sets:
            raise ValueError("one parameter dict and one (state, inc) per particle set")
        self.field_set = field_set
        params = [Parameters(p) for p in parameters]
        p0 = params[0]
        self._shared = {k: getattr(p0, k) for k in ("time_step", "initial_time", "frozen_eulerian_fields",
                                                      "rng_method")}
        self.dim = particle_sets[0].dim
        self._pc0 = None
        self._check(particle_sets, params, initial=True)
        if nsteps is None:
            nsteps = [_steps_of(SimuTime(False, p.initial_time, p.final_time, p.time_step)) for p in params]
        self.time = float(p0.initial_time)
        self.dt = float(p0.time_step)
        self._next_id = 0
        self.scen = [self._new_scenario(ps, p, k) for ps, p, k in zip(particle_sets, params, nsteps)]
        fs = field_set
        fs.frozen_hydro = bool(p0.frozen_eulerian_fields == 1)
        self.ctx = _hip.Context(device)
        lib, h = self.ctx.lib, self.ctx.h
        upload_mesh(self.ctx, fs.triangular_mesh)
        upload_fields(self.ctx, fs)
        self.ctx.check(lib.cyl_fields_reset(h, int(fs.frozen_hydro), self.time), "cyl_fields_reset")
        if not fs.frozen_hydro:
            self.ctx.check(lib.cyl_fields_interpolate(h, self.time), "cyl_fields_interpolate")
        rows = self._rows(particle_sets)
        P = _hip.ptr
        pc = _pconst_struct(particle_sets[0])
        n = int(sum(ps.npart for ps in particle_sets))
        self.ctx.check(lib.cyl_particles_upload(h, n, self.dim, P(rows["position"]), P(rows["velocity"]),
                                                P(rows["fluid_velocity_seen"]), P(rows["inactive"], "i"),
                                                P(rows["tri"], "i"), P(rows["lowerlayer"], "i"),
                                                P(rows["diameter"]), P(rows["volume"]), P(rows["init_tri"], "i"),
                                                C.byref(pc)), "cyl_particles_upload")
        base = np.concatenate([[0], np.cumsum([ps.npart for ps in particle_sets])]).astype(np.int32)
        pars, st, inc = self._tables(particle_sets, params, rng_states)
        ns = np.ascontiguousarray(nsteps, dtype=np.int32)
        self.ctx.check(lib.cyl_scenarios_set(h, len(particle_sets), P(base, "i"), C.cast(pars, C.c_void_p), st, inc,
                                             P(ns, "i")), "cyl_scenarios_set")
        self.set_tuning(sort_interval, 0, 0, use_graph)
        self.ctx.check(lib.cyl_init_particles(h, C.byref(pars[0]), self.time, self.dt), "cyl_init_particles")

    def set_tuning(self, sort_interval, blocks_per_cu=0, inline_normals=0, use_graph=None):
        """Element re-sort interval (steps, 0 = never), resident blocks per CU of
        the persistent kernels (0 = as many as fit), the source of the 2D fused
        kernels' normals (0 pool-order generator, 1 inline sampling, 2
        slot-order generator), and graph replay of multi-step calls (None:
        keep); results do not depend on them."""
        if use_graph is not None:
            self.use_graph = int(use_graph)
        self.ctx.check(self.ctx.lib.cyl_set_tuning(self.ctx.h, int(sort_interval), int(blocks_per_cu),
                                                   int(inline_normals), getattr(self, "use_graph", 0)),
                       "cyl_set_tuning")
        self.tuning = (int(sort_interval), int(blocks_per_cu), int(inline_normals))

    # -- scenario table -----------------------------------------------------------
    def _new_scenario(self, pset, params, nsteps):
        sc = _Scenario(self._next_id, pset, params, nsteps)
        self._next_id += 1
        return sc

    def _check(self, particle_sets, params, initial=False):
        for p in params:
            p.check_values()
            for k, v in self._shared.items():
                if k == "initial_time" and not initial:
                    continue
                if getattr(p, k) != v:
                    raise ValueError("the scenarios of a batch share {}".format(k))
        for ps, p in zip(particle_sets, params):
            check_setup(self.field_set, ps, p)
            if ps.dim != self.dim:
                raise ValueError("the particle sets of a batch share one dimension")
            ps.rho0 = p.water_density
            # the LSM-2 coefficients (functions of the shared physical constants)
            # are set for every scenario: a later LSM-2 arrival uses the context's
            ps._initialize_model_coefficients()
        # physical particle constants must agree
        phys = [f for f, _ in _hip.CylPconst._fields_[:8]]
        pcs = [tuple(getattr(_pconst_struct(ps), f) for f in phys) for ps in particle_sets]
        if self._pc0 is None:
            self._pc0 = pcs[0]
        if any(pc != self._pc0 for pc in pcs):
            raise ValueError("the particle sets of a batch share their particle constants")

    def _rows(self, particle_sets):
        """The particle sets' rows concatenated, every active row marked for
        initialize_particle_state (solver.pyx:193-197)."""
        fs = self.field_set
        for ps in particle_sets:
            ps._sync()
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([np.asarray(ps._h[k], dtype=dt)
                                                                 for ps in particle_sets]))
        out = {k: cat(k, np.float64) for k in ("position", "velocity", "fluid_velocity_seen", "diameter", "volume")}
        out.update({k: cat(k, np.int32) for k in ("inactive", "tri", "lowerlayer")})
        act = out["inactive"] == 0
        out["tri"][act] = -2
        init_tri = np.full(act.size, -3, dtype=np.int32)
        if fs.trifinder is not None:
            pos, lo = out["position"], 0
            for ps in particle_sets:
                hi = lo + ps.npart
                rows = lo + np.nonzero(act[lo:hi])[0]
                lo = hi
                if not len(rows):
                    continue
                p = pos[rows]
                if np.all(p == p[0]):  # a point release: one lookup
                    init_tri[rows] = int(fs.trifinder(p[0, 0], p[0, 1]))
                else:
                    init_tri[rows] = np.asarray(fs.trifinder(p[:, 0], p[:, 1]), dtype=np.int32)
        out["init_tri"] = init_tri
        return out

    @staticmethod
    def _tables(particle_sets, params, rng_states):
        pars = (_hip.CylParams * len(params))(*[_params_struct(p, ps) for p, ps in zip(params, particle_sets)])
        st = (C.c_uint64 * len(params))(*[int(s) for s, _ in rng_states])
        inc = (C.c_uint64 * len(params))(*[int(i) for _, i in rng_states])
        return pars, st, inc

    @property
    def nscen(self):
        return len(self.scen)

    @property
    def particle_sets(self):
        return [sc.pset for sc in self.scen]

    @property
    def ids(self):
        """Stable ids of the scenarios on the device, in device order."""
        return [sc.id for sc in self.scen]

    def index(self, sid):
        for s, sc in enumerate(self.scen):
            if sc.id == sid:
                return s
        raise KeyError(sid)

    def steps_left(self, s=None):
        if s is None:
            return [sc.nsteps - sc.done for sc in self.scen]
        return self.scen[s].nsteps - self.scen[s].done

    @property
    def npart(self):
        return int(sum(sc.pset.npart for sc in self.scen))

    # -- stepping -----------------------------------------------------------------
    def run(self, nsteps=None):
        """nsteps steps of the batch (default: until every scenario is done)."""
        if nsteps is None:
            nsteps = max(self.steps_left())
        if nsteps <= 0:
            return
        nopen = self.field_set.triangular_mesh.nopen
        stats = np.zeros(max(1, nsteps * self.nscen * max(nopen, 1)), dtype=np.int32)
        t = C.c_double(self.time)
        rc = self.ctx.lib.cyl_batch_step(self.ctx.h, C.byref(t), self.dt, int(nsteps), _hip.ptr(stats, "i"))
        self.ctx.check(rc, "cyl_batch_step")
        self.time = t.value
        if nopen:
            blk = stats[:nsteps * self.nscen * nopen].reshape(nsteps, self.nscen, nopen)
        for s, sc in enumerate(self.scen):
            k = min(nsteps, max(0, sc.nsteps - sc.done))
            if nopen and k:
                sc.stats.append(blk[:k, s, :].copy())
            sc.done += k

    @property
    def bnd_stats(self):
        """Per scenario (device order): boundary counts of every step it ran, (k, nopen)."""
        nopen = max(self.field_set.triangular_mesh.nopen, 1)
        return [np.concatenate(sc.stats) if sc.stats else np.zeros((0, nopen), dtype=np.int32) for sc in self.scen]

    def rng_state(self, s):
        st, inc = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.lib.cyl_scenario_rng_get(self.ctx.h, s, C.byref(st), C.byref(inc)),
                       "cyl_scenario_rng_get")
        return st.value, inc.value

    # -- results ------------------------------------------------------------------
    def collect(self, s, allocate=True):
        """Write scenario s's rows back into its particle set (tags of deleted
        rows renumbered in deletion order, lagrangian_particle_set.pyx:786-788).
        allocate=False stages the rows in the library's static device array
        instead of device memory allocated for the call (scenarios of at most
        CYL_NOALLOC_MAX_ROWS rows)."""
        ps, dim = self.scen[s].pset, self.dim
        n = ps.npart
        h = ps._h
        for k in ("position", "velocity", "fluid_velocity_seen"):
            h[k] = np.zeros((n, dim))
        for k in ("inactive", "tri", "lowerlayer"):
            h[k] = np.zeros(n, dtype=np.int32)
        dele = np.zeros(max(n, 1), dtype=np.int32)
        cnt = C.c_int()
        P = _hip.ptr
        fn = "cyl_batch_collect" if allocate else "cyl_batch_collect_noalloc"
        self.ctx.check(getattr(self.ctx.lib, fn)(
            self.ctx.h, s, n, P(h["position"]), P(h["velocity"]), P(h["fluid_velocity_seen"]), P(h["inactive"], "i"),
            P(h["tri"], "i"), P(h["lowerlayer"], "i"), P(dele, "i"), n, C.byref(cnt)), fn)
        rows = dele[:cnt.value]
        tag = h["tag"]
        tag[rows] = np.arange(ps.maxtag + 1, ps.maxtag + 1 + len(rows), dtype=np.int32)
        ps.maxtag += len(rows)
        inact = h["inactive"] != 0
        h["diameter"][inact] = 0.
        h["volume"][inact] = 0.
        ps.npart_active = int(np.count_nonzero(~inact))
        ps._stale = False
        return ps

    def results(self):
        """Every scenario's rows written back into its particle set."""
        return [self.collect(s) for s in range(self.nscen)]

    # -- continuous batching ------------------------------------------------------
    def reshape(self, keep, particle_sets=(), parameters=(), rng_states=(), nsteps=None):
        """Keep the scenarios at device positions `keep` (their rows stay on the
        device), drop the others, and admit new scenarios after them; returns
        the new scenarios' ids."""
        keep = sorted(set(int(s) for s in keep))
        params = [Parameters(p) for p in parameters]
        psets = list(particle_sets)
        if not (len(psets) == len(params) == len(rng_states)):
            raise ValueError("one parameter dict and one (state, inc) per particle set")
        if psets:
            self._check(psets, params)
        if nsteps is None:
            nsteps = [_steps_of(SimuTime(False, self.time, p.final_time, p.time_step)) for p in params]
        new = [self._new_scenario(ps, p, k) for ps, p, k in zip(psets, params, nsteps)]
        lib, h, P = self.ctx.lib, self.ctx.h, _hip.ptr
        kp = np.ascontiguousarray(keep, dtype=np.int32)
        if psets:
            rows = self._rows(psets)
            sizes = np.ascontiguousarray([ps.npart for ps in psets], dtype=np.int32)
            pars, st, inc = self._tables(psets, params, rng_states)
            ns = np.ascontiguousarray(nsteps, dtype=np.int32)
            rc = lib.cyl_batch_reshape(h, len(keep), P(kp, "i"), len(psets), P(sizes, "i"), C.cast(pars, C.c_void_p),
                                       st, inc, P(ns, "i"), P(rows["position"]), P(rows["velocity"]),
                                       P(rows["fluid_velocity_seen"]), P(rows["inactive"], "i"), P(rows["tri"], "i"),
                                       P(rows["lowerlayer"], "i"), P(rows["diameter"]), P(rows["volume"]),
                                       P(rows["init_tri"], "i"), self.time, self.dt)
        else:
            rc = lib.cyl_batch_reshape(h, len(keep), P(kp, "i"), 0, None, None, None, None, None, None, None, None,
                                       None, None, None, None, None, None, self.time, self.dt)
        self.ctx.check(rc, "cyl_batch_reshape")
        self.scen = [self.scen[s] for s in keep] + new
        return [sc.id for sc in new]

    def admit(self, particle_sets, parameters, rng_states, nsteps=None):
        """New scenarios join the running batch; returns their ids."""
        return self.reshape(range(self.nscen), particle_sets, parameters, rng_states, nsteps)

    def retire(self, positions):
        """Scenarios at these device positions leave the device (collect them first)."""
        gone = set(int(s) for s in positions)
        self.reshape([s for s in range(self.nscen) if s not in gone])

    def close(self):
        self.ctx.close()
