"""BatchSolver: many release scenarios on one mesh and field set, advanced
together by one device context.

A scenario is a LagrangianParticleSet with its own parameters (LSM1 or LSM2,
diffusion, boundary conditions...), its own random stream (state, inc) and its
own number of steps.  The scenarios' pools are concatenated into one device
pool; every step advances all scenarios that have steps left in a few kernel
launches (one per kernel class: the fused 2D LSM-1 kernel, the generic kernel
per model), and the element re-sort, the event tables and the normal generator
are shared.  Each scenario's rows, random stream, deleted-row tags and
boundary statistics are the ones a plain Solver run of that scenario alone
gives (the reference's nearest analogue is SolverParallel's split of one
release into independently seeded ranks, solver_parallel.pyx:99-130).

All scenarios share the time step, the initial time, the frozen-fields flag,
the particle constants (density, drag, buoyancy...) and one rng_method.
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


class BatchSolver:
    def __init__(self, field_set, particle_sets, parameters, rng_states, nsteps=None, device=0,
                 sort_interval=8):
        if not (len(particle_sets) == len(parameters) == len(rng_states)) or not particle_sets:
            raise ValueError("one parameter dict and one (state, inc) per particle set")
        self.field_set = field_set
        self.particle_sets = list(particle_sets)
        self.parameters = [Parameters(p) for p in parameters]
        for p in self.parameters:
            p.check_values()
            if p.model == 3:
                raise ValueError("LSM3 (model=3) is not available in the MI355X step")
        p0 = self.parameters[0]
        keys = ("time_step", "initial_time", "frozen_eulerian_fields", "rng_method")
        for p in self.parameters[1:]:
            for k in keys:
                if getattr(p, k) != getattr(p0, k):
                    raise ValueError("the scenarios of a batch share {}".format(k))
        for ps, p in zip(self.particle_sets, self.parameters):
            check_setup(field_set, ps, p)
        if len({ps.dim for ps in self.particle_sets}) != 1:
            raise ValueError("the particle sets of a batch share one dimension")
        for ps, p in zip(self.particle_sets, self.parameters):
            ps.rho0 = p.water_density
            if p.model == 2:
                ps._initialize_model_coefficients()
        # physical particle constants must agree; the LSM-2 coefficients come
        # from an LSM-2 scenario (LSM-1 does not read them)
        phys = [f for f, _ in _hip.CylPconst._fields_[:8]]
        pcs = [tuple(getattr(_pconst_struct(ps), f) for f in phys) for ps in self.particle_sets]
        if any(pc != pcs[0] for pc in pcs[1:]):
            raise ValueError("the particle sets of a batch share their particle constants")
        self._pc_set = next((ps for ps, p in zip(self.particle_sets, self.parameters) if p.model == 2),
                            self.particle_sets[0])
        self.simutimes = [SimuTime(False, p.initial_time, p.final_time, p.time_step) for p in self.parameters]
        self.nsteps = list(nsteps) if nsteps is not None else [_steps_of(st) for st in self.simutimes]
        self.time = float(p0.initial_time)
        self.dt = float(p0.time_step)
        self.steps_done = 0
        fs = field_set
        fs.frozen_hydro = bool(p0.frozen_eulerian_fields == 1)
        self.ctx = _hip.Context(device)
        lib, h = self.ctx.lib, self.ctx.h
        upload_mesh(self.ctx, fs.triangular_mesh)
        upload_fields(self.ctx, fs)
        self.ctx.check(lib.cyl_fields_reset(h, int(fs.frozen_hydro), self.time), "cyl_fields_reset")
        if not fs.frozen_hydro:
            self.ctx.check(lib.cyl_fields_interpolate(h, self.time), "cyl_fields_interpolate")
        # one pool: the scenarios' pools in order
        sizes = [ps.npart for ps in self.particle_sets]
        self.base = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
        n, dim = int(self.base[-1]), self.particle_sets[0].dim
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([np.asarray(ps._h[k], dtype=dt)
                                                                 for ps in self.particle_sets]))
        for ps in self.particle_sets:
            ps._sync()
        pos, vel, us = cat("position", np.float64), cat("velocity", np.float64), cat("fluid_velocity_seen", np.float64)
        inactive, tri, ll = cat("inactive", np.int32), cat("tri", np.int32), cat("lowerlayer", np.int32)
        diam, vol = cat("diameter", np.float64), cat("volume", np.float64)
        act = inactive == 0
        tri[act] = -2  # initialize_particle_state for every active row (solver.pyx:193-197)
        init_tri = np.full(n, -3, dtype=np.int32)
        if fs.trifinder is not None:
            for s in range(len(sizes)):
                lo, hi = int(self.base[s]), int(self.base[s + 1])
                rows = lo + np.nonzero(act[lo:hi])[0]
                if not len(rows):
                    continue
                p = pos[rows]
                if np.all(p == p[0]):  # a point release: one lookup
                    init_tri[rows] = int(fs.trifinder(p[0, 0], p[0, 1]))
                else:
                    init_tri[rows] = np.asarray(fs.trifinder(p[:, 0], p[:, 1]), dtype=np.int32)
        P = _hip.ptr
        pc = _pconst_struct(self._pc_set)
        self.ctx.check(lib.cyl_particles_upload(h, n, dim, P(pos), P(vel), P(us), P(inactive, "i"), P(tri, "i"),
                                                P(ll, "i"), P(diam), P(vol), P(init_tri, "i"), C.byref(pc)),
                       "cyl_particles_upload")
        pars = (_hip.CylParams * len(sizes))(*[_params_struct(p, ps)
                                                 for p, ps in zip(self.parameters, self.particle_sets)])
        st = (C.c_uint64 * len(sizes))(*[int(s) for s, _ in rng_states])
        inc = (C.c_uint64 * len(sizes))(*[int(i) for _, i in rng_states])
        ns = np.ascontiguousarray(self.nsteps, dtype=np.int32)
        self.ctx.check(lib.cyl_scenarios_set(h, len(sizes), P(self.base, "i"), C.cast(pars, C.c_void_p), st, inc,
                                             P(ns, "i")),
                       "cyl_scenarios_set")
        lib.cyl_set_tuning(h, int(sort_interval), 256)
        self.ctx.check(lib.cyl_init_particles(h, C.byref(pars[0]), self.time, self.dt), "cyl_init_particles")
        self._maxtag0 = [ps.maxtag for ps in self.particle_sets]
        self.bnd_stats = []

    @property
    def nscen(self):
        return len(self.particle_sets)

    def run(self, nsteps=None):
        """nsteps steps of the batch (default: until every scenario is done)."""
        if nsteps is None:
            nsteps = max(self.nsteps) - self.steps_done
        if nsteps <= 0:
            return
        nopen = self.field_set.triangular_mesh.nopen
        stats = np.zeros(max(1, nsteps * self.nscen * max(nopen, 1)), dtype=np.int32)
        t = C.c_double(self.time)
        rc = self.ctx.lib.cyl_batch_step(self.ctx.h, C.byref(t), self.dt, int(nsteps), _hip.ptr(stats, "i"))
        self.ctx.check(rc, "cyl_batch_step")
        self.time = t.value
        self.steps_done += nsteps
        if nopen:
            self.bnd_stats.append(stats[:nsteps * self.nscen * nopen].reshape(nsteps, self.nscen, nopen).copy())

    def rng_state(self, s):
        st, inc = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.lib.cyl_scenario_rng_get(self.ctx.h, s, C.byref(st), C.byref(inc)),
                       "cyl_scenario_rng_get")
        return st.value, inc.value

    def results(self):
        """Download the pool and write every scenario's rows back into its
        particle set (tags of deleted rows renumbered in deletion order)."""
        lib, h, P = self.ctx.lib, self.ctx.h, _hip.ptr
        n, dim = int(self.base[-1]), self.particle_sets[0].dim
        out = {k: np.zeros((n, dim)) for k in ("position", "velocity", "fluid_velocity_seen", "last_position",
                                                  "last_velocity")}
        out.update({k: np.zeros(n, dtype=np.int32) for k in ("inactive", "tri", "lowerlayer", "last_tri")})
        self.ctx.check(lib.cyl_particles_download(
            h, n, P(out["position"]), P(out["velocity"]), P(out["fluid_velocity_seen"]), P(out["inactive"], "i"),
            P(out["tri"], "i"), P(out["lowerlayer"], "i"), P(out["last_position"]), P(out["last_velocity"]),
            P(out["last_tri"], "i"), P(None)), "cyl_particles_download")
        dele = np.zeros(n, dtype=np.int32)
        cnt = C.c_int()
        self.ctx.check(lib.cyl_particles_deleted(h, P(dele, "i"), n, C.byref(cnt)), "cyl_particles_deleted")
        dele = dele[:cnt.value]
        for s, ps in enumerate(self.particle_sets):
            lo, hi = int(self.base[s]), int(self.base[s + 1])
            for k, v in out.items():
                ps._h[k] = v[lo:hi].copy()
            rows = dele[(dele >= lo) & (dele < hi)] - lo
            tag = ps._h["tag"]
            tag[rows] = np.arange(ps.maxtag + 1, ps.maxtag + 1 + len(rows), dtype=np.int32)
            ps.maxtag += len(rows)
            inact = ps._h["inactive"] != 0
            ps._h["diameter"][inact] = 0.
            ps._h["volume"][inact] = 0.
            ps.npart_active = int(np.count_nonzero(~inact))
            ps._stale = False
        return self.particle_sets
