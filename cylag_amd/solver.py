"""Solver: the reference's driver (cylag/core/solver.pyx:42-690) with the
per-particle loop replaced by the HIP step of libcylag_hip.so.

Solver.forward(dt) and solve() keep the reference's semantics: field time
interpolation at the start-of-step time, the pool-order loop with the model
update and boundary conditions, SimuTime.increment, and the global PCG32
stream advanced exactly as the serial loop would.  solve() runs the steps
in device-resident batches and only reads results back at the end.
"""
import ctypes as C

import numpy as np

from . import _hip, random_utils
from .control_shapes import ControlSection, section_segments
from .particles_io import (prepare_output, write_deposited_particles_txt, write_particles,
                           write_stranded_particles_txt)
from .write_statistics import prepare_secstats_output, write_secstats_buffer
from .parameters import Parameters
from .simutime import SimuTime


def _params_struct(p, pset):
    return _hip.CylParams(
        int(p.model), int(p.time_scheme), int(p.particle_velocity_init), int(p.rng_method),
        int(p.diffusion_model), int(p.diffusion_lsm2_option), float(p.horizontal_diffusivity),
        float(p.vertical_diffusivity), float(p.schmidt_number), int(bool(p.boundary_conditions)),
        int(p.boundary_conditions_type), int(p.max_wall_rebounds), int(bool(p.bnd_statistics)),
        float(p.rebound_damping_coef), float(p.wall_roughness_angle),
        float(p.stranding_probability), int(bool(pset.compute_depth)),
        int(bool(p.frozen_eulerian_fields)), int(p.lsm3_edge_opt), float(p.diffusion_lsm3_tl_horizontal),
        float(p.diffusion_lsm3_tl_vertical), int(bool(p.stranding_output)), int(bool(p.bottom_deposition)))


def _pconst_struct(ps):
    return _hip.CylPconst(float(ps.rho0), float(ps.density), float(ps.drag_coef),
                          float(ps.added_mass_coef), float(ps.buoyancy_velocity_value),
                          int(ps.drag_coef_model), int(bool(ps.added_mass_force)),
                          int(ps.buoyancy_velocity_model), float(ps.a0c), float(ps.a1c),
                          float(ps.a2c), float(ps.a3c))


def upload_mesh(ctx, m):
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    arrs = [d(m.x), d(m.y), i(m.triangles), i(m.edges), i(m.edgetri_map), i(m.triedge_map),
            i(m.edges_labels), d(m.det), d(m.gradx), d(m.grady)]
    P = _hip.ptr
    ctx.check(ctx.lib.cyl_mesh_upload(
        ctx.h, m.nv, P(arrs[0]), P(arrs[1]), m.nt, P(arrs[2], "i"), m.ne, P(arrs[3], "i"),
        P(arrs[4], "i"), P(arrs[5], "i"), P(arrs[6], "i"), P(arrs[7]), P(arrs[8]), P(arrs[9]),
        int(m.nopen)), "cyl_mesh_upload")


def upload_fields(ctx, f):
    P = _hip.ptr
    fids = np.array([f.fid_zs, f.fid_ux, f.fid_uy, f.fid_uz, f.fid_turbeng, f.fid_dissip,
                     f.fid_zb], dtype=np.int32)
    opt = f.opt_fields if f.nopt > 0 else None
    w = f.velocity_z if f.dim == 3 else None
    ctx.check(ctx.lib.cyl_fields_upload(
        ctx.h, f.dim, f.nlayers, f.npoints, f.ntimes, P(f.times), P(f.elevation),
        P(f.velocity_x), P(f.velocity_y), P(w), f.nopt, P(opt), P(fids, "i"),
        int(f.friction_model), float(f.friction_coefficient)), "cyl_fields_upload")


def check_setup(fs, ps, p):
    """The Solver constructor's sanity checks on a field set, particle set and
    parameters (solver.pyx:120-177); raises ValueError with the reference's texts."""
    if fs.dim != ps.dim:
        if fs.dim == 2 and ps.dim == 3:
            if fs.fid_zs == -1 or fs.fid_zb == -1:
                raise ValueError("2D EulerianFieldSet with 3D LagrangianParticleSet : "
                                 "EulerianFieldSet should contain free surface and bottom fields")
        else:
            raise ValueError("EulerianFieldSet dimension ({}) does not match "
                             "LagrangianParticleSet set dimension ({})".format(fs.dim, ps.dim))
    if p.frozen_eulerian_fields != 1 and p.final_time > fs.times[fs.ntimes - 1]:
        raise ValueError("final_time > EulerianFieldSet maximum time")
    if fs.ntimes == 1 and p.frozen_eulerian_fields != 1:
        raise ValueError("EulerianFieldSet has only one time step, use frozen_eulerian_fields=True")
    if ps.additional_velocity and p.model in (2, 3):
        raise ValueError("Additional velocity only compatible with LSM1")
    if ps.additional_force and p.model == 1:
        raise ValueError("Additional force not compatible with LSM1")
    if ps.additional_velocity or ps.additional_force:
        # add_velocity_i / add_force_i of particle-set subclasses (lagrangian_particle_set.pyx:892-900)
        raise ValueError("additional velocity / force fields are not supported by the device step")
    if p.diffusion_model in (2, 3) and (fs.fid_turbeng == -1 or fs.fid_dissip == -1):
        raise ValueError(
            "LSM: diffusion_model={} requires 'TURBULENT ENERG.' and 'DISSIPATION' fields in "
            "EulerianFieldSet (fid_turbeng={}, fid_dissip={})".format(
                p.diffusion_model, fs.fid_turbeng, fs.fid_dissip))


class Solver:
    def __init__(self, field_set, particle_set, parameters={}, sources=None,
                 control_sections=None, device=0):
        self.parameters = Parameters(parameters)
        self.parameters.check_values()
        p = self.parameters
        self.simutime = SimuTime(p.listing, p.initial_time, p.final_time, p.time_step)
        self.sources = sources
        if control_sections is not None:
            control_sections = [c if isinstance(c, ControlSection) else ControlSection(c)
                                for c in control_sections]
        self.control_sections = control_sections
        p.control_sections = control_sections is not None
        p.ncsec = len(control_sections) if control_sections is not None else 0
        self.particle_set = particle_set
        self.field_set = field_set
        fs, ps = field_set, particle_set
        check_setup(fs, ps, p)
        fs.frozen_hydro = bool(p.frozen_eulerian_fields == 1)
        ps.rho0 = p.water_density
        if p.model in (2, 3):
            ps._initialize_model_coefficients()
        self._par = _params_struct(p, ps)
        self.ctx = _hip.Context(device)
        upload_mesh(self.ctx, fs.triangular_mesh)
        upload_fields(self.ctx, fs)
        if control_sections:
            nseg_of, seg = section_segments(control_sections)
            self.ctx.check(self.ctx.lib.cyl_control_sections_set(self.ctx.h, len(control_sections),
                                                                 _hip.ptr(nseg_of, "i"), _hip.ptr(seg)),
                           "cyl_control_sections_set")
        self.ctx.check(self.ctx.lib.cyl_fields_reset(self.ctx.h, int(fs.frozen_hydro),
                                                     float(p.initial_time)), "cyl_fields_reset")
        if not fs.frozen_hydro:
            self.ctx.check(self.ctx.lib.cyl_fields_interpolate(self.ctx.h, float(p.initial_time)),
                           "cyl_fields_interpolate")
        # initialize_particle_state for every active row (solver.pyx:193-197)
        ps._sync()
        h = ps._h
        act = h["inactive"] == 0
        h["tri"][act] = -2
        self._n_deleted_seen = 0
        ps._solver = self
        self._push(init=True)
        self._deleted_rows_to_tags()
        if p.output_file:  # output directory, header and the initial record (solver.pyx:200-225)
            prepare_output(ps, p.output_rep, p.output_file_format, p.output_file_name, p.model)
        if p.bnd_statistics or p.control_sections:
            prepare_secstats_output(p.bnd_statistics, fs.triangular_mesh.nopen, p.output_rep,
                                    p.bnd_statistics_file_name, p.control_sections, p.ncsec,
                                    p.control_sections_file_name)
        if p.output_file:
            write_particles(ps, self.simutime.time, 0, p)
        nopen = fs.triangular_mesh.nopen
        nt_buf = self.simutime.nt
        self._bnd_stats_times = np.zeros(nt_buf)
        self._bnd_stats_data = np.zeros((nt_buf, max(1, nopen)), dtype=np.int32)
        self._bnd_stats_count = 0
        self._bnd_flushed = []
        self._deposited_flushed = []
        self._csec_stats_times = np.zeros(nt_buf)
        self._csec_stats_data = np.zeros((nt_buf, max(1, p.ncsec)), dtype=np.int32)
        self._csec_stats_count = 0
        self._csec_flushed = []

    # -- host <-> device ----------------------------------------------------------
    def _init_tri(self, rows):
        """Elements of new rows from the field set's trifinder (solver.pyx:262-268);
        -3 asks the device for the reference's KD-tree search."""
        ps, fs = self.particle_set, self.field_set
        it = np.full(ps.npart, -3, dtype=np.int32)
        if fs.trifinder is not None and len(rows):
            pos = ps._h["position"]
            it[rows] = np.asarray(fs.trifinder(pos[rows, 0], pos[rows, 1]), dtype=np.int32)
        return it

    def _push(self, init=False):
        """Upload the host rows.  init: initialize the tri == -2 rows now (the
        constructor, initialize_particle_state); otherwise the next step
        initializes each of them right before updating it, in pool order, as
        forward() does for arrivals (solver.pyx:458-467)."""
        ps = self.particle_set
        h = ps._h
        P = _hip.ptr
        new = np.nonzero((h["inactive"] == 0) & (h["tri"] == -2))[0]
        it = self._init_tri(new)
        d = lambda k: np.ascontiguousarray(h[k], dtype=np.float64)
        i = lambda k: np.ascontiguousarray(h[k], dtype=np.int32)
        pc = _pconst_struct(ps)
        arr = [d("position"), d("velocity"), d("fluid_velocity_seen"), i("inactive"), i("tri"),
               i("lowerlayer"), d("diameter"), d("volume")]
        self.ctx.check(self.ctx.lib.cyl_particles_upload(
            self.ctx.h, ps.npart, ps.dim, P(arr[0]), P(arr[1]), P(arr[2]), P(arr[3], "i"),
            P(arr[4], "i"), P(arr[5], "i"), P(arr[6]), P(arr[7]), P(it, "i"), C.byref(pc)),
            "cyl_particles_upload")
        self._n_deleted_seen = 0
        ps._host_dirty = False
        if init and len(new):
            self._rng_to_device()
            self.ctx.check(self.ctx.lib.cyl_init_particles(
                self.ctx.h, C.byref(self._par), float(self.simutime.time),
                float(self.simutime.time_step)), "cyl_init_particles")
            self._rng_from_device()
            ps._stale = True

    def _rng_to_device(self):
        st, inc = random_utils.get_state()
        self.ctx.check(self.ctx.lib.cyl_rng_set(self.ctx.h, st, inc), "cyl_rng_set")

    def _rng_from_device(self):
        st, inc = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.ctx.lib.cyl_rng_get(self.ctx.h, C.byref(st), C.byref(inc)), "cyl_rng_get")
        random_utils.set_state(st.value, inc.value)

    def _pull(self):
        """Download the device rows into the particle set's host arrays."""
        ps = self.particle_set
        h, n, dim = ps._h, ps.npart, ps.dim
        P = _hip.ptr
        for k in ("position", "velocity", "fluid_velocity_seen", "last_position", "last_velocity"):
            h[k] = np.zeros((n, dim))
        for k in ("inactive", "tri", "lowerlayer", "last_tri"):
            h[k] = np.zeros(n, dtype=np.int32)
        depth = np.zeros(n) if ps.compute_depth else None
        self.ctx.check(self.ctx.lib.cyl_particles_download(
            self.ctx.h, n, P(h["position"]), P(h["velocity"]), P(h["fluid_velocity_seen"]),
            P(h["inactive"], "i"), P(h["tri"], "i"), P(h["lowerlayer"], "i"),
            P(h["last_position"]), P(h["last_velocity"]), P(h["last_tri"], "i"), P(depth)),
            "cyl_particles_download")
        if depth is not None:
            h["depth"] = depth
        self._deleted_rows_to_tags()
        inact = h["inactive"] != 0
        h["diameter"][inact] = 0.
        h["volume"][inact] = 0.
        ps.npart_active = int(np.count_nonzero(~inact))

    def _deleted_rows_to_tags(self):
        """Rows deleted on the device get new tags in deletion order
        (lagrangian_particle_set.pyx:786-788)."""
        n = C.c_longlong()
        self.ctx.check(self.ctx.lib.cyl_deleted_count(self.ctx.h, C.byref(n)), "cyl_deleted_count")
        if n.value == self._n_deleted_seen:
            return
        ps = self.particle_set
        buf = np.zeros(ps.npart, dtype=np.int32)
        cnt = C.c_int()
        self.ctx.check(self.ctx.lib.cyl_particles_deleted(self.ctx.h, _hip.ptr(buf, "i"), ps.npart,
                                                          C.byref(cnt)), "cyl_particles_deleted")
        rows = buf[:cnt.value]
        tag = ps._h["tag"]
        tag[rows] = np.arange(ps.maxtag + 1, ps.maxtag + 1 + len(rows), dtype=np.int32)
        ps.maxtag += len(rows)
        self._n_deleted_seen = n.value

    # -- stepping -------------------------------------------------------------------
    def initialize_particle_state(self, i):
        """(Re)initialize row i: localization and velocity (solver.pyx:238-380)."""
        ps = self.particle_set
        ps._sync()
        ps._h["tri"][i] = -2
        self._push(init=True)

    def _run(self, nsteps):
        ps, st, p = self.particle_set, self.simutime, self.parameters
        if ps._host_dirty:
            ps._sync()
            self._push()
        nopen = self.field_set.triangular_mesh.nopen
        stats = np.zeros(max(1, nsteps * max(nopen, 1)), dtype=np.int32)
        t = C.c_double(st.time)
        t0 = st.time
        self._rng_to_device()
        rc = self.ctx.lib.cyl_step(self.ctx.h, C.byref(self._par), C.byref(t), float(st.time_step),
                                   int(nsteps), _hip.ptr(stats, "i"))
        self._rng_from_device()
        ps._stale = True
        self.ctx.check(rc, "cyl_step")
        times = []
        for _ in range(nsteps):
            times.append(st.time)
            st.increment()
        st.time = t.value if nsteps else st.time
        if p.bnd_statistics:
            for s in range(nsteps):
                self._bnd_stats_times[self._bnd_stats_count] = times[s]
                self._bnd_stats_data[self._bnd_stats_count, :nopen] = stats[s * nopen:(s + 1) * nopen]
                self._bnd_stats_count += 1
        if p.stranding_output and p.stranding_probability > 0.:
            self._collect_stranded()
        if p.bottom_deposition and self.particle_set.dim == 3:
            self._collect_deposited(st.iteration - nsteps + 1)
        if self.control_sections:
            nc = p.ncsec
            counts = np.zeros(max(1, nsteps * nc), dtype=np.int32)
            nout = C.c_int()
            self.ctx.check(self.ctx.lib.cyl_control_counts(self.ctx.h, _hip.ptr(counts, "i"), counts.size,
                                                           C.byref(nout)), "cyl_control_counts")
            for s in range(nsteps):
                self._csec_stats_times[self._csec_stats_count] = times[s]
                self._csec_stats_data[self._csec_stats_count, :nc] = counts[s * nc:(s + 1) * nc]
                self._csec_stats_count += 1
        if st.print_progress:
            for it in range(st.iteration - nsteps + 1, st.iteration + 1):
                if p.listing_printout_period and it % p.listing_printout_period == 0:
                    st.progress()
        return t0

    def _collect_stranded(self):
        """Append the positions of the rows stranded by the last call to the
        particle set's buffer, in the reference's order (compute_bc.pyx:107-122)."""
        ps = self.particle_set
        n = C.c_int()
        self.ctx.check(self.ctx.lib.cyl_stranded_get(self.ctx.h, None, 0, C.byref(n)), "cyl_stranded_get")
        if n.value == 0:
            return
        xyz = np.zeros((n.value, 3))
        self.ctx.check(self.ctx.lib.cyl_stranded_get(self.ctx.h, _hip.ptr(xyz), n.value, C.byref(n)),
                       "cyl_stranded_get")
        for r in xyz:
            if ps.stranded_count < ps.max_stranded:
                ps.stranded_positions[ps.stranded_count] = r
                ps.stranded_count += 1
            else:
                print("Warning: stranded particle buffer is full. Increase max_stranded in LagrangianParticleSet")

    def _collect_deposited(self, first_iteration):
        """Append the rows deposited on the bed by the last call (bottom_deposition,
        an extension) to the particle set's buffer: position on the bed and the
        iteration of the step, in the serial order of the loop."""
        ps = self.particle_set
        n = C.c_int()
        self.ctx.check(self.ctx.lib.cyl_deposited_get(self.ctx.h, None, None, 0, C.byref(n)), "cyl_deposited_get")
        if n.value == 0:
            return
        xyz = np.zeros((n.value, 3))
        step = np.zeros(n.value, dtype=np.int32)
        self.ctx.check(self.ctx.lib.cyl_deposited_get(self.ctx.h, _hip.ptr(xyz), _hip.ptr(step, "i"), n.value,
                                                      C.byref(n)), "cyl_deposited_get")
        for r, s in zip(xyz, step):
            if ps.deposited_count < ps.max_deposited:
                ps.deposited_positions[ps.deposited_count] = r
                ps.deposited_iterations[ps.deposited_count] = first_iteration + s
                ps.deposited_count += 1
            else:
                print("Warning: deposited particle buffer is full. Increase max_deposited in LagrangianParticleSet")

    def forward(self, time_step=None):
        """One time step (solver.pyx:413-585); time_step is unused, as in the reference."""
        deferred = self._release_sources()
        self._run(1)
        self._activate(deferred)
        self._output_if_due()

    def _activate(self, deferred):
        """Arrivals that grew the pool join from the next step on."""
        if deferred is None:
            return
        rows, keep = deferred
        ps = self.particle_set
        ps._sync()
        h = ps._h
        for k, v in keep.items():
            h[k][rows] = v
        h["inactive"][rows] = 0
        h["tri"][rows] = -2
        ps.npart_active += len(rows)
        ps._host_dirty = True

    def _output_if_due(self):
        """Particle output every output_printout_period iterations (solver.pyx:556-565)."""
        p, st = self.parameters, self.simutime
        if p.output_file and st.iteration % p.output_printout_period == 0:
            write_particles(self.particle_set, st.time, st.iteration, p)

    def _release_sources(self):
        """Release the sources due at this step.  forward() loops over the rows
        the pool had when the step began (solver.pyx:432,458), so arrivals that
        grow the pool are neither initialized nor stepped before the next step:
        their rows are returned and kept inactive for this one."""
        if self.sources is None:
            return None
        ps = self.particle_set
        n0 = ps.npart
        fs = self.field_set
        if fs.ntimes > 1:  # current time record, for sources that read the fields
            t, rec = fs.times, 0
            for i in range(fs.ntimes - 1):
                if t[i] <= self.simutime.time < t[i + 1]:
                    rec = i
                    break
            else:
                rec = fs.ntimes - 2 if self.simutime.time >= t[-1] else 0
            fs.record = rec
        for isrc, source in enumerate(self.sources, start=1):
            if source.release_condition(self.simutime, self.field_set):
                new = source.init_positions(self.field_set, self.particle_set.dim)
                self.particle_set.add(new)
                if self.parameters.listing:
                    print(' ~~> Adding {} particles from source {}'.format(np.shape(new)[0], isrc))
        if ps.npart > n0:
            h = ps._h
            rows = n0 + np.nonzero(h["inactive"][n0:] == 0)[0]
            keep = {k: h[k][rows].copy() for k in ("position", "velocity", "fluid_velocity_seen", "diameter",
                                                   "volume")}
            h["inactive"][rows] = 1
            ps.npart_active -= len(rows)
            ps._host_dirty = True
            return rows, keep
        return None

    def _steps_left(self):
        """Number of forward() calls solve() makes: its loop runs nt times and
        skips once SimuTime is finished (solver.pyx:399-401, simutime.pyx:84),
        which float accumulation of the time can reach early."""
        st = self.simutime
        t, it, fin, n = st.time, st.iteration, st.is_finished, 0
        for _ in range(st.nt):
            if fin:
                break
            t += st.time_step
            it += 1
            n += 1
            fin = t >= st.final_time or it >= st.maxiteration
        return n

    def solve(self, batch=None):
        """All remaining steps (solver.pyx:382-411), in device-resident batches."""
        st = self.simutime
        remaining = self._steps_left()
        p = self.parameters
        while remaining > 0:
            deferred = None
            if self.sources is not None:
                deferred = self._release_sources()
                n = 1
            else:
                n = remaining if batch is None else min(batch, remaining)
            if p.output_file:  # stop at the next output iteration
                n = min(n, p.output_printout_period - st.iteration % p.output_printout_period)
            self._run(n)
            self._activate(deferred)
            self._output_if_due()
            remaining -= n
        self.flush_statistics()

    def flush_statistics(self):
        """Write the buffered per-step statistics to their files and reset the
        buffers (solver.pyx:587-630); the flushed rows stay readable through
        bnd_stats and csec_stats."""
        p = self.parameters
        n = self._bnd_stats_count
        if n:
            if p.bnd_statistics:
                write_secstats_buffer(self._bnd_stats_times, self._bnd_stats_data, n,
                                      self.field_set.triangular_mesh.nopen, p.output_rep,
                                      p.bnd_statistics_file_name)
            self._bnd_flushed.append((self._bnd_stats_times[:n].copy(),
                                      self._bnd_stats_data[:n].copy()))
        self._bnd_stats_count = 0
        if p.stranding_output:
            ps = self.particle_set
            write_stranded_particles_txt(p.output_rep, p.stranding_output_file_name, ps.stranded_positions,
                                         ps.stranded_count)
            ps.stranded_count = 0
        if p.bottom_deposition:
            ps = self.particle_set
            n = ps.deposited_count
            write_deposited_particles_txt(p.output_rep, p.deposition_output_file_name, ps.deposited_positions,
                                          ps.deposited_iterations, n)
            self._deposited_flushed.append((ps.deposited_positions[:n].copy(), ps.deposited_iterations[:n].copy()))
            ps.deposited_count = 0
        n = self._csec_stats_count
        if n:
            write_secstats_buffer(self._csec_stats_times, self._csec_stats_data, n, p.ncsec, p.output_rep,
                                  p.control_sections_file_name)
            self._csec_flushed.append((self._csec_stats_times[:n].copy(),
                                       self._csec_stats_data[:n].copy()))
        self._csec_stats_count = 0

    @property
    def csec_stats(self):
        """(times, counts[nstep, ncsec]) of every step run so far."""
        parts = self._csec_flushed + [(self._csec_stats_times[:self._csec_stats_count],
                                       self._csec_stats_data[:self._csec_stats_count])]
        return (np.concatenate([t for t, _ in parts]),
                np.concatenate([d for _, d in parts])[:, :max(self.parameters.ncsec, 1)])

    @property
    def deposited(self):
        """(positions[n, 3], iterations[n]) of every deposit on the bed so far
        (bottom_deposition, an extension), flushed or not, in the loop's order."""
        ps = self.particle_set
        parts = self._deposited_flushed + [(ps.deposited_positions[:ps.deposited_count],
                                            ps.deposited_iterations[:ps.deposited_count])]
        return np.concatenate([q[0] for q in parts]), np.concatenate([q[1] for q in parts])

    @property
    def bnd_stats(self):
        """(times, counts[nstep, nopen]) of every step run so far."""
        parts = self._bnd_flushed + [(self._bnd_stats_times[:self._bnd_stats_count],
                                      self._bnd_stats_data[:self._bnd_stats_count])]
        nopen = self.field_set.triangular_mesh.nopen
        return (np.concatenate([t for t, _ in parts]),
                np.concatenate([d for _, d in parts])[:, :max(nopen, 1)])
