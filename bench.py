"""Throughput of the plain particle time step on one MI355X.

Workload (2D LSM1 on an unstructured mesh): Delaunay triangulation of 1M
jittered points in a 20 x 5 km channel (~2M elements, seed 7), two velocity
frames (t = 0 and 10000 s) of a cellular vortex (amplitude 1 m/s, wavelength
2 km) interpolated in time, depth 2-20 m, k-epsilon diffusivity K(x, y) in
[0.1, 5] m2/s with its gradient drift, 10M FP64 particles, dt = 5 s, 2000
steps in all (warm-up included), default Ziggurat normals, schmidt number
0.72.  A "step" is one Solver.forward over the whole particle set.  Inputs
are in HBM before the timed region, which brackets K steps with device syncs.

Prints one JSON line with the roofline of the fused step kernel (its launch
durations measured with HIP events on the context's stream) and a CPU
baseline (the oracle, one core, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tests.cases import c2_workload, c3_release, c3_workload  # noqa: E402  (workload builders shared with the tests)

HBM_PEAK = 8.0e12
METRIC = "particle-steps/s plain vs batched (1 GPU); p50/p99 scenario latency at fixed QPS"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_lsm1_2d_unstructured_keps_10M_counters.md")
TRAFFIC_FILE_C3 = os.path.join(ROOT, "profiles", "k_fast_3d_channel_settling_5M_counters.md")


def kernel_source_digest():
    """sha256 (16 hex digits) of the device sources the library is built from
    (cylag_amd/csrc/*.h, *.hip, in name order)."""
    import glob
    import hashlib
    h = hashlib.sha256()
    for f in sorted(glob.glob(os.path.join(ROOT, "cylag_amd", "csrc", "*.h")) +
                    glob.glob(os.path.join(ROOT, "cylag_amd", "csrc", "*.hip"))):
        with open(f, "rb") as fh:
            h.update(os.path.basename(f).encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def measured_traffic(path=TRAFFIC_FILE):
    """HBM bytes per launch of the fused step kernel from a committed counter
    summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration),
    with the source digest the summary was measured at; the figure is None
    when the summary is absent or was measured on other kernel sources."""
    fields = {}
    try:
        with open(path) as f:
            for line in f:
                if ":" in line and line.split(":", 1)[0] in ("traffic_bytes_per_launch", "kernel_source_sha256"):
                    k, v = line.split(":", 1)
                    fields[k] = v.strip()
    except OSError:
        return None, None
    digest = fields.get("kernel_source_sha256")
    if "traffic_bytes_per_launch" not in fields or digest != kernel_source_digest():
        return None, digest
    return float(fields["traffic_bytes_per_launch"]), digest


def calibrate_copy(nbytes=1 << 30):
    """A plain device-to-device copy of known size (the HIP runtime's blit
    kernel), profiled in the same run as the step kernel to scale the
    FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("calibration copy: 3 x {} bytes".format(nbytes))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_particle_step(nt, nv, npart):
    """Algorithmic bytes per particle-step: the evolving state read and written
    once (x, y f64 + element id i32 = 20 B each way), plus the mesh and node
    fields a step must read once, shared by the N particles (element record
    84 B: P1 coefficients 72 B + 3 node ids; node: x, y and two frames of u,
    v, k, eps)."""
    state = 2 * (16 + 4)
    mesh = nt * 84 + nv * (16 + 2 * 2 * 16)
    return state + mesh / npart, state, mesh


def cpu_threads():
    """Threads the CPU baseline may use: this process's CPU affinity set,
    capped by OMP_NUM_THREADS when set (a GPU box grants each GPU a share of
    its host's cores and sets that variable to the share)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(x, y, tri, fields, pos, params, nsample=100_000, budget_s=15., nthreads=None):
    """The oracle (the serial C restatement of the reference's step) timed on
    this host, on bounded samples of the same workload:

    * one core: the first `nsample` particles, one pool-order stream, stepped
      until ~budget_s of wall time (the reference's serial Solver.solve loop,
      solver.pyx:459);
    * all granted cores: thread r steps its own `nsample` particles (rows
      r*nsample...) from the stream advanced by r << 48, with its own field set
      state, as SolverParallel splits a release over ranks
      (solver_parallel.pyx:113-115,127-179); every thread runs as many steps as
      the one-core run did, and the clock spans them all.

    The releases are located by the reference's KD-tree search (no trifinder,
    as in the GPU legs)."""
    import threading
    from oracle import refpy as R
    t0 = time.perf_counter()
    mesh = R.RefMesh(x, y, tri)
    locate = mesh.kdtree_locator()

    def solver(rank):
        rows = pos[rank * nsample:(rank + 1) * nsample].copy()
        rng = R.RefRng()
        rng.advance(rank << 48)
        return R.RefSolver(R.RefFields(mesh, **fields), R.RefParticles(rows, dim=2), params, trifinder=locate,
                           rng=rng)

    s = solver(0)
    log("cpu baseline: oracle ready ({:.1f} s)".format(time.perf_counter() - t0))
    t0 = time.perf_counter()
    nsteps, shown = 0, 0.
    while True:
        s.forward(1)
        nsteps += 1
        el = time.perf_counter() - t0
        if el - shown >= 3.:
            shown = el
            log("cpu baseline, 1 core: {} steps, {:.1f} s".format(nsteps, el))
        if el >= budget_s:
            break
    one = dict(value=nsample * nsteps / el, unit="particle-steps/s", cores=1,
               sample="{} particles x {} steps of the same workload ({:.1f} s, serial pool-order loop)".format(
                   nsample, nsteps, el))
    nth = cpu_threads() if nthreads is None else nthreads
    nth = min(nth, pos.shape[0] // nsample)
    out = dict(one, kind="port", cpu_model=cpu_model(), host_cpus=os.cpu_count())
    if nth <= 1:
        out["single_core"] = dict(one)
        return out
    t0 = time.perf_counter()
    solvers = [solver(r) for r in range(nth)]
    log("cpu baseline: {} oracle solvers ready ({:.1f} s)".format(nth, time.perf_counter() - t0))
    start = threading.Barrier(nth + 1)
    errors = []

    def work(sv):
        start.wait()
        try:
            sv.forward(nsteps)
        except Exception as e:  # reported after the join
            errors.append(e)

    th = [threading.Thread(target=work, args=(sv,)) for sv in solvers]
    for t in th:
        t.start()
    start.wait()
    t0 = time.perf_counter()
    for t in th:
        t.join()
    el = time.perf_counter() - t0
    if errors:
        raise errors[0]
    log("cpu baseline, {} threads: {} x {} particles x {} steps in {:.1f} s".format(nth, nth, nsample, nsteps, el))
    out.update(value=nth * nsample * nsteps / el, cores=nth,
               sample="{} threads x {} particles x {} steps of the same workload ({:.1f} s; thread r on rows "
                      "r*{}.. with the stream advanced by r<<48)".format(nth, nsample, nsteps, el, nsample),
               single_core=one)
    return out


def c3_bytes_per_particle_step(nt, nv, nl, npart):
    """Algorithmic bytes per particle-step of the 3D LSM-2 step: the evolving
    state read and written once (x, y, z and the particle velocity f64, element
    and layer ids i32: 2 x 56 B) and the diameter read (8 B), plus the mesh and
    the one (frozen) frame of the node fields read once per step, shared by the
    N particles (element record 84 B; 2D node x, y 16 B; per node and plane
    elevation, u, v, w, k, eps 48 B)."""
    state = 2 * (6 * 8 + 2 * 4) + 8
    mesh = nt * 84 + nv * 16 + nv * nl * 6 * 8
    return state + mesh / npart, state, mesh


def c3_leg(cy, steps=100, warmup=10, npart=5_000_000, sort_interval=8):
    """3D LSM-2 Langevin with settling: the 3D channel (500k elements x 10
    sigma planes), 5M sediment particles (density 2650, settling diameters),
    direct integrator, k-epsilon K(z) with its gradient drift, dt = 2 s,
    frozen fields; `steps` steps timed after `warmup`."""
    t0 = time.time()
    x, y, tri, h, fields = c3_workload()
    pos, d = c3_release(x, y, h, npart)
    mesh = cy.TriangularMesh(x, y, tri)
    fs = cy.EulerianFieldSet(mesh, **fields)
    ps = cy.LagrangianParticleSet(pos, dim=3, particle_density=2650., buoyancy_velocity_model=1)
    ps.diameter = d
    ps.volume = np.pi * d ** 3 / 6.
    params = dict(final_time=6000., time_step=2., frozen_eulerian_fields=True, model=2, time_scheme=5,
                  diffusion_lsm2_option=1, diffusion_model=3, particle_velocity_init=1)
    cy.pcg32_seed(42, 0xda3e06d11bb2b203)
    solver = cy.Solver(fs, ps, params)
    solver.ctx.lib.cyl_set_tuning(solver.ctx.h, sort_interval, 0, 0, 0)
    log("c3: {} elements x {} planes, {} particles, ready ({:.1f} s)".format(tri.shape[0], fields["nlayers"], npart,
                                                                            time.time() - t0))
    solver._run(warmup)
    t0 = time.perf_counter()
    solver._run(steps)
    wall = time.perf_counter() - t0
    k_ms, _, nlaunch = solver.ctx.last_timing()
    prep_ms = solver.ctx.last_prep_ms()
    fused = solver.ctx.last_path()[0]
    ps._sync()
    nact = ps.npart_active
    solver.ctx.close()
    b_alg, _, _ = c3_bytes_per_particle_step(mesh.nt, mesh.nv, fields["nlayers"], nact)
    traffic, digest = measured_traffic(TRAFFIC_FILE_C3)
    avg_ms = k_ms / max(nlaunch, 1)
    achieved = b_alg * nact / (avg_ms * 1e-3)
    log("c3: {} steps in {:.3f} s, fused kernel {:.3f} ms/launch ({} fused launches), draw starts {:.3f} ms".format(
        steps, wall, avg_ms, fused, prep_ms / max(nlaunch, 1)))
    return dict(value=nact * steps / wall, unit="particle-steps/s", ms_per_step=wall * 1e3 / steps, steps=steps,
                warmup=warmup, particles=int(nact),
                workload="3D LSM-2 direct integrator with settling (lognormal settling velocities, density 2650), "
                         "k-eps K(z) with drift, sigma-layer channel {} elements x {} planes, {} particles, dt=2 s, "
                         "frozen".format(mesh.nt, fields["nlayers"], nact),
                roofline={"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                          "frac": achieved / HBM_PEAK, "traffic": traffic,
                          "traffic_source": "profiles/k_fast_3d_channel_settling_5M_counters.md",
                          "traffic_kernel_sources": digest, "kernel": "k_fast_3d<2,2,true>", "kernel_avg_ms": avg_ms,
                          "bytes_per_particle_step": b_alg, "draw_start_avg_ms": prep_ms / max(nlaunch, 1)})


def c4_scenarios(nscen=256, seed=11, lx=20000., ly=5000.):
    """Release scenarios on the unstructured channel: particle counts
    log-uniform in [1e3, 1e6], 500-5000 steps, model 1 or 2, a random point
    release in the interior."""
    rng = np.random.default_rng(seed)
    npart = np.exp(rng.uniform(np.log(1e3), np.log(1e6), nscen)).astype(np.int64)
    nsteps = rng.integers(500, 5001, nscen)
    model = rng.integers(1, 3, nscen)
    xy = np.stack([rng.uniform(0.05 * lx, 0.95 * lx, nscen), rng.uniform(0.05 * ly, 0.95 * ly, nscen)], 1)
    return npart, nsteps, model, xy


def c4_params(model):
    """LSM1 as the plain workload (k-eps K with drift), or 2D LSM2 with the
    exact Ornstein-Uhlenbeck update of the particle velocity (time_scheme 5,
    diffusion_lsm2_option 1); dt = 2 s so that 5000 steps stay within the two
    field frames (t = 0 and 10000 s)."""
    p = dict(final_time=10000., time_step=2., diffusion_model=2, schmidt_number=0.72)
    if model == 1:
        p.update(model=1)
    else:
        p.update(model=2, time_scheme=5, diffusion_lsm2_option=1)
    return p


def c4_square_scenarios(nscen=256, seed=11, length=100.):
    """The same scenario set (counts, steps, models from the same draws) as
    point releases on the config-1 square (101 x 101 nodes, 20,000 elements)."""
    npart, nsteps, model, _ = c4_scenarios(nscen, seed)
    rng = np.random.default_rng(seed + 1)
    xy = rng.uniform(0.05 * length, 0.95 * length, (nscen, 2))
    return npart, nsteps, model, xy


def c4_square_params(model):
    """The config-1 parameters (uniform flow 0.1 m/s, constant K = 0.01 m2/s,
    dt = 1 s, frozen fields, walls), LSM1, or 2D LSM2 with the direct
    integrator (time_scheme 5, diffusion_lsm2_option 1)."""
    p = dict(final_time=6000., time_step=1., frozen_eulerian_fields=True, diffusion_model=1,
             horizontal_diffusivity=0.01)
    if model == 1:
        p.update(model=1)
    else:
        p.update(model=2, time_scheme=5, diffusion_lsm2_option=1)
    return p


def c1_square_workload():
    """The config-1 mesh and fields: 101 x 101 nodes on a 100 m square, each
    square split into two triangles, uniform u = 0.1 m/s, v = 0, depth 1 m."""
    from tests.cases import structured_square
    x, y, tri = structured_square(101, 100.)
    n = x.size
    fields = dict(dim=2, nlayers=1, times=np.zeros(1), elevation=np.ones((1, n)),
                  velocity_x=np.full((1, n), 0.1), velocity_y=np.zeros((1, n)))
    return x, y, tri, fields


def scenario_streams(cy, nscen):
    """Scenario s draws from the default stream advanced by s << 48, the split
    SolverParallel gives its ranks (solver_parallel.pyx:113-115)."""
    out = []
    for s in range(nscen):
        cy.pcg32_seed(42, 0xda3e06d11bb2b203)
        cy.pcg32_advance(s << 48)
        out.append(cy.random_utils.get_state())
    return out


def point_release(cy, xy, n):
    return cy.LagrangianParticleSet(np.repeat(np.asarray(xy, dtype=np.float64)[None], int(n), 0), dim=2)


def batched_leg(cy, mesh, fields, sort_interval, max_steps=None, chunk=250, tune=True, scenarios=c4_scenarios,
                params=c4_params):
    """The whole batched job: every scenario runs all its steps (or the first
    max_steps of them); every 250 steps the finished scenarios' rows come back
    to the host and the scenarios leave the device.  The clock brackets the
    stepping, collection and retirement (inputs resident on the device)."""
    npart, nsteps, model, xy = scenarios()
    fs = cy.EulerianFieldSet(mesh, **fields)
    streams = scenario_streams(cy, len(npart))
    t0 = time.time()
    psets = [point_release(cy, xy[s], npart[s]) for s in range(len(npart))]
    b = cy.BatchSolver(fs, psets, [params(m) for m in model], streams, nsteps=nsteps,
                       sort_interval=sort_interval)
    log("batched: {} scenarios, {} particles, {}..{} steps, ready ({:.1f} s)".format(
        len(npart), int(npart.sum()), int(nsteps.min()), int(nsteps.max()), time.time() - t0))
    total = int(nsteps.max()) if max_steps is None else min(int(max_steps), int(nsteps.max()))
    t0 = time.perf_counter()
    done, ncollected, tuned, graph_steps = 0, 0, [], 0
    tune_at = [0, total // 2] if tune else []
    while done < total:
        if tune_at and done >= tune_at[0] and total - done > 400:
            tune_at.pop(0)
            # the autotuner's trial steps are steps of the job
            s0 = max(b.steps_left())
            si, bpc, inl, _ = cy.autotune(b, trial_steps=24, log=log)
            done += s0 - max(b.steps_left())
            tuned.append(dict(step=done, sort_interval=si, blocks_per_cu=bpc, inline_normals=inl))
            continue
        k = min(chunk - done % chunk, total - done)
        b.run(k)
        graph_steps += b.ctx.last_graph_steps()
        done += k
        # finished scenarios: rows back to the host, then off the device
        fin = [s for s in range(b.nscen) if b.steps_left(s) == 0]
        if fin and done < total:
            for s in fin:
                b.collect(s)
            ncollected += len(fin)
            if len(fin) < b.nscen:
                b.retire(fin)
        log("batched: {} / {} steps, {} scenarios on the device, {:.1f} s".format(
            done, total, b.nscen, time.perf_counter() - t0))
    for s in range(b.nscen):
        b.collect(s)
    wall = time.perf_counter() - t0
    k_ms, tot_ms, nlaunch = b.ctx.last_timing()
    work = float(np.sum(npart * np.minimum(nsteps, total)))
    b.ctx.close()
    return dict(scenarios=len(npart), particles=int(npart.sum()), lsm2_scenarios=int(np.sum(model == 2)),
                particle_steps=work, steps=total, wall_s=wall, value=work / wall, graph_steps=graph_steps,
                tuning=tuned if tune else [dict(step=0, sort_interval=sort_interval, blocks_per_cu=0)])


def plain_sequential_leg(cy, mesh, fields, ksteps=40, warm=3, nsteps_cap=None, scenarios=c4_scenarios,
                         params=c4_params):
    """The same scenarios, one after another, each with the plain step
    (cyl_step on a context holding the mesh and fields, the scenario's pool
    uploaded and initialized as Solver does).  Each scenario is timed over its
    first `ksteps` steps and its job time taken as T_s times that per-step
    time."""
    import ctypes as C
    from cylag_amd import _hip
    from cylag_amd.parameters import Parameters
    from cylag_amd.solver import _params_struct, _pconst_struct, upload_fields, upload_mesh
    npart, nsteps, model, xy = scenarios()
    if nsteps_cap is not None:
        nsteps = np.minimum(nsteps, nsteps_cap)
    fs = cy.EulerianFieldSet(mesh, **fields)
    streams = scenario_streams(cy, len(npart))
    ctx = _hip.Context(0)
    lib, h, P = ctx.lib, ctx.h, _hip.ptr
    upload_mesh(ctx, mesh)
    upload_fields(ctx, fs)
    t_job = 0.
    t0 = time.perf_counter()
    for s in range(len(npart)):
        p = Parameters(params(model[s]))
        ps = point_release(cy, xy[s], npart[s])
        ps.rho0 = p.water_density
        if p.model == 2:
            ps._initialize_model_coefficients()
        par, pc = _params_struct(p, ps), _pconst_struct(ps)
        n = int(npart[s])
        h_ = ps._h
        d = lambda k: np.ascontiguousarray(h_[k], dtype=np.float64)
        tri = np.full(n, -2, dtype=np.int32)
        init_tri = np.full(n, -3, dtype=np.int32)
        frozen = bool(p.frozen_eulerian_fields)
        ctx.check(lib.cyl_fields_reset(h, int(frozen), 0.), "cyl_fields_reset")
        if not frozen:
            ctx.check(lib.cyl_fields_interpolate(h, 0.), "cyl_fields_interpolate")
        ctx.check(lib.cyl_particles_upload(h, n, 2, P(d("position")), P(d("velocity")),
                                           P(d("fluid_velocity_seen")), P(np.zeros(n, np.int32), "i"), P(tri, "i"),
                                           P(np.zeros(n, np.int32), "i"), P(d("diameter")), P(d("volume")),
                                           P(init_tri, "i"), C.byref(pc)), "cyl_particles_upload")
        ctx.check(lib.cyl_rng_set(h, *streams[s]), "cyl_rng_set")
        dt = float(p.time_step)
        ctx.check(lib.cyl_init_particles(h, C.byref(par), 0., dt), "cyl_init_particles")
        t = C.c_double(0.)
        ctx.check(lib.cyl_step(h, C.byref(par), C.byref(t), dt, warm, None), "cyl_step")
        ta = time.perf_counter()
        ctx.check(lib.cyl_step(h, C.byref(par), C.byref(t), dt, ksteps, None), "cyl_step")
        per_step = (time.perf_counter() - ta) / ksteps
        t_job += per_step * int(nsteps[s])
        if s % 32 == 31:
            log("plain sequential: {} / {} scenarios ({:.1f} s)".format(s + 1, len(npart), time.perf_counter() - t0))
    ctx.close()
    work = float(np.sum(npart * nsteps))
    return dict(value=work / t_job, job_s_estimated=t_job, steps_timed_per_scenario=ksteps)


def c5_requests(cy, x, y, h, window, lam=20., seed=13, dt=10.):
    """Poisson arrivals (lam per second) over `window` seconds; each request a
    point release of N ~ logU(1e3, 2e6) particles at a random interior point
    and depth, simulated for U[1, 48] h at dt = 10 s, LSM1 or LSM2 (3D,
    k-epsilon K(z)); settling diameters per particle; request i draws from
    the default stream advanced by i << 48."""
    from functools import partial
    from scipy.spatial import cKDTree
    lx, ly = 20000., 5000.
    rng = np.random.default_rng(seed)
    arr, t = [], 0.
    while True:
        t += rng.exponential(1. / lam)
        if t > window:
            break
        arr.append(t)
    m = len(arr)
    npart = np.exp(rng.uniform(np.log(1e3), np.log(2e6), m)).astype(np.int64)
    nsteps = np.round(rng.uniform(1., 48., m) * 3600. / dt).astype(np.int64)
    model = rng.integers(1, 3, m)
    px, py = rng.uniform(0.05 * lx, 0.95 * lx, m), rng.uniform(0.05 * ly, 0.95 * ly, m)
    hn = h[cKDTree(np.stack([x, y], 1)).query(np.stack([px, py], 1))[1]]
    pz = -hn * rng.uniform(0.1, 0.9, m)
    streams = scenario_streams(cy, m)

    def make(i):
        n = int(npart[i])
        ps = cy.LagrangianParticleSet(np.repeat([[px[i], py[i], pz[i]]], n, 0), dim=3, particle_density=2650.,
                                      buoyancy_velocity_model=1)
        from tests.cases import settling_diameters
        d = settling_diameters(np.random.default_rng((seed, i)), n)
        ps.diameter = d
        ps.volume = np.pi * d ** 3 / 6.
        return ps

    def params(mdl):
        p = dict(final_time=1e9, time_step=dt, frozen_eulerian_fields=True, diffusion_model=3)
        if mdl == 1:
            p.update(model=1)
        else:
            p.update(model=2, time_scheme=5, diffusion_lsm2_option=1, particle_velocity_init=1)
        return p

    return [cy.Request(None, params(model[i]), streams[i], int(nsteps[i]), arrival=arr[i], make=partial(make, i),
                       npart=int(npart[i]), dim=3) for i in range(m)]


def c5_mean_work(dt=10.):
    """Expected particle-steps of one request: E[N] E[steps] for N ~ logU(1e3,
    2e6) and steps = round(U[1, 48] h / dt) (independent draws)."""
    lo, hi = 1e3, 2e6
    return (hi - lo) / np.log(hi / lo) * (24.5 * 3600. / dt)


def served_run(cy, fs, x, y, h, lam, window, horizon, max_rows, tick_steps, sort_interval):
    """One request loop on the 3D channel: Poisson arrivals at `lam` per second
    over `window` seconds, stopped `horizon` seconds after the start."""
    from cylag_amd.serve import latency_stats
    reqs = c5_requests(cy, x, y, h, window, lam=lam)
    last = [0.]

    def progress(msg):
        now = time.time()
        if now - last[0] > 10.:
            last[0] = now
            log(msg)

    loop = cy.ServeLoop(fs, tick_steps=tick_steps, sort_interval=sort_interval, max_rows=max_rows,
                        keep_results=False, log=progress)
    done = loop.serve(reqs, horizon=horizon)
    loop.close()
    st = latency_stats(done, loop.unfinished, horizon=horizon)
    offered = sum(r.npart * r.nsteps for r in reqs) / window
    out = dict(lam_per_s=lam, window_s=window, horizon_s=horizon, arrived=len(reqs),
               admitted=sum(r.admitted is not None for r in reqs), completed=st["completed"],
               completed_frac=st.get("completed_frac"),
               latency_p50_s=st["p50_s"], latency_p99_s=st["p99_s"],
               latency_p50_censored_s=st.get("p50_censored_s"), latency_p99_censored_s=st.get("p99_censored_s"),
               value=loop.particle_steps / loop.elapsed, unit="particle-steps/s",
               offered_particle_steps_per_s=offered, max_running=loop.max_running, capacity_rows=loop.capacity,
               tick_steps=tick_steps)
    log("served: lambda {} /s: {} arrived, {} admitted, {} completed ({:.0%}), p50 {} s, p99 {} s, censored p50 {} s, "
        "p99 {} s, {:.3e} particle-steps/s".format(lam, out["arrived"], out["admitted"], out["completed"],
                                                  out["completed_frac"] or 0., out["latency_p50_s"],
                                                  out["latency_p99_s"], out["latency_p50_censored_s"],
                                                  out["latency_p99_censored_s"], out["value"]))
    return out


def served_leg(cy, window=60., horizon=75., lam=20., max_rows=20_000_000, tick_steps=10, sort_interval=8,
               stable_window=40., stable_horizon=75., stable_margin=0.7):
    """The request loop on the 3D channel at the stated rate (lam = 20
    requests/s, far above what one GPU serves), then at a rate below
    saturation: the saturated throughput over the requests' mean work gives
    the largest stable rate, and a second run at stable_margin times that
    rate measures the latencies there.  Latencies are reported over the
    completed requests and over every arrival (unfinished requests counted
    with horizon - arrival, which makes those percentiles lower bounds)."""
    t0 = time.time()
    x, y, tri, h, fields = c3_workload()
    mesh = cy.TriangularMesh(x, y, tri)
    fs = cy.EulerianFieldSet(mesh, **fields)
    log("served: 3D channel {} elements ({:.1f} s)".format(tri.shape[0], time.time() - t0))
    out = served_run(cy, fs, x, y, h, lam, window, horizon, max_rows, tick_steps, sort_interval)
    work = c5_mean_work()
    lam_stable = out["value"] / work
    out["workload"] = ("3D sigma-layer channel (~500k elements, 10 planes, frozen); Poisson {} req/s (seed 13); "
                       "N ~ logU(1e3, 2e6), 1-48 h at dt = 10 s, LSM1 or LSM2 with settling; admission by free HBM "
                       "capped at {} rows, small requests pass a blocked large one until it has waited 10 s").format(
                           lam, max_rows)
    out["mean_request_particle_steps"] = work
    out["lambda_stable_per_s"] = lam_stable
    out["stable"] = served_run(cy, fs, x, y, h, round(stable_margin * lam_stable, 3), stable_window, stable_horizon,
                               max_rows, tick_steps, sort_interval)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1990)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--npart", type=int, default=10_000_000)
    ap.add_argument("--sort-interval", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=None,
                    help="threads of the all-core CPU baseline (default: the CPU share granted to this process)")
    ap.add_argument("--normals", type=int, default=0, choices=[0, 1, 2],
                    help="plain leg: where the fused kernel takes its normals (0: pool-order generator, 1: sampled "
                         "inline, 2: slot-order generator)")
    ap.add_argument("--graph", type=int, default=0, choices=[0, 1],
                    help="plain leg: replay the step loop from a captured HIP graph")
    ap.add_argument("--legs", default="plain,c3,batched,sequential,batched_square,served",
                    help="comma-separated: plain (the 10M-particle step), c3 (3D LSM-2 with settling, 5M particles), "
                         "batched (256 scenarios in one batch), "
                         "sequential (the same scenarios one after another with the plain step), batched_square "
                         "(the scenario set on the config-1 square, batched and one after another), served "
                         "(Poisson requests on the 3D channel)")
    ap.add_argument("--served-window", type=float, default=60.,
                    help="served leg: seconds of arrivals (the loop stops 15 s later)")
    ap.add_argument("--batch-steps", type=int, default=None,
                    help="run only the first steps of the batched job (profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the legs, copy 1 GiB device to device (for counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the step runs on one MI355X holding the whole mesh "
              "and particle set; --gpus {} is not supported".format(args.gpus))
        sys.exit(2)
    legs = set(args.legs.split(","))
    import cylag_amd as cy
    if legs & {"plain", "batched", "sequential"}:
        t0 = time.time()
        x, y, tri, fields, pos, params = c2_workload(npart=args.npart if "plain" in legs else 1)
        log("workload built: {} nodes, {} triangles, {} particles ({:.1f} s)".format(
            x.size, tri.shape[0], pos.shape[0], time.time() - t0))
        mesh = cy.TriangularMesh(x, y, tri)
    line = {"metric": METRIC, "unit": "particle-steps/s", "n_gpus": 1, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    if "plain" in legs:
        line.update(plain_leg(cy, args, mesh, fields, pos, params))
    if "c3" in legs:
        line["c3"] = c3_leg(cy)
    bt = None
    if "batched" in legs:
        bt = batched_leg(cy, mesh, fields, args.sort_interval, args.batch_steps)
        log("batched: {:.3e} particle-steps/s over {:.1f} s".format(bt["value"], bt["wall_s"]))
        bt["workload"] = ("256 point releases on the same mesh and fields, N ~ logU(1e3, 1e6) (seed 11), "
                          "500-5000 steps of 2 s, LSM1 or 2D LSM2 per scenario, streams 42 advanced by s<<48")
        line["batched"] = bt
    if "sequential" in legs:
        seq = plain_sequential_leg(cy, mesh, fields, nsteps_cap=args.batch_steps)
        log("plain sequential: {:.3e} particle-steps/s (job {:.1f} s estimated)".format(
            seq["value"], seq["job_s_estimated"]))
        if bt is not None:
            bt["plain_sequential"] = seq
            bt["speedup_vs_plain_sequential"] = bt["value"] / seq["value"]
        else:
            line["plain_sequential"] = seq
    if "batched_square" in legs:
        xs, ys, ts, fsq = c1_square_workload()
        msq = cy.TriangularMesh(xs, ys, ts)
        bq = batched_leg(cy, msq, fsq, args.sort_interval, args.batch_steps, scenarios=c4_square_scenarios,
                         params=c4_square_params)
        seq = plain_sequential_leg(cy, msq, fsq, nsteps_cap=args.batch_steps, scenarios=c4_square_scenarios,
                                   params=c4_square_params)
        bq["plain_sequential"] = seq
        bq["speedup_vs_plain_sequential"] = bq["value"] / seq["value"]
        bq["workload"] = ("the batched scenario set on the config-1 square (101 x 101 nodes, 20,000 elements, "
                          "u = 0.1 m/s, K = 0.01 m2/s, dt = 1 s, frozen, walls): 256 point releases, "
                          "N ~ logU(1e3, 1e6), 500-5000 steps, LSM1 or 2D LSM2")
        log("batched on the square: {:.3e} particle-steps/s, {:.1f}x one after another".format(
            bq["value"], bq["speedup_vs_plain_sequential"]))
        line["batched_square"] = bq
    if "served" in legs:
        line["served"] = served_leg(cy, window=args.served_window, horizon=args.served_window + 15.)
    if "plain" in legs and not args.no_cpu:
        log("cpu baseline ...")
        cb = cpu_baseline(x, y, tri, fields, pos, params, nthreads=args.cpu_threads)
        cb["gpu_over_cpu"] = line["value"] / cb["value"]
        cb["gpu_over_one_core"] = line["value"] / cb["single_core"]["value"]
        line["cpu_baseline"] = cb
    if args.calibrate_copy:
        calibrate_copy()
    print(json.dumps(line), flush=True)


def plain_leg(cy, args, mesh, fields, pos, params):
    fs = cy.EulerianFieldSet(mesh, **fields)
    ps = cy.LagrangianParticleSet(pos, dim=2)
    cy.pcg32_seed(42, 0xda3e06d11bb2b203)
    t0 = time.time()
    solver = cy.Solver(fs, ps, params)
    solver.ctx.lib.cyl_set_tuning(solver.ctx.h, args.sort_interval, 0, args.normals, args.graph)
    log("solver ready ({:.1f} s): {}".format(time.time() - t0, solver.ctx.device_info()[0]))
    solver._run(args.warmup)
    log("warmup done")
    t0 = time.perf_counter()
    solver._run(args.steps)
    wall = time.perf_counter() - t0
    k_ms, tot_ms, nlaunch = solver.ctx.last_timing()
    prep_ms = solver.ctx.last_prep_ms()
    ps._sync()
    nact = ps.npart_active
    log("timed: {} steps in {:.3f} s, step kernel {:.3f} ms/launch over {} launches, normals {:.3f} ms".format(
        args.steps, wall, k_ms / max(nlaunch, 1), nlaunch, prep_ms / max(nlaunch, 1)))
    solver.ctx.close()
    psteps = nact * args.steps / wall
    b_alg, b_state, b_mesh = bytes_per_particle_step(mesh.nt, mesh.nv, nact)
    traffic, digest = measured_traffic()
    avg_ms = k_ms / max(nlaunch, 1)
    achieved = b_alg * nact / (avg_ms * 1e-3)
    return {
        "value": psteps,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "config": {"workload": "2D LSM1 k-eps, Delaunay 1M pts / {} elements, {} particles, 2 frames, "
                               "dt=5 s, Ziggurat".format(mesh.nt, nact),
                   "elements": int(mesh.nt), "nodes": int(mesh.nv), "particles": int(nact),
                   "rng_method": 2, "sort_interval": args.sort_interval},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": "profiles/k_lsm1_2d_unstructured_keps_10M_counters.md",
                     "traffic_kernel_sources": digest,
                     "kernel": "k_lsm1_2d<2,true>", "kernel_avg_ms": avg_ms,
                     "bytes_per_particle_step": b_alg,
                     "normals_kernel_avg_ms": prep_ms / max(nlaunch, 1)},
    }


if __name__ == "__main__":
    main()
