"""Batched scenarios: each scenario's result in the batch equals the result of
running it alone with a plain Solver (same stream, same steps) and the
oracle's serial run of it (the reference's step with that stream): element
ids, flags, tags and random stream identical, positions within 1e-9
relative."""
import numpy as np
import pytest

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import open_channel_case, square_case, vortex_case

pytestmark = pytest.mark.gpu

TOL = 1e-9


def streams(n):
    """Scenario s: the default stream advanced by s << 48 (solver_parallel.pyx:113-115)."""
    out = []
    for s in range(n):
        g = R.RefRng(42, 0xda3e06d11bb2b203)
        g.advance(s << 48)
        out.append(g.state())
    return out


def pick(case, npart, seed):
    rng = np.random.default_rng(seed)
    return case.position[rng.choice(case.position.shape[0], npart, replace=False)].copy()


def make_pset(case, npart, seed):
    return cy.LagrangianParticleSet(pick(case, npart, seed), dim=2)


def run_oracle(case, pos, params, state, nsteps, trifinder=True, pset_kw=None, diameter=None):
    """The scenario through the oracle: the reference's serial step on its own
    pool, from the scenario's stream (solver_parallel.pyx:113-115 gives each
    rank such a stream)."""
    mesh = R.RefMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = R.RefFields(mesh, **case.fields)
    ps = R.RefParticles(pos, dim=pos.shape[1], diameter=diameter, **(pset_kw or {}))
    rng = R.RefRng()
    rng.set_state(*state)
    o = R.RefSolver(fs, ps, params, trifinder=case.trifinder() if trifinder else None, rng=rng)
    o.forward(nsteps)
    return dict(position=ps.position.copy(), tri=ps.tri.copy(), inactive=ps.inactive.copy(), tag=ps.tag.copy(),
                velocity=ps.velocity.copy(), lowerlayer=ps.lowerlayer.copy(), rng=rng.state(),
                bnd=np.array(o.bnd_history, dtype=np.int64).reshape(nsteps, mesh.nopen))


def run_alone(case, pset, params, state, nsteps):
    mesh = cy.TriangularMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    cy.random_utils.set_state(*state)
    s = cy.Solver(fs, pset, params)
    for _ in range(nsteps):
        s.forward()
    ps = s.particle_set
    out = dict(position=ps.position.copy(), tri=ps.tri.copy(), inactive=ps.inactive.copy(),
               tag=ps.tag.copy(), velocity=ps.velocity.copy(), rng=cy.random_utils.get_state())
    s.ctx.close()
    return out


def compare(b, a, rng_b):
    assert np.array_equal(b.inactive, a["inactive"])
    assert np.array_equal(b.tri, a["tri"])
    assert np.array_equal(b.tag, a["tag"])
    if "lowerlayer" in a:
        assert np.array_equal(b.lowerlayer, a["lowerlayer"])
    scale = np.maximum(np.abs(a["position"]), 1.)
    assert np.max(np.abs(b.position - a["position"]) / scale) <= TOL
    vs = np.maximum(np.abs(a["velocity"]), 1.)
    assert np.max(np.abs(b.velocity - a["velocity"]) / vs) <= TOL
    assert tuple(rng_b) == tuple(a["rng"])


MIX = [
    (dict(model=1, diffusion_model=2), 1500, 30),                     # fused kernel, k-eps
    (dict(model=1, diffusion_model=1, horizontal_diffusivity=0.3), 900, 12),  # fused kernel, constant K
    (dict(model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1,
          particle_velocity_init=1), 700, 25),                         # generic LSM2 DI
    (dict(model=1, time_scheme=4, diffusion_model=0), 500, 20),        # generic LSM1 RK4
    (dict(model=1, diffusion_model=2), 1200, 0),                       # no steps at all
]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("normals", [0, 2], ids=["pool_generator", "slot_generator"])
def test_batch_equals_alone(normals, rng_method):
    case = vortex_case(npart=4000)
    base = dict(final_time=10000., time_step=5., rng_method=rng_method)
    params = [dict(base, **p) for p, _, _ in MIX]
    nsteps = [k for _, _, k in MIX]
    st = streams(len(MIX))
    psets = [make_pset(case, n, 100 + i) for i, (_, n, _) in enumerate(MIX)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    b = cy.BatchSolver(fs, psets, params, st, nsteps=nsteps, sort_interval=3)
    b.set_tuning(3, 0, normals)
    b.run()
    out = b.results()
    for s, (p, n, k) in enumerate(MIX):
        alone = run_alone(case, make_pset(case, n, 100 + s), params[s], st[s], k)
        compare(out[s], alone, b.rng_state(s))
        compare(out[s], run_oracle(case, pick(case, n, 100 + s), params[s], st[s], k), b.rng_state(s))
    b.ctx.close()


def test_batch_deletions_tags_and_statistics(tmp_path):
    """Open boundaries: per-scenario deletions, tag renumbering and statistics."""
    case = open_channel_case()
    base = dict(final_time=1e9, time_step=1., frozen_eulerian_fields=True, model=1, bnd_statistics=True,
                output_rep=str(tmp_path))
    mix = [(dict(diffusion_model=1, horizontal_diffusivity=2.), 400, 70),
           (dict(diffusion_model=1, horizontal_diffusivity=0.5, stranding_probability=0.3), 300, 40),
           (dict(time_scheme=2, diffusion_model=1, horizontal_diffusivity=1.), 350, 55)]
    params = [dict(base, **p) for p, _, _ in mix]
    st = streams(len(mix))
    mesh = lambda: cy.TriangularMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points,
                                     nopen=case.nopen)
    fs = cy.EulerianFieldSet(mesh(), trifinder=case.trifinder(), **case.fields)
    psets = [make_pset(case, n, 200 + i) for i, (_, n, _) in enumerate(mix)]
    b = cy.BatchSolver(fs, psets, params, st, nsteps=[k for _, _, k in mix], sort_interval=5)
    b.run()
    out = b.results()
    stats = b.bnd_stats
    for s, (p, n, k) in enumerate(mix):
        pset = make_pset(case, n, 200 + s)
        fs1 = cy.EulerianFieldSet(mesh(), trifinder=case.trifinder(), **case.fields)
        cy.random_utils.set_state(*st[s])
        solo = cy.Solver(fs1, pset, params[s])
        for _ in range(k):
            solo.forward()
        ps = solo.particle_set
        alone = dict(position=ps.position.copy(), tri=ps.tri.copy(), inactive=ps.inactive.copy(),
                     tag=ps.tag.copy(), velocity=ps.velocity.copy(), rng=cy.random_utils.get_state())
        assert alone["inactive"].sum() > 0
        compare(out[s], alone, b.rng_state(s))
        np.testing.assert_array_equal(stats[s], solo.bnd_stats[1][:k])
        solo.ctx.close()
        if params[s].get("stranding_probability", 0.) == 0.:
            # stranding draws come from per-row sub-streams on the device, so a
            # scenario that strands is compared with its plain run only
            ref = run_oracle(case, pick(case, n, 200 + s), params[s], st[s], k)
            compare(out[s], ref, b.rng_state(s))
            np.testing.assert_array_equal(stats[s], ref["bnd"])
    b.ctx.close()


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_continuous_batching_equals_alone(rng_method):
    """Scenarios leave a running batch and others join it (two frames of
    time-dependent fields): each one's result equals a plain Solver run of that
    scenario alone, built at the batch time at which it joined."""
    case = vortex_case(npart=4000)
    dt = 5.
    base = dict(final_time=10000., time_step=dt, rng_method=rng_method)
    spec = [  # (parameters, particles, steps, joins after batch step)
        (dict(model=1, diffusion_model=2), 1500, 12, 0),
        (dict(model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1, particle_velocity_init=1), 700, 30, 0),
        (dict(model=1, time_scheme=4, diffusion_model=0), 500, 30, 0),
        (dict(model=1, diffusion_model=1, horizontal_diffusivity=0.3), 900, 15, 12),
        (dict(model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=2, particle_velocity_init=1), 600, 10, 12),
    ]
    st = streams(len(spec))
    params = [dict(base, **p) for p, _, _, _ in spec]
    psets = [make_pset(case, n, 300 + i) for i, (_, n, _, _) in enumerate(spec)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    b = cy.BatchSolver(fs, psets[:3], params[:3], st[:3], nsteps=[k for _, _, k, _ in spec[:3]], sort_interval=4)
    b.run(12)
    done = {}
    s0 = b.index(0)
    done[0] = (b.collect(s0), b.rng_state(s0))
    b.retire([s0])
    ids = b.admit(psets[3:], params[3:], st[3:], nsteps=[k for _, _, k, _ in spec[3:]])
    assert ids == [3, 4]
    b.run(10)
    s4 = b.index(4)
    assert b.steps_left(s4) == 0
    done[4] = (b.collect(s4), b.rng_state(s4))
    b.retire([s4])
    b.run()
    for sid in (1, 2, 3):
        s = b.index(sid)
        done[sid] = (b.collect(s), b.rng_state(s))
    b.close()
    for sid, (p, n, k, join) in enumerate(spec):
        pj = dict(params[sid], initial_time=join * dt)
        alone = run_alone(case, make_pset(case, n, 300 + sid), pj, st[sid], k)
        compare(done[sid][0], alone, done[sid][1])
        compare(done[sid][0], run_oracle(case, pick(case, n, 300 + sid), pj, st[sid], k), done[sid][1])


def test_batch_graph_replay_equals_alone():
    """Frozen fields: the batch replays captured graphs of its step loop; every
    scenario still equals its plain run alone."""
    case = vortex_case(npart=4000)
    base = dict(final_time=1e9, time_step=5., rng_method=2, frozen_eulerian_fields=True)
    params = [dict(base, **p) for p, _, _ in MIX]
    nsteps = [k + 20 for _, _, k in MIX]
    st = streams(len(MIX))
    psets = [make_pset(case, n, 100 + i) for i, (_, n, _) in enumerate(MIX)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    b = cy.BatchSolver(fs, psets, params, st, nsteps=nsteps, sort_interval=3, use_graph=True)
    b.run()
    assert b.ctx.last_graph_steps() > 0
    out = b.results()
    for s, (p, n, k) in enumerate(MIX):
        alone = run_alone(case, make_pset(case, n, 100 + s), params[s], st[s], nsteps[s])
        compare(out[s], alone, b.rng_state(s))
        compare(out[s], run_oracle(case, pick(case, n, 100 + s), params[s], st[s], nsteps[s]), b.rng_state(s))
    b.ctx.close()


def square_scenarios(nscen=8, seed=11):
    """Release scenarios on the plain workload's mesh (101 x 101 nodes, 20,000
    triangles, uniform 0.1 m/s flow, walls): uneven particle counts
    (log-uniform in [5, 5000]), uneven step counts, LSM-1 or 2D LSM-2 per
    scenario, random interior release points and one on the centre vertex."""
    rng = np.random.default_rng(seed)
    npart = np.exp(rng.uniform(np.log(5), np.log(5000), nscen)).astype(int)
    nsteps = rng.integers(10, 300, nscen)
    model = np.array([1, 2] * (nscen // 2))
    xy = np.stack([rng.uniform(5., 95., nscen), rng.uniform(5., 95., nscen)], 1)
    xy[0] = 50.
    return npart, nsteps, model, xy


def square_params(model, rng_method):
    p = dict(final_time=1e9, time_step=1., frozen_eulerian_fields=True, diffusion_model=1,
             horizontal_diffusivity=0.01, rng_method=rng_method)
    if model == 1:
        p.update(model=1)
    else:
        p.update(model=2, time_scheme=5, diffusion_lsm2_option=1, particle_velocity_init=1)
    return p


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_batch_square_mesh_uneven_mixed(rng_method):
    """Uneven scenarios of mixed LSM-1 / LSM-2 on the 20,000-triangle square,
    each against the oracle and against its plain run alone; the device
    locates the releases itself (no trifinder), as the batched workload does."""
    case = square_case(n=101, npart=1)
    case.use_trifinder = False
    npart, nsteps, model, xy = square_scenarios()
    params = [square_params(m, rng_method) for m in model]
    st = streams(len(npart))
    pos = [np.repeat(xy[s][None], npart[s], 0) for s in range(len(npart))]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, **case.fields)
    b = cy.BatchSolver(fs, [cy.LagrangianParticleSet(q.copy(), dim=2) for q in pos], params, st, nsteps=nsteps,
                       sort_interval=5)
    b.run()
    out = b.results()
    for s in range(len(npart)):
        ref = run_oracle(case, pos[s], params[s], st[s], int(nsteps[s]), trifinder=False)
        compare(out[s], ref, b.rng_state(s))
        alone = run_alone(case, cy.LagrangianParticleSet(pos[s].copy(), dim=2), params[s], st[s], int(nsteps[s]))
        compare(out[s], alone, b.rng_state(s))
    b.ctx.close()


@pytest.mark.parametrize("rng_method", [2, 3])
def test_batch_dead_zone_general_rerun(rng_method):
    """Rows in a zone without turbulence draw nothing, so the draws per row are
    not uniform: each batched step is rerun in general mode, whose event tables
    are laid out from the counted units of every scenario.  Each scenario
    against its plain run alone and against the oracle."""
    case = vortex_case(npart=4000, dead_zone=True)
    base = dict(final_time=10000., time_step=5., rng_method=rng_method, model=1, diffusion_model=2)
    counts, steps = [900, 700, 400], [18, 11, 0]
    params = [dict(base) for _ in counts]
    st = streams(len(counts))
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    psets = [make_pset(case, n, 300 + i) for i, n in enumerate(counts)]
    assert all((ps.position[:, 0] < 500.).any() for ps in psets), "every scenario has rows in the dead zone"
    b = cy.BatchSolver(fs, psets, params, st, nsteps=steps, sort_interval=4)
    b.run()
    out = b.results()
    for s, n in enumerate(counts):
        alone = run_alone(case, make_pset(case, n, 300 + s), params[s], st[s], steps[s])
        compare(out[s], alone, b.rng_state(s))
        compare(out[s], run_oracle(case, pick(case, n, 300 + s), params[s], st[s], steps[s]), b.rng_state(s))
    b.ctx.close()


@pytest.mark.parametrize("rng_method", [2, 3])
def test_batch_large_scenarios_many_spans(rng_method):
    """Three scenarios of 1e5-3e5 particles: each one's stream window covers
    many spans of the event tables, whose reach keys, counts and bucket
    entries carry across spans and scenario boundaries.  Each scenario against
    its plain run alone and against the oracle."""
    case = vortex_case(nx=201, ny=51, lx=10000., ly=2500., npart=600_000, seed=12, wavelength=2000.)
    base = dict(final_time=10000., time_step=5., rng_method=rng_method, model=1, diffusion_model=2)
    counts, steps = [300_000, 180_000, 100_000], [3, 2, 3]
    params = [dict(base) for _ in counts]
    st = streams(len(counts))
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    b = cy.BatchSolver(fs, [make_pset(case, n, 400 + i) for i, n in enumerate(counts)], params, st, nsteps=steps,
                       sort_interval=2)
    b.run()
    out = b.results()
    for s, n in enumerate(counts):
        alone = run_alone(case, make_pset(case, n, 400 + s), params[s], st[s], steps[s])
        compare(out[s], alone, b.rng_state(s))
        compare(out[s], run_oracle(case, pick(case, n, 400 + s), params[s], st[s], steps[s]), b.rng_state(s))
    b.ctx.close()
