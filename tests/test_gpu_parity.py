"""HIP step (through the C ABI) against the oracle on the same seeded inputs:
identical element ids, pool flags, tags and random-stream state; positions
within 1e-9 relative (FP64)."""
from types import SimpleNamespace

import numpy as np
import pytest

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import (SEED, open_channel_case, reference_square_case, square_case, vortex_3d_case,
                         vortex_case)
from tests.test_oracle_golden import (GOLDEN_3D, GOLDEN_3D_IDS, LSM2_GOLDEN, P4, golden_3d_setup,
                                      lsm2_params)

pytestmark = pytest.mark.gpu

TOL = 1e-9


def assert_parity(g, r, check_velocity=True):
    assert np.array_equal(g.inactive, r.inactive), "pool flags differ"
    assert np.array_equal(g.tri, r.tri), "element ids differ at rows {}".format(
        np.nonzero(g.tri != r.tri)[0][:10])
    assert np.array_equal(g.lowerlayer, r.lowerlayer)
    assert np.array_equal(g.tag, r.tag), "tags differ"
    scale = np.maximum(np.abs(r.position), 1.)
    err = np.max(np.abs(g.position - r.position) / scale) if r.position.size else 0.
    assert err <= TOL, "position error {:.3e}".format(err)
    if check_velocity:
        verr = np.max(np.abs(g.velocity - r.velocity) / np.maximum(np.abs(r.velocity), 1.))
        assert verr <= TOL, "velocity error {:.3e}".format(verr)
    assert cy.random_utils.get_state() == r.rng.state(), "random stream state differs"


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_plain_diffusion_from_a_vertex(rng_method):
    """All particles start on a mesh vertex: the first walk circles the vertex fan."""
    case = square_case(n=21, npart=400)
    params = dict(final_time=60., time_step=1., frozen_eulerian_fields=True, model=1,
                  diffusion_model=1, horizontal_diffusivity=0.5, rng_method=rng_method)
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
def test_lsm1_schemes_reference_mesh(scheme):
    pos = np.array([[5000., 0.], [5000., 500.], [1000., 0.], [1000., 500.], [-7000., 3000.]])
    case = reference_square_case(pos, pool=8)
    params = dict(final_time=3600., time_step=100., time_scheme=scheme,
                  frozen_eulerian_fields=True, model=1, diffusion_model=1,
                  horizontal_diffusivity=20.)
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


def test_lsm1_golden_rk4():
    # tests/test_20_lsm1_analytical_advection_2d.py:242-246 through the product
    pos = np.array([[5000., 0.], [5000., 500.], [1000., 0.], [1000., 500.]])
    case = reference_square_case(pos, pool=1000)
    params = dict(final_time=3600., time_step=200., time_scheme=4, frozen_eulerian_fields=True,
                  model=1, diffusion_model=0)
    g = case.run_gpu(cy, params)
    np.testing.assert_allclose(g.position[g.inactive == 0], np.array(
        [[4.99888531e+03, 3.71825833e+00], [4.99851348e+03, 5.03606789e+02],
         [9.99777062e+02, 7.43651666e-01], [9.99405236e+02, 5.00632183e+02]]), rtol=1e-7)


def test_model0_rebound_golden():
    pos = np.full((11, 2), 7500.)
    vel = 10. * np.array([[1., 0.1 * k] for k in range(11)])
    case = reference_square_case(pos, vel, pool=12)
    params = dict(final_time=5000., time_step=50., time_scheme=1, frozen_eulerian_fields=True,
                  model=0, particle_velocity_init=0, boundary_conditions=True,
                  max_wall_rebounds=4, rebound_damping_coef=0.)
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    assert_parity(g, r)
    xp = g.position[np.argsort(g.inactive, kind="stable")][:int((g.inactive == 0).sum())]
    from tests.test_oracle_golden import REBOUND_EXPECTED
    np.testing.assert_allclose(xp, REBOUND_EXPECTED, rtol=1e-7)


@pytest.mark.parametrize("ts,dt,dm,opt,tf,expected", LSM2_GOLDEN,
                         ids=["euler", "rk2", "rk4", "di1", "di_opt2", "euler_maruyama"])
def test_lsm2_2d_golden(ts, dt, dm, opt, tf, expected):
    """LSM2 (velocity as state) on the reference tests' rotation field."""
    case = reference_square_case(P4, pool=1000)
    params = lsm2_params(ts, dt, dm, opt, tf)
    g = case.run_gpu(cy, params)
    assert_parity(g, case.run_oracle(R, params))
    np.testing.assert_allclose(g.position[g.inactive == 0], np.array(expected), rtol=1e-7)


@pytest.mark.parametrize("kind,model,ts,dt,dm,opt,expected", GOLDEN_3D, ids=GOLDEN_3D_IDS)
def test_3d_and_pseudo3d_golden(kind, model, ts, dt, dm, opt, expected):
    """3D sigma-layer fields (prism interpolation, z-localizer) and 3D particles
    in 2D fields (log-law pseudo-3D corrections)."""
    case, params = golden_3d_setup(kind, model, ts, dt, dm, opt)
    g = case.run_gpu(cy, params)
    assert_parity(g, case.run_oracle(R, params))
    np.testing.assert_allclose(g.position[g.inactive == 0], np.array(expected), rtol=1e-7)


V3D = [
    dict(model=1, diffusion_model=1, horizontal_diffusivity=0.1, vertical_diffusivity=1e-3),
    dict(model=1, diffusion_model=3),
    dict(model=2, time_scheme=5, diffusion_model=3, diffusion_lsm2_option=1, particle_velocity_init=1),
    dict(model=2, time_scheme=5, diffusion_model=3, diffusion_lsm2_option=2, particle_velocity_init=1),
    dict(model=2, time_scheme=1, diffusion_model=1, diffusion_lsm2_option=2, particle_velocity_init=1,
         horizontal_diffusivity=0.1, vertical_diffusivity=1e-3),
]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("which", range(len(V3D)), ids=["lsm1_const", "lsm1_keps3d", "lsm2_di1", "lsm2_di2",
                                                        "lsm2_em"])
def test_3d_stochastic(which, rng_method):
    """Stochastic 3D steps: K(z) from k-epsilon planes and its gradient,
    sigma-layer localization, bottom and surface reflections, two frames."""
    case = vortex_3d_case()
    params = dict(final_time=300., time_step=5., rng_method=rng_method, **V3D[which])
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


@pytest.mark.parametrize("opt", [1, 2, 3])
def test_lsm2_2d_keps_stochastic(opt):
    case = vortex_case(npart=2000)
    params = dict(final_time=200., time_step=5., model=2, time_scheme=5, diffusion_model=2,
                  diffusion_lsm2_option=opt, particle_velocity_init=1)
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("drag", [0, 1, 2])
def test_lsm2_2d_direct_rebounds(drag, rng_method):
    """2D LSM-2, direct integrator without velocity noise (the fused kernel's
    LSM-2 path): drag models, k-epsilon position diffusion, damped wall
    rebounds that reflect the particle velocity, two frames."""
    case = vortex_case(npart=2500)
    case.pset_kw = dict(drag_coefficient_model=drag, particle_diameter=2e-4, particle_density=1800.)
    params = dict(final_time=300., time_step=5., model=2, time_scheme=5, diffusion_model=2,
                  diffusion_lsm2_option=1, particle_velocity_init=1, rng_method=rng_method,
                  rebound_damping_coef=0.3)
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_unstructured_keps_two_frames(rng_method):
    case = vortex_case(npart=3000)
    params = dict(final_time=200., time_step=5., model=1, diffusion_model=2,
                  rng_method=rng_method)
    g, r = case.run_gpu(cy, params), case.run_oracle(R, params)
    assert_parity(g, r)
    # fluid_velocity_seen (not stored by the fused 2D LSM-1 step, derived on download)
    us_err = np.max(np.abs(g.fluid_velocity_seen - r.fluid_velocity_seen) / np.maximum(np.abs(r.fluid_velocity_seen), 1.))
    assert us_err <= TOL


def test_seen_velocity_materialized_on_model_switch():
    """A context stepped with 2D LSM-1 (seen velocity not stored, derived from
    the velocity on download) and then with the test model, which neither
    sets nor reads it: the stored seen velocity must be the last LSM-1
    velocity, materialized at the switch."""
    case = vortex_case(npart=1500)
    p1 = dict(final_time=100., time_step=5., model=1, diffusion_model=2)
    s = case.build_gpu(cy, p1)
    for _ in range(10):
        s.forward()
    ps = s.particle_set
    up_lsm1 = ps.velocity.copy()
    from cylag_amd.parameters import Parameters
    from cylag_amd.solver import _params_struct
    s.parameters = Parameters(dict(p1, model=0, particle_velocity_init=0, boundary_conditions=False))
    s._par = _params_struct(s.parameters, ps)
    s.forward()
    act = ps.inactive == 0
    np.testing.assert_array_equal(ps.fluid_velocity_seen[act], up_lsm1[act])


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_keps_dead_zone_non_uniform_draws(rng_method):
    """Particles in a zone without turbulence draw nothing: the uniform draw
    assignment is rejected and the step reruns with pool-order offsets."""
    case = vortex_case(npart=3000, dead_zone=True)
    params = dict(final_time=200., time_step=5., model=1, diffusion_model=2,
                  rng_method=rng_method)
    r = case.run_oracle(R, params)
    assert_parity(case.run_gpu(cy, params), r)


@pytest.mark.parametrize("rng_method", [2, 3])
def test_large_pool_two_steps(rng_method):
    """2.2M particles on a 20k-element mesh for two steps: multi-tile scans of
    ranks and event tables, many generator chunks, persistent-grid strides."""
    case = vortex_case(nx=201, ny=51, lx=10000., ly=2500., npart=2_200_000, seed=11,
                       wavelength=2000.)
    params = dict(final_time=10., time_step=5., model=1, diffusion_model=2,
                  rng_method=rng_method)
    r = case.run_oracle(R, params)
    assert_parity(case.run_gpu(cy, params), r)


@pytest.mark.parametrize("rng_method", [1, 2])
def test_open_boundary_deletions_and_tags(rng_method, tmp_path):
    """Particles leave through open boundaries: deletions, tag renumbering,
    boundary statistics and the pool-order random stream after them."""
    case = open_channel_case()
    params = dict(final_time=70., time_step=1., frozen_eulerian_fields=True, model=1,
                  diffusion_model=1, horizontal_diffusivity=2., rng_method=rng_method,
                  bnd_statistics=True, output_rep=str(tmp_path))
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    assert (r.inactive[:400] == 1).sum() > 50, "the case must delete particles"
    assert_parity(g, r)
    gs = g.solver.bnd_stats[1]
    assert gs.sum() == np.sum(r.solver.bnd_history)
    np.testing.assert_array_equal(gs, np.array(r.solver.bnd_history))


@pytest.mark.parametrize("which", ["keps_fused", "keps_rk2_generic", "open_deletions"])
def test_element_resort_does_not_change_results(which, tmp_path):
    """Re-sorting the slots by element every 3 steps (state gather, swapped
    per-slot buffers, rank recompute) leaves every result unchanged."""
    if which == "open_deletions":
        case = open_channel_case()
        params = dict(final_time=70., time_step=1., frozen_eulerian_fields=True, model=1,
                      diffusion_model=1, horizontal_diffusivity=2., bnd_statistics=True,
                      output_rep=str(tmp_path))
    else:
        case = vortex_case(npart=2000)
        params = dict(final_time=150., time_step=5., model=1, diffusion_model=2)
        if which == "keps_rk2_generic":
            params["time_scheme"] = 2
    s = case.build_gpu(cy, params)
    s.ctx.lib.cyl_set_tuning(s.ctx.h, 3, 0, 0, 0)
    s.solve()
    r = case.run_oracle(R, params)
    ps = s.particle_set
    g = SimpleNamespace(position=ps.position, velocity=ps.velocity, tri=ps.tri,
                        inactive=ps.inactive, lowerlayer=ps.lowerlayer, tag=ps.tag)
    assert_parity(g, r)
    for name in ("last_position", "last_velocity"):
        gv, rv = getattr(ps, name), getattr(r, name)
        assert np.max(np.abs(gv - rv) / np.maximum(np.abs(rv), 1.)) <= TOL, name


@pytest.mark.parametrize("which,fused", [(0, True), (1, True), (2, True), (3, False), (4, False)],
                         ids=["lsm1_const", "lsm1_keps3d", "lsm2_di1", "lsm2_di2", "lsm2_em"])
def test_3d_kernel_path(which, fused):
    """The fused 3D kernel steps the configurations it covers (and parity holds
    there, test_3d_stochastic); the others run the generic kernel."""
    case = vortex_3d_case()
    params = dict(final_time=50., time_step=5., **V3D[which])
    g = case.run_gpu(cy, params)
    f, gen = g.solver.ctx.last_path()
    assert (f > 0 and gen == 0) if fused else (f == 0 and gen > 0), (f, gen)


@pytest.mark.parametrize("model", [1, 2])
def test_2d_kernel_path(model):
    case = vortex_case(npart=500)
    p = dict(model=1, diffusion_model=2) if model == 1 else dict(
        model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1, particle_velocity_init=1)
    g = case.run_gpu(cy, dict(final_time=50., time_step=5., **p))
    f, gen = g.solver.ctx.last_path()
    assert f > 0 and gen == 0, (f, gen)


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("mode", [1, 2], ids=["inline", "slot_generator"])
def test_inline_normals_keps_two_frames(mode, rng_method):
    """The 2D fused kernel with its normals sampled inline from each slot's
    prepared stream position (mode 1), or drawn per slot at its rank's stream
    position by the slot-order generator (mode 2)."""
    case = vortex_case(npart=3000)
    params = dict(final_time=200., time_step=5., model=1, diffusion_model=2, rng_method=rng_method)
    s = case.build_gpu(cy, params)
    s.ctx.lib.cyl_set_tuning(s.ctx.h, 4, 0, mode, 0)
    s.solve()
    ps = s.particle_set
    g = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                        inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy())
    assert s.ctx.last_path()[0] > 0
    assert_parity(g, case.run_oracle(R, params))


GRAPH_CASES = [
    ("2d_lsm1", lambda: vortex_case(npart=3000), dict(model=1, diffusion_model=2)),
    ("2d_lsm2", lambda: vortex_case(npart=2000), dict(model=2, time_scheme=5, diffusion_model=2,
                                                      diffusion_lsm2_option=1, particle_velocity_init=1)),
    ("3d_lsm1", lambda: vortex_3d_case(), dict(model=1, diffusion_model=3)),
    ("3d_lsm2", lambda: vortex_3d_case(), dict(model=2, time_scheme=5, diffusion_model=3,
                                               diffusion_lsm2_option=1, particle_velocity_init=1)),
]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("frozen", [True, False], ids=["frozen", "two_frames"])
@pytest.mark.parametrize("which", range(len(GRAPH_CASES)), ids=[c[0] for c in GRAPH_CASES])
def test_graph_replay(which, frozen, rng_method):
    """Multi-step calls replayed from a captured HIP graph (device-side step
    counters, a period of two re-sorts) give the oracle's results, with frozen
    fields and with two frames interpolated in time (each replayed step's
    record and weight from the host's table)."""
    name, mk, p = GRAPH_CASES[which]
    case = mk()
    params = dict(final_time=300., time_step=5., frozen_eulerian_fields=frozen, rng_method=rng_method, **p)
    s = case.build_gpu(cy, params)
    s.ctx.lib.cyl_set_tuning(s.ctx.h, 3, 0, 0, 1)
    s._run(s._steps_left())
    assert s.ctx.last_graph_steps() >= 2 * 6
    ps = s.particle_set
    g = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                        inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy())
    assert_parity(g, case.run_oracle(R, params))


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("model", [1, 2])
def test_arrivals_midrun(model, rng_method):
    """Rows added mid-run (first-fit into inactive rows) are initialized by the
    next step right before their own update, in pool order, as forward() does
    (solver.pyx:458-467): with LSM-2's initial velocity fluctuations the
    initialization draws interleave with the other rows' step draws."""
    case = vortex_case(npart=1200)
    case.pool = 1800
    p = dict(model=1, diffusion_model=2) if model == 1 else dict(
        model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=2, particle_velocity_init=1)
    params = dict(final_time=300., time_step=5., rng_method=rng_method, **p)
    s = case.build_gpu(cy, params)
    mesh = R.RefMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = R.RefFields(mesh, **{k: v for k, v in case.fields.items() if k != "trifinder"})
    ps = R.RefParticles(case.position, case.velocity, dim=2, initial_pool_size=case.pool, **case.pset_kw)
    rng = R.RefRng(*SEED)
    o = R.RefSolver(fs, ps, params, trifinder=case.trifinder(), rng=rng)
    rs = np.random.default_rng(9)
    for k in range(3):
        for _ in range(5):
            s.forward()
        o.forward(5)
        new = np.stack([rs.uniform(100., 1900., 150), rs.uniform(50., 450., 150)], 1)
        s.particle_set.add(new)
        o.add(new)
    for _ in range(5):
        s.forward()
    o.forward(5)
    g = s.particle_set
    out = SimpleNamespace(position=g.position.copy(), velocity=g.velocity.copy(), tri=g.tri.copy(),
                          inactive=g.inactive.copy(), lowerlayer=g.lowerlayer.copy(), tag=g.tag.copy())
    ref = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                          inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy(),
                          rng=rng)
    assert_parity(out, ref)


CSEC_CHANNEL = [np.array([[50.3, -1.], [50.3, 101.]]),      # across the channel
                np.array([[-1., -1.], [1., 1.]]),            # through the origin: rows deleted in the step
                np.array([[20., 10.], [80., 90.], [0., 0.]])]  # diagonal (its 2nd segment is not tested)


@pytest.mark.parametrize("rng_method", [1, 2])
@pytest.mark.parametrize("sort_interval", [0, 3])
def test_control_sections_channel(rng_method, sort_interval, tmp_path):
    """Per-step control-section counts (solver.pyx:540-549) equal the oracle's:
    diffusing particles crossing a line, a diagonal, and a section through the
    origin, where rows deleted at the open boundary count with their zeroed
    positions.  The counts reach the statistics files unchanged."""
    case = open_channel_case()
    params = dict(final_time=70., time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=1,
                  horizontal_diffusivity=2., rng_method=rng_method, bnd_statistics=True,
                  output_rep=str(tmp_path))
    s = case.build_gpu(cy, params, control_sections=[cy.ControlSection(q) for q in CSEC_CHANNEL])
    s.ctx.lib.cyl_set_tuning(s.ctx.h, sort_interval, 0, 0, 1)
    s.solve()
    assert s.ctx.last_graph_steps() == 0  # per-step counts: eager steps
    r = case.run_oracle(R, params, control_sections=CSEC_CHANNEL)
    times, counts = s.csec_stats
    ref = np.array(r.solver.csec_history)
    assert counts.shape == ref.shape == (70, 3)
    np.testing.assert_array_equal(counts, ref)
    assert ref[:, 0].sum() > 100 and ref[:, 1].sum() > 20 and ref[:, 2].sum() > 20
    # every open-boundary exit is a deletion (walls delete too, after max_wall_rebounds)
    assert np.all(ref[:, 1] >= np.array(r.solver.bnd_history).sum(1))
    for k in range(3):
        d = np.loadtxt(tmp_path / "cs_stats_{}.txt".format(k + 1), delimiter=',', skiprows=2, ndmin=2)
        np.testing.assert_array_equal(d[:, 1].astype(int), counts[:, k])
        np.testing.assert_allclose(d[:, 0], times, rtol=0, atol=1e-8)


@pytest.mark.parametrize("model", [1, 2])
def test_control_sections_vortex(model):
    """Counts on the fused 2D kernels (LSM-1 and LSM-2, k-epsilon diffusion)
    with mid-run arrivals, step by step against the oracle."""
    case = vortex_case(npart=1500)
    case.pool = 1700
    p = dict(model=1, diffusion_model=2) if model == 1 else dict(
        model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1)
    params = dict(final_time=200., time_step=5., **p)
    secs = [np.array([[1000., -10.], [1000., 510.]]), np.array([[0., 250.], [2000., 250.]])]
    s = case.build_gpu(cy, params, control_sections=secs)
    mesh = R.RefMesh(case.x, case.y, case.tri, boundary_points=case.boundary_points, nopen=case.nopen)
    fs = R.RefFields(mesh, **{k: v for k, v in case.fields.items() if k != "trifinder"})
    ps = R.RefParticles(case.position, case.velocity, dim=2, initial_pool_size=case.pool, **case.pset_kw)
    rng = R.RefRng(*SEED)
    o = R.RefSolver(fs, ps, params, trifinder=case.trifinder(), rng=rng, control_sections=secs)
    for _ in range(10):
        s.forward()
    o.forward(10)
    new = np.stack(np.meshgrid(np.linspace(900., 1100., 10), np.linspace(200., 300., 10)), -1).reshape(-1, 2)
    s.particle_set.add(new)
    o.add(new)
    for _ in range(20):
        s.forward()
    o.forward(20)
    assert s.ctx.last_path()[0] > 0  # the fused kernel ran
    counts = s.csec_stats[1]
    ref = np.array(o.csec_history)
    assert ref.sum() > 100
    np.testing.assert_array_equal(counts, ref)


@pytest.mark.parametrize("scheme", [1, 4])
def test_stranded_positions_output(scheme, tmp_path):
    """stranding_output (compute_bc.pyx:107-122, write_particles.pyx:339-388):
    flow toward the top wall with stranding probability 1, so every wall hit
    strands whatever the draw; the recorded positions, in the reference's
    order, equal the oracle's, and the flush writes them to the file."""
    case = open_channel_case(u=0.3)
    top, bot = np.nonzero(np.isclose(case.y, 100.))[0], np.nonzero(np.isclose(case.y, 0.))[0]
    walls = np.concatenate([top, bot])
    case.boundary_points = np.concatenate([case.boundary_points, np.stack([walls, 0 * walls], 1)]).astype(np.int32)
    case.fields = dict(case.fields, velocity_y=np.full_like(case.fields["velocity_x"], 0.8))
    params = dict(final_time=90., time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=0,
                  time_scheme=scheme, stranding_probability=1.0, stranding_output=True, output_rep=str(tmp_path))
    s = case.build_gpu(cy, params)
    for _ in range(30):  # forward() steps, then a batched solve of the rest
        s.forward()
    s.solve()
    r = case.run_oracle(R, params)
    ps = s.particle_set
    np.testing.assert_array_equal(ps.inactive, r.inactive)
    np.testing.assert_array_equal(ps.tag, r.tag)
    nr = r.solver.p.s.stranded_count
    assert nr > 200
    ref = r.solver.p.stranded[:nr]
    got = np.loadtxt(tmp_path / "stranded_particles.txt", delimiter=',', skiprows=2, ndmin=2)
    assert got.shape == ref.shape
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.)) <= TOL
    assert ps.stranded_count == 0  # reset by the flush


@pytest.mark.parametrize("model", [1, 2])
def test_graph_replay_crosses_field_records(model):
    """Replayed steps whose times cross several field records (61 frames,
    one every 167 s, over 600 s of steps)."""
    case = vortex_case(npart=2000, ntimes=61)
    p = dict(model=1, diffusion_model=2) if model == 1 else dict(
        model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1, particle_velocity_init=1)
    params = dict(final_time=600., time_step=5., **p)
    s = case.build_gpu(cy, params)
    s.ctx.lib.cyl_set_tuning(s.ctx.h, 4, 0, 0, 1)
    s._run(s._steps_left())
    assert s.ctx.last_graph_steps() >= 100
    ps = s.particle_set
    g = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                        inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy())
    assert_parity(g, case.run_oracle(R, params))
