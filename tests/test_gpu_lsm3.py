"""LSM-3 on the device (generic step kernel) against the oracle: Euler-Maruyama
and direct integrators, both coefficient options, constant-K and k-epsilon
seen-velocity models, added mass, 2D, 3D and pseudo-3D, arrival fluctuations,
wall rebounds that reflect the seen velocity, and the reference's golden
vectors (tests/test_40, test_42, test_46)."""
import numpy as np
import pytest

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import reference_square_case, reference_square_p3d_case, vortex_3d_case, vortex_case
from tests.test_gpu_batch import streams as batch_streams
from tests.test_gpu_parity import TOL, assert_parity
from tests.test_oracle_lsm3 import P4, small_square

pytestmark = pytest.mark.gpu

# The explicit scheme needs a1 dt < 1: 5 mm grains and a 0.02 s step
EM = dict(time_step=0.02, final_time=3.)
EM_GRAIN = dict(particle_diameter=5e-3, particle_density=1100.)
L3_2D = [
    dict(time_scheme=5, diffusion_model=2, lsm3_edge_opt=2),
    dict(time_scheme=5, diffusion_model=2, lsm3_edge_opt=1),
    dict(time_scheme=5, diffusion_model=1, horizontal_diffusivity=0.5, diffusion_lsm3_tl_horizontal=40.),
    dict(time_scheme=1, diffusion_model=2, **EM),
    dict(time_scheme=1, diffusion_model=1, horizontal_diffusivity=0.5, diffusion_lsm3_tl_horizontal=40., **EM),
]
L3_2D_IDS = ["di_keps_clipped", "di_keps_series", "di_constK", "em_keps", "em_constK"]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("which", range(len(L3_2D)), ids=L3_2D_IDS)
def test_lsm3_2d_vortex(which, rng_method):
    """Two frames (the mean-flow acceleration includes dU/dt), drag of sediment
    grains, arrival fluctuations of the seen velocity, damped wall rebounds."""
    case = vortex_case(npart=2000)
    em = L3_2D[which]["time_scheme"] == 1
    case.pset_kw = EM_GRAIN if em else dict(particle_diameter=2e-4, particle_density=1800.)
    params = dict(dict(final_time=200., time_step=5.), model=3, particle_velocity_init=1, rng_method=rng_method,
                  rebound_damping_coef=0.3)
    params.update(L3_2D[which])
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    assert_parity(g, r)
    us_err = np.max(np.abs(g.solver.particle_set.fluid_velocity_seen - r.solver.p.us) /
                    np.maximum(np.abs(r.solver.p.us), 1.))
    assert us_err <= TOL, us_err
    f, gen = g.solver.ctx.last_path()
    assert f == 0 and gen > 0  # LSM-3 runs the generic kernel


@pytest.mark.parametrize("added_mass", [False, True])
def test_lsm3_added_mass_neutral(added_mass):
    """Neutral particles with added mass keep Up = Us under the direct
    integrator (tests/test_46_lsm3_coupling_regressions.py:72-90)."""
    case, params = small_square(2, 2, 5, diffusion_model=1, edge_opt=1, added_mass=added_mass)
    case.velocity = np.full((4, 2), 0.1)
    case.pset_kw = dict(case.pset_kw, fluid_velocity_seen=np.full((4, 2), 0.1))
    s = case.build_gpu(cy, params)
    for _ in range(4):
        s.forward()
    ps = s.particle_set
    assert ps.npart_active == 4
    if added_mass:
        np.testing.assert_array_equal(ps.velocity, ps.fluid_velocity_seen)
    assert np.abs(ps.fluid_velocity_seen - 0.1).max() > 1e-6


L3_3D = [
    dict(time_scheme=5, diffusion_model=3),
    dict(time_scheme=5, diffusion_model=2, vertical_diffusivity=1e-3, diffusion_lsm3_tl_vertical=20.),
    dict(time_scheme=5, diffusion_model=2, vertical_diffusivity=1e-3, diffusion_lsm3_tl_vertical=20.,
         lsm3_edge_opt=1),
    dict(time_scheme=1, diffusion_model=3, **EM),
]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
@pytest.mark.parametrize("which", range(len(L3_3D)), ids=["di_keps3", "di_keps2_clipped", "di_keps2_series",
                                                          "em_keps3"])
def test_lsm3_3d(which, rng_method):
    """Sigma-layer 3D: seen-velocity vertical component with its own TL,
    settling (A0 g), bottom and surface reflections of Up and Us."""
    case = vortex_3d_case()
    em = L3_3D[which]["time_scheme"] == 1
    case.pset_kw = EM_GRAIN if em else dict(particle_diameter=1e-4, particle_density=1500.)
    params = dict(final_time=300., time_step=5., model=3, particle_velocity_init=1, rng_method=rng_method)
    params.update(L3_3D[which])
    assert_parity(case.run_gpu(cy, params), case.run_oracle(R, params))


def test_lsm3_2d_direct_golden():
    """tests/test_40_lsm3_analytical_advection_2d.py:76-109."""
    case = reference_square_case(P4, pool=1000)
    n = case.x.size
    case.fields.update(opt_fields=1e-8 * np.ones((2, 1, n)), opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    params = dict(final_time=3600., time_step=1., time_scheme=5, frozen_eulerian_fields=True, model=3,
                  particle_velocity_init=0, diffusion_model=2)
    g = case.run_gpu(cy, params)
    assert_parity(g, case.run_oracle(R, params))
    xpa = np.array([[5.07977645e+03, 1.08730031e+01], [5.07860134e+03, 5.18836293e+02],
                    [1.02296934e+03, 2.34944380e+00], [1.02123508e+03, 5.13524275e+02]])
    np.testing.assert_allclose(g.position[g.inactive == 0], xpa, rtol=1e-2)


def test_lsm3_p3d_direct_golden():
    """tests/test_42_lsm3_analytical_advection_p3d.py:75-117."""
    case = reference_square_p3d_case()
    n = case.x.size
    case.fields.update(opt_fields=np.zeros((3, 1, n)),
                       opt_fields_names=['BOTTOM', 'TURBULENT ENERG.', 'DISSIPATION'])
    params = dict(final_time=3600., time_step=1., time_scheme=5, frozen_eulerian_fields=True, model=3,
                  particle_velocity_init=1, diffusion_model=2)
    g = case.run_gpu(cy, params)
    assert_parity(g, case.run_oracle(R, params))
    xpa = np.array([[4.75450561e+03, -1.83294844e+03, 5.00000000e-01],
                    [2.33409108e+03, -4.58085258e+03, 9.99900000e-01],
                    [9.57929441e+02, -3.69170032e+02, 5.00000000e-01],
                    [8.55082445e+02, -7.73654172e+02, 9.99900000e-01]])
    np.testing.assert_allclose(g.position[g.inactive == 0], xpa, rtol=1e-7)


def test_lsm3_batch_equals_alone():
    """LSM-3 scenarios batched with LSM-1 and LSM-2 ones step as they do alone."""
    case = vortex_case(npart=800)
    mix = [dict(model=3, time_scheme=5, diffusion_model=2, particle_velocity_init=1),
           dict(model=1, diffusion_model=2),
           dict(model=3, time_scheme=5, diffusion_model=1, horizontal_diffusivity=0.3,
                diffusion_lsm3_tl_horizontal=30., particle_velocity_init=1, lsm3_edge_opt=1)]
    base = dict(final_time=1e4, time_step=5.)
    params = [dict(base, **m) for m in mix]
    streams = batch_streams(3)
    fs = cy.EulerianFieldSet(cy.TriangularMesh(case.x, case.y, case.tri), trifinder=case.trifinder(),
                             **case.fields)
    mk = lambda: cy.LagrangianParticleSet(case.position, dim=2, particle_diameter=2e-4, particle_density=1800.)
    b = cy.BatchSolver(fs, [mk() for _ in mix], params, streams, nsteps=[20, 20, 20])
    b.run()
    for s in range(3):
        out = b.collect(s)
        cy.random_utils.set_state(*streams[s])
        fs1 = cy.EulerianFieldSet(cy.TriangularMesh(case.x, case.y, case.tri), trifinder=case.trifinder(),
                                  **case.fields)
        a = cy.Solver(fs1, mk(), params[s])
        for _ in range(20):
            a.forward()
        ps = a.particle_set
        np.testing.assert_array_equal(out.tri, ps.tri)
        scale = np.maximum(np.abs(ps.position), 1.)
        assert np.max(np.abs(out.position - ps.position) / scale) <= TOL
        assert tuple(b.rng_state(s)) == tuple(cy.random_utils.get_state())
        a.ctx.close()
    b.close()
