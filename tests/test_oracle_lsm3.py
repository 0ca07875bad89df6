"""The oracle's LSM-3 against matrix exponentials, the reference's one-step
closed forms and its golden vectors (tests/test_40, test_42, test_46)."""
import numpy as np
import pytest
from scipy.linalg import expm

from oracle import refpy as R
from tests.cases import Case, reference_square_case, reference_square_p3d_case


def drift(a1, tl, a2):
    return np.array([[-1. / tl, 0., 0.], [a1 - a2 / tl, -a1, 0.], [0., 1., 0.]])


def van_loan(a1, tl, a2, dt):
    """Covariance of the noise (1, a2, 0) dW over dt (Van Loan's block exponential)."""
    A = drift(a1, tl, a2)
    g = np.array([1., a2, 0.])
    M = np.zeros((6, 6))
    M[:3, :3] = A
    M[:3, 3:] = np.outer(g, g)
    M[3:, 3:] = -A.T
    E = expm(dt * M)
    return E[:3, 3:] @ E[:3, :3].T


def force_response(a1, tl, a2, dt):
    M = np.zeros((4, 4))
    M[:3, :3] = drift(a1, tl, a2)
    M[1, 3] = 1.
    return expm(dt * M)[:3, 3]


# -- coefficients (tests/test_46_lsm3_coupling_regressions.py:84-180) -----------------
def test_physical_transition_matches_expm():
    a1, tl, a2, dt = 0.3, 1., 0.5, 1e-18
    tr, cov, resp = R.lsm3_direct_coefficients(a1, tl, dt, a2)
    np.testing.assert_allclose(tr, expm(dt * drift(a1, tl, a2)), rtol=3e-13, atol=0.)
    np.testing.assert_allclose(resp, force_response(a1, tl, a2, dt), rtol=3e-13, atol=0.)
    np.testing.assert_array_equal(cov, cov.T)


def test_zero_physical_drift_coupling():
    tr, _, _ = R.lsm3_direct_coefficients(2., 0.25, 1e-6, 0.5)
    assert tr[1, 0] == 0. and tr[2, 0] == 0.


def test_covariance_matches_van_loan():
    _, cov, _ = R.lsm3_direct_coefficients(0.3, 1., 0.2, 0.5)
    np.testing.assert_allclose(cov, van_loan(0.3, 1., 0.5, 0.2), rtol=3e-13, atol=3e-16)


@pytest.mark.parametrize("a1,tl,a2,dt", [(1., 1., 0.5, 0.3), (50., 1e-2, 2.5, 1e-3)])
def test_clipped_option_matches_reference(a1, tl, a2, dt):
    tr, cov, resp = R.lsm3_direct_coefficients(a1, tl, dt, a2, 2)
    np.testing.assert_allclose(tr, expm(dt * drift(a1, tl, a2)), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(resp, force_response(a1, tl, a2, dt), rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(cov, van_loan(a1, tl, a2, dt), rtol=5e-6, atol=1e-12 * np.abs(cov).max())


def test_clipped_option_is_finite_and_positive():
    tr, cov, resp = R.lsm3_direct_coefficients(1e8, 1e-8, 1e-6, 0.5, 2)
    assert np.isfinite(tr).all() and np.isfinite(cov).all() and np.isfinite(resp).all()
    ev = np.linalg.eigvalsh(cov)
    assert ev.min() >= -1e-10 * max(ev.max(), 1e-300)


def test_invalid_inputs_raise():
    with pytest.raises(ValueError):
        R.lsm3_direct_coefficients(1., 1., 0.1, 0., 3)
    with pytest.raises(ValueError):
        R.lsm3_direct_coefficients(-1., 1., 0.1, 0.)
    with pytest.raises(ValueError):
        R.lsm3_direct_coefficients(1., 0., 0.1, 0., 2)


@pytest.mark.parametrize("edge_opt", [1, 2])
def test_coefficient_sweep_matches_exponentials(edge_opt):
    """Rates from 1e-3 to 20 per step, both relaxation orders, with and without
    added mass: the series (doublings) and closed forms against expm and Van Loan."""
    rng = np.random.default_rng(11)
    for _ in range(60):
        dt = 10 ** rng.uniform(-3, 0)
        a1 = 10 ** rng.uniform(-3, 1.3) / dt
        tl = dt / 10 ** rng.uniform(-3, 1.3)
        a2 = rng.choice([0., 0.5, 1., 1.7])
        if edge_opt == 2 and abs(a1 - 1. / tl) < 0.05 * (a1 + 1. / tl):
            continue  # clipped near a1 = 1/tl: covered above at its own tolerance
        tr, cov, resp = R.lsm3_direct_coefficients(a1, tl, dt, a2, edge_opt)
        np.testing.assert_allclose(tr, expm(dt * drift(a1, tl, a2)), rtol=1e-9, atol=1e-13 * dt)
        np.testing.assert_allclose(resp, force_response(a1, tl, a2, dt), rtol=1e-9, atol=1e-15 * dt * dt)
        ref = van_loan(a1, tl, a2, dt)
        ref = 0.5 * (ref + ref.T)  # expm of the block matrix is itself good to about 1e-7 here
        np.testing.assert_allclose(cov, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())


# -- one-step closed forms on the reference's two-triangle square ----------------------
def small_square(dim, field_dim, scheme, density=1000., added_mass=True, diffusivity=0.1, diffusion_model=1,
                 edge_opt=2):
    """tests/test_46_lsm3_coupling_regressions.py:9-47"""
    x = np.array([-100., 100., 100., -100.])
    y = np.array([-100., -100., 100., 100.])
    tri = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int32)
    nl = 2 if field_dim == 3 else 1
    shape = (1, 4 * nl)
    elev = np.array([[-100.] * 4 + [100.] * 4]) if field_dim == 3 else np.full(shape, 100.)
    opt = np.empty((3,) + shape)
    opt[0], opt[1], opt[2] = -100., (0.5 + 0.75 * 1.2) * 0.7 * 0.2, 0.2
    fields = dict(dim=field_dim, nlayers=nl, times=np.zeros(1), elevation=elev, velocity_x=np.zeros(shape),
                  velocity_y=np.zeros(shape), opt_fields=opt,
                  opt_fields_names=['BOTTOM', 'TURBULENT ENERG.', 'DISSIPATION'])
    if field_dim == 3:
        fields["velocity_z"] = np.zeros(shape)
    case = Case(x, y, tri, fields, np.zeros((4, dim)), pool=4, pdim=dim,
                pset_kw=dict(particle_density=density, particle_diameter=0.001, drag_coefficient_model=1,
                             added_mass_force=added_mass))
    params = dict(model=3, time_scheme=scheme, time_step=0.005, final_time=0.02, frozen_eulerian_fields=True,
                  particle_velocity_init=0, diffusion_model=diffusion_model, horizontal_diffusivity=diffusivity,
                  vertical_diffusivity=0.5 * diffusivity, diffusion_lsm3_tl_horizontal=0.7,
                  diffusion_lsm3_tl_vertical=1.3, rng_method=2, water_density=1000., boundary_conditions=False,
                  lsm3_edge_opt=edge_opt)
    return case, params


def oracle_solver(case, params, seed=(42, 7)):
    mesh = R.RefMesh(case.x, case.y, case.tri)
    fs = R.RefFields(mesh, **case.fields)
    ps = R.RefParticles(case.position, dim=case.pdim, initial_pool_size=case.pool, **case.pset_kw)
    return R.RefSolver(fs, ps, params, trifinder=case.trifinder(), rng=R.RefRng(*seed)), ps


def test_em_uses_initial_seen_velocity():
    """tests/test_46_lsm3_coupling_regressions.py:49-69: Up(t+dt) = Up + a1 (Us - Up) dt
    + A2 (Us(t+dt) - Us), Xp(t+dt) = Xp + Up(t+dt) dt."""
    case, params = small_square(2, 2, 1)
    s, ps = oracle_solver(case, params)
    ps.velocity[:4] = -0.1
    ps.us[:4] = 0.2
    xp0, up0, us0 = ps.position.copy(), ps.velocity.copy(), ps.us.copy()
    slip = np.linalg.norm(us0[0] - up0[0])
    a1 = ps.s.a1c * 1.5 / 0.001 * slip * R.lib.ref_drag_coefficient(ps.s, slip, 0.001)
    dt = params["time_step"]
    s.forward(1)
    assert ps.npart_active == 4
    expected = up0 + a1 * (us0 - up0) * dt + ps.s.a2c * (ps.us - us0)
    np.testing.assert_allclose(ps.velocity, expected, rtol=3e-14, atol=3e-16)
    np.testing.assert_allclose(ps.position, xp0 + expected * dt, rtol=3e-14, atol=3e-16)
    assert np.abs(ps.us - us0).max() > 0.  # the seen velocity moved


@pytest.mark.parametrize("dim,field_dim,dm,edge_opt", [(2, 2, 1, 1), (2, 2, 1, 2), (3, 2, 3, 2), (3, 3, 2, 1)])
def test_neutral_particles_follow_seen_velocity(dim, field_dim, dm, edge_opt):
    """tests/test_46_lsm3_coupling_regressions.py:72-90, 184-196: with A2 = 1 (neutral
    particle, added mass) the direct integrator keeps Up = Us."""
    case, params = small_square(dim, field_dim, 5, diffusion_model=dm, edge_opt=edge_opt)
    s, ps = oracle_solver(case, params)
    assert ps.s.a2c == 1.
    ps.velocity[:4] = 0.1
    ps.us[:4] = 0.1
    for _ in range(4):
        s.forward(1)
        assert ps.npart_active == 4 and np.isfinite(ps.position).all()
        if edge_opt == 1:
            np.testing.assert_array_equal(ps.velocity, ps.us)
        else:
            np.testing.assert_allclose(ps.velocity, ps.us, rtol=1e-12, atol=1e-14)
    assert np.abs(ps.us - 0.1).max() > 1e-6  # the seen velocity diffused


# -- golden vectors ---------------------------------------------------------------------
P4 = np.array([[5000., 0.], [5000., 500.], [1000., 0.], [1000., 500.]])


def test_lsm3_2d_direct_golden():
    """tests/test_40_lsm3_analytical_advection_2d.py:76-109 (k = eps = 1e-8)."""
    case = reference_square_case(P4, pool=1000)
    n = case.x.size
    case.fields.update(opt_fields=1e-8 * np.ones((2, 1, n)), opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    params = dict(final_time=3600., time_step=1., time_scheme=5, frozen_eulerian_fields=True, model=3,
                  particle_velocity_init=0, diffusion_model=2)
    out = case.run_oracle(R, params)
    xpa = np.array([[5.07977645e+03, 1.08730031e+01], [5.07860134e+03, 5.18836293e+02],
                    [1.02296934e+03, 2.34944380e+00], [1.02123508e+03, 5.13524275e+02]])
    np.testing.assert_allclose(out.position[out.inactive == 0], xpa, rtol=1e-2)


def test_lsm3_p3d_direct_golden():
    """tests/test_42_lsm3_analytical_advection_p3d.py:75-117 (zero k and eps,
    velocity initialization with the seen-velocity fluctuations)."""
    case = reference_square_p3d_case()
    n = case.x.size
    case.fields.update(opt_fields=np.zeros((3, 1, n)),
                       opt_fields_names=['BOTTOM', 'TURBULENT ENERG.', 'DISSIPATION'])
    params = dict(final_time=3600., time_step=1., time_scheme=5, frozen_eulerian_fields=True, model=3,
                  particle_velocity_init=1, diffusion_model=2)
    out = case.run_oracle(R, params)
    xpa = np.array([[4.75450561e+03, -1.83294844e+03, 5.00000000e-01],
                    [2.33409108e+03, -4.58085258e+03, 9.99900000e-01],
                    [9.57929441e+02, -3.69170032e+02, 5.00000000e-01],
                    [8.55082445e+02, -7.73654172e+02, 9.99900000e-01]])
    np.testing.assert_allclose(out.position[out.inactive == 0], xpa, rtol=1e-7)
