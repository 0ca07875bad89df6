"""Settling laws and fluid acceleration in the oracle, pinned by what the
physics fixes rather than by the formulas themselves.

* buoyancy_velocity (lagrangian_particle_set.pyx:834-890): each analytic
  branch is the terminal velocity of a force balance, buoyant weight =
  drag, (s - 1) g (pi d^3 / 6) = Cd(Re) (pi d^2 / 4) w^2 / 2, for the drag law
  the branch stands for: Stokes' Cd = 24 / Re below 1e-4 m, a constant Cd
  above 1e-3 m, and for Ruby-Zanke the law Cd = 80 / (3 Re) + 4 / 3 that the
  formula solves exactly.  The balance is solved here by bracketing root
  search, so a dropped term, a wrong power or constant fails it.
* drag_coefficient (:790-832): model 0 constant; models 1 (Schiller-Naumann)
  and 2 (Almedeij) reduce to Stokes' 24 / Re at vanishing Re and reach the
  Newton plateau of the standard drag curve; model 2 falls through the drag
  crisis.
* An LSM-2 particle released at rest in still water reaches Stokes' settling
  velocity (update_state_lsm2.pyx:1329-1543 with the gravity term -g A0 and
  the drag relaxation rate of model 1).
* _interpolate_acceleration_2d/3d (eulerian_field_set.pyx:1396-1579) on the
  closed forms of the reference's tests/test_17_core_acceleration.py:56-88.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.optimize import brentq

from oracle import refpy as R

GRAV, NU0 = 9.80665, 1.3e-6


def pset(**kw):
    p = R.RefParticles(np.zeros((1, 3)), dim=3, **kw)
    p.s.rho0 = 1000.
    return p


def wb(p, d, model_hr=0):
    return R.lib.ref_buoyancy_velocity(C.byref(p.s), d, model_hr, 1000.)


def terminal_velocity(s, d, cd):
    """Root w > 0 of (s-1) g (pi d^3/6) = cd(Re) (pi d^2/4) w^2 / 2, Re = w d / nu."""
    f = lambda w: abs(s - 1.) * GRAV * d * 4. / 3. - cd(w * d / NU0) * w * w
    return brentq(f, 1e-14, 100., xtol=1e-300, rtol=1e-15, maxiter=500)


STOKES = lambda re: 24. / re
ZANKE = lambda re: 80. / (3. * re) + 4. / 3.


@pytest.mark.parametrize("rho_p", [2650., 1200., 1010.])
@pytest.mark.parametrize("d", [1e-6, 7e-6, 3e-5, 9.99e-5])
def test_buoyancy_stokes_branch_is_stokes_terminal_velocity(rho_p, d):
    p = pset(particle_density=rho_p, buoyancy_velocity_model=1)
    w = terminal_velocity(rho_p / 1000., d, STOKES)
    assert wb(p, d) == pytest.approx(-w, rel=1e-12)


@pytest.mark.parametrize("rho_p", [2650., 1500.])
@pytest.mark.parametrize("d", [1e-4, 2.5e-4, 6e-4, 1e-3])
def test_buoyancy_ruby_zanke_branch_solves_its_drag_law(rho_p, d):
    """1e-4 <= d <= 1e-3, both ends included (the branch tests are < and >)."""
    p = pset(particle_density=rho_p, buoyancy_velocity_model=1)
    w = terminal_velocity(rho_p / 1000., d, ZANKE)
    assert wb(p, d) == pytest.approx(-w, rel=1e-9)


def test_buoyancy_ruby_zanke_limits():
    """Ruby-Zanke lies below both of its asymptotes, (s-1) g d^2 / (20 nu) for
    small d and sqrt((s-1) g d) for large d; at d = 1e-4 it takes over from
    Stokes' law, whose asymptote is larger by the factor 20 / 18."""
    s = 2.65
    p = pset(particle_density=1000. * s, buoyancy_velocity_model=1)
    lo = -wb(p, 1e-4) / ((s - 1.) * GRAV * 1e-8 / (20. * NU0))
    hi = -wb(p, 1e-3) / np.sqrt((s - 1.) * GRAV * 1e-3)
    assert 0.5 < lo < 1.0       # Zanke's law is below its Stokes-like asymptote
    assert 0.3 < hi < 1.0       # and below its Newton asymptote
    # Stokes (1/18) and Ruby-Zanke (1/20) asymptotes differ by the factor 18/20
    below = -wb(p, 1e-4 * (1. - 1e-12))
    assert below / -wb(p, 1e-4) > 20. / 18. * 0.99


@pytest.mark.parametrize("cd", [0.44, 1.2])
@pytest.mark.parametrize("d", [1.0001e-3, 4e-3, 2e-2])
def test_buoyancy_constant_drag_branch(cd, d):
    p = pset(particle_density=2650., buoyancy_velocity_model=1, drag_coefficient=cd)
    w = terminal_velocity(2.65, d, lambda re: cd)
    assert wb(p, d, 0) == pytest.approx(-w, rel=1e-12)


def test_buoyancy_van_rijn_branch():
    """Van Rijn's coarse-sediment law is 1.1 times the Newton asymptote of
    Ruby-Zanke, sqrt((s-1) g d); reached with model_hr = 1 only."""
    s = 2.65
    p = pset(particle_density=1000. * s, buoyancy_velocity_model=1)
    for d in (1.5e-3, 5e-3, 3e-2):
        zanke_far = terminal_velocity(s, d * 1e4, ZANKE) / np.sqrt(1e4)   # Newton regime of the law
        assert -wb(p, d, 1) == pytest.approx(1.1 * zanke_far, rel=2e-2)
        assert -wb(p, d, 1) == pytest.approx(1.1 * np.sqrt((s - 1.) * GRAV * d), rel=1e-14)


def test_buoyancy_sign_and_constant_model():
    light = pset(particle_density=900., buoyancy_velocity_model=1)
    heavy = pset(particle_density=1100., buoyancy_velocity_model=1)
    for d in (5e-5, 5e-4, 5e-3):
        assert wb(light, d) > 0. > wb(heavy, d)
    neutral = pset(particle_density=1000., buoyancy_velocity_model=1)
    assert wb(neutral, 5e-4) == 0.
    const = pset(particle_density=2650., buoyancy_velocity_model=0, buoyancy_velocity=-3.5e-3)
    assert wb(const, 5e-4) == -3.5e-3 and wb(const, 5e-2, 1) == -3.5e-3


def drag(model, re, cd0=0.44, dp=1e-3):
    p = pset(drag_coefficient_model=model, drag_coefficient=cd0)
    return R.lib.ref_drag_coefficient(C.byref(p.s), re * NU0 / dp, dp)


def test_drag_constant_model():
    for re in (1e-3, 1., 1e3, 1e6):
        assert drag(0, re, 0.7) == 0.7


@pytest.mark.parametrize("model", [1, 2])
def test_drag_stokes_limit(model):
    for re in (1e-6, 1e-4):
        assert drag(model, re) * re / 24. == pytest.approx(1., rel=0.2 * re ** 0.65)


def test_drag_schiller_naumann_plateau_and_continuity():
    assert drag(1, 1e3 * (1 + 1e-9)) == 0.44 and drag(1, 5e4) == 0.44
    # the correlation meets the plateau at Re = 1000 to within 1 %
    assert drag(1, 1e3) == pytest.approx(0.44, rel=1e-2)
    re = np.logspace(-3, 3, 40)
    cd = np.array([drag(1, r) for r in re])
    assert np.all(np.diff(cd) < 0.)


def test_drag_almedeij_standard_curve():
    re = np.logspace(-3, 5, 60)
    cd = np.array([drag(2, r) for r in re])
    assert np.all(np.diff(cd[re <= 2e3]) < 0.)
    plateau = cd[(re >= 2e3) & (re <= 1e5)]
    assert plateau.min() > 0.38 and plateau.max() < 0.5
    # drag crisis: past Re ~ 3e5 the coefficient falls to about 0.2
    assert drag(2, 1e6) < 0.25 < 0.45 < drag(2, 1e5)
    # the intermediate range stays within 15 % of Schiller-Naumann, the other fit of the curve
    mid = (re >= 1.) & (re <= 1e3)
    sn = np.array([drag(1, r) for r in re[mid]])
    assert np.max(np.abs(cd[mid] / sn - 1.)) < 0.15


def still_water_3d(nlayers=3, h=10.):
    x, y = np.array([0., 100., 100., 0.]), np.array([0., 0., 100., 100.])
    tri = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    z = np.concatenate([np.full(4, -h + k * h / (nlayers - 1)) for k in range(nlayers)])
    zeros = np.zeros((1, 4 * nlayers))
    return x, y, tri, dict(dim=3, nlayers=nlayers, times=np.zeros(1), elevation=z[None], velocity_x=zeros,
                           velocity_y=zeros, velocity_z=zeros)


@pytest.mark.parametrize("d", [1e-5, 2e-5])
def test_lsm2_heavy_particle_reaches_stokes_settling_velocity(d):
    """Direct integrator, drag model 1, still water: the vertical particle
    velocity relaxes to -(s-1) g d^2 / (18 nu), Stokes' law (Re below 5e-3
    here, so Schiller-Naumann's correction is below 4e-3)."""
    x, y, tri, fields = still_water_3d()
    mesh = R.RefMesh(x, y, tri)
    fs = R.RefFields(mesh, **fields)
    ps = R.RefParticles(np.array([[50., 30., -1.], [30., 50., -2.]]), dim=3, particle_density=2650.,
                        particle_diameter=d, drag_coefficient_model=1)
    params = dict(model=2, time_scheme=5, diffusion_model=0, frozen_eulerian_fields=True, time_step=0.01,
                  final_time=1., particle_velocity_init=1)
    s = R.RefSolver(fs, ps, params, rng=R.RefRng())
    s.forward(100)
    w_stokes = (2.65 - 1.) * GRAV * d * d / (18. * NU0)
    re = w_stokes * d / NU0
    assert re < 5e-3
    np.testing.assert_allclose(ps.velocity[:, 2], -w_stokes, rtol=0.2 * re ** 0.687)
    np.testing.assert_allclose(ps.velocity[:, :2], 0., atol=1e-15)


def accel_fields(times, ux, uy):
    xn, yn = np.array([-100., 100., 100., -100.]), np.array([-100., -100., 100., 100.])
    mesh = R.RefMesh(xn, yn, np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    shape = (len(times), 4)
    fs = R.RefFields(mesh, dim=2, nlayers=1, times=np.array(times), elevation=np.full(shape, 100.),
                     velocity_x=np.broadcast_to(ux, shape).copy(), velocity_y=np.broadcast_to(uy, shape).copy())
    return mesh, fs, xn, yn


def accel_2d(mesh, fs, x, y, dt, ua, va):
    ax, ay = C.c_double(), C.c_double()
    k = mesh.brute(x, y)
    e = R.lib.ref_interp_acceleration_2d(C.byref(fs.s), C.byref(mesh.s), k, x, y, dt, ua, va, C.byref(ax),
                                         C.byref(ay))
    assert e == 0
    return ax.value, ay.value


def test_acceleration_2d_convective_term():
    """tests/test_17_core_acceleration.py:56-64"""
    xn, yn = np.array([-100., 100., 100., -100.]), np.array([-100., -100., 100., 100.])
    mesh, fs, _, _ = accel_fields((0.,), 1. + 0.1 * xn, 0.3 + 0.2 * yn)
    ax, ay = accel_2d(mesh, fs, 10., 20., 0.1, 0.7, -0.4)
    assert ax == pytest.approx(0.1 * 0.7, rel=1e-12)
    assert ay == pytest.approx(0.2 * -0.4, rel=1e-12)


def test_acceleration_2d_time_derivative():
    """tests/test_17_core_acceleration.py:67-79: u = 1 + 0.1 x + 0.5 t, so the
    material derivative is 0.5 + 0.1 u_a once two interpolations were made."""
    xn = np.array([-100., 100., 100., -100.])
    times = (0., 1., 2.)
    ux = np.array([1. + 0.1 * xn + 0.5 * t for t in times])
    mesh, fs, _, _ = accel_fields(times, ux, np.zeros_like(ux))
    fs.s.frozen = 0
    fs.s.record = R.lib.ref_lower_index(R.P(fs.times, "d"), 3, 1.1, 0)
    assert R.lib.ref_time_interpolation(C.byref(fs.s), 1.1) == 0
    assert R.lib.ref_time_interpolation(C.byref(fs.s), 1.2) == 0
    assert fs.s.record == 1
    ax, ay = accel_2d(mesh, fs, 10., 20., 0.1, 0.7, 0.)
    assert ax == pytest.approx(0.5 + 0.1 * 0.7, rel=1e-9)
    assert ay == 0.


@pytest.mark.parametrize("slope", [0., 0.1])
def test_acceleration_3d_convective_term(slope):
    """tests/test_17_core_acceleration.py:82-95: u = a x + b z, v = c y + d z
    on two planes with sloping bed and surface."""
    a, b, c, d = 0.1, 0.02, 0.05, 0.03
    xn, yn = np.array([-100., 100., 100., -100.]), np.array([-100., -100., 100., 100.])
    mesh = R.RefMesh(xn, yn, np.array([[0, 1, 2], [0, 2, 3]], np.int32))
    z = np.concatenate([-100. + slope * xn, 100. + slope * xn])
    x2, y2 = np.tile(xn, 2), np.tile(yn, 2)
    fs = R.RefFields(mesh, dim=3, nlayers=2, times=np.zeros(1), elevation=z[None],
                     velocity_x=(a * x2 + b * z)[None], velocity_y=(c * y2 + d * z)[None],
                     velocity_z=np.zeros((1, 8)))
    x, y = 10., 20.
    zp = 5. + slope * x
    k = mesh.brute(x, y)
    out = [C.c_double() for _ in range(3)]
    e = R.lib.ref_interp_acceleration_3d(C.byref(fs.s), C.byref(mesh.s), k, 0, x, y, zp, 0.1, 0.7, -0.4, 0.2,
                                         *[C.byref(o) for o in out])
    assert e == 0
    ax, ay, az = (o.value for o in out)
    assert ax == pytest.approx(a * 0.7 + b * 0.2, rel=1e-10)
    assert ay == pytest.approx(c * -0.4 + d * 0.2, rel=1e-10)
    assert az == 0.
