"""The oracle against the reference's own golden vectors and closed forms."""
import ctypes as C

import numpy as np
import pytest

from oracle import refpy as R
from tests.cases import reference_square_3d_case, reference_square_case, reference_square_p3d_case

P4 = np.array([[5000., 0.], [5000., 500.], [1000., 0.], [1000., 500.]])


@pytest.mark.parametrize("scheme,dt,expected", [
    # tests/test_20_lsm1_analytical_advection_2d.py:104-108, 173-177, 242-246
    (1, 10., [[5.28176591e+03, 3.36909449e+00], [5.28142900e+03, 5.31545686e+02],
              [1.05635318e+03, 6.73818897e-01], [1.05601627e+03, 5.28850410e+02]]),
    (2, 200., [[5130.63522395, -632.91944187], [5193.92716813, -119.85591948],
               [1026.12704479, -126.58388837], [1089.41898898, 386.47963402]]),
    (4, 200., [[4.99888531e+03, 3.71825833e+00], [4.99851348e+03, 5.03606789e+02],
               [9.99777062e+02, 7.43651666e-01], [9.99405236e+02, 5.00632183e+02]]),
])
def test_lsm1_advection_golden(scheme, dt, expected):
    case = reference_square_case(P4, pool=1000)
    params = dict(final_time=3600., time_step=dt, time_scheme=scheme,
                  frozen_eulerian_fields=True, model=1, diffusion_model=0)
    out = case.run_oracle(R, params)
    xp = out.position[out.inactive == 0]
    np.testing.assert_allclose(xp, np.array(expected), rtol=1e-7)


def test_lsm1_euler_maruyama_zero_diffusivity_golden():
    # tests/test_20_lsm1_analytical_advection_2d.py:312-316: K = 0 draws nothing
    case = reference_square_case(P4, pool=1000)
    params = dict(final_time=3600., time_step=10., time_scheme=1, frozen_eulerian_fields=True,
                  model=1, diffusion_model=1, horizontal_diffusivity=0., vertical_diffusivity=0.)
    out = case.run_oracle(R, params)
    np.testing.assert_allclose(out.position[out.inactive == 0], np.array(
        [[5.28176591e+03, 3.36909449e+00], [5.28142900e+03, 5.31545686e+02],
         [1.05635318e+03, 6.73818897e-01], [1.05601627e+03, 5.28850410e+02]]), rtol=1e-7)
    assert out.rng.state() == R.RefRng(42, 0xda3e06d11bb2b203).state()


LSM2_GOLDEN = [
    # tests/test_30_lsm2_analytical_advection_2d.py: (time_scheme, dt, diffusion_model,
    # diffusion_lsm2_option, final_time, expected positions)
    (1, 1., 0, 1, 3600., [[5.02240714e+03, 1.59416032e-02], [5.02231846e+03, 5.02247697e+02],
                          [1.01149457e+03, 3.88291522e-02], [1.01090280e+03, 5.05494818e+02]]),  # :106-110
    (2, 1., 0, 1, 3600., [[5.04996265e+03, 1.42800005e-01], [5.04986094e+03, 5.05129760e+02],
                          [1.01703735e+03, 8.91474852e-02], [1.01641878e+03, 5.08313138e+02]]),  # :177-181
    (4, 1., 0, 1, 3600., [[5.04996154e+03, 1.58905815e-01], [5.04985822e+03, 5.05145755e+02],
                          [1.01703715e+03, 9.23901032e-02], [1.01641697e+03, 5.08316281e+02]]),  # :248-252
    (5, 20., 0, 1, 3600., [[5.58482914e+03, 1.46424824e+01], [5.58333674e+03, 5.73120494e+02],
                           [1.12027542e+03, 3.17155011e+00], [1.11839840e+03, 5.63135678e+02]]),  # :320-324
    (5, 20., 1, 2, 3600., [[5.58482914e+03, 1.46424824e+01], [5.58333674e+03, 5.73120494e+02],
                           [1.12027542e+03, 3.17155011e+00], [1.11839840e+03, 5.63135678e+02]]),  # :389-393
    (1, 0.005, 1, 2, 50., [[4981.651198, -435.835882], [5025.233608, 62.329215],
                           [996.422963, -87.174511], [1040.00244, 411.033803]]),  # :460-464
]


def lsm2_params(ts, dt, dm, opt, tf):
    return dict(final_time=tf, time_step=dt, time_scheme=ts, frozen_eulerian_fields=True, model=2,
                particle_velocity_init=1, diffusion_model=dm, diffusion_lsm2_option=opt,
                horizontal_diffusivity=0., vertical_diffusivity=0.)


@pytest.mark.parametrize("ts,dt,dm,opt,tf,expected", LSM2_GOLDEN,
                         ids=["euler", "rk2", "rk4", "di1", "di_opt2", "euler_maruyama"])
def test_lsm2_advection_golden(ts, dt, dm, opt, tf, expected):
    case = reference_square_case(P4, pool=1000)
    out = case.run_oracle(R, lsm2_params(ts, dt, dm, opt, tf))
    np.testing.assert_allclose(out.position[out.inactive == 0], np.array(expected), rtol=1e-7)


GOLDEN_3D = [
    # (case, model, time_scheme, dt, diffusion_model, diffusion_lsm2_option, expected)
    # tests/test_21_lsm1_analytical_advection_3d.py:126, 200, 274, 349
    ("3d", 1, 1, 10., 0, 1, [[-5.06900817e+03, -4.04228110e-01, 0.5], [5.28103619e+03, 5.34857127e+02, 0.9999],
                             [-1.01380163e+03, -8.08456219e-02, 0.5], [1.05567231e+03, 5.29507814e+02, 0.9999]]),
    ("3d", 1, 2, 200., 0, 1, [[-5.00982318e+03, 7.91738560e+01, 0.5], [5.19393649e+03, -1.16404904e+02, 0.9999],
                              [-1.00196464e+03, 1.58347712e+01, 0.5], [1.08914758e+03, 3.87197910e+02, 0.9999]]),
    ("3d", 1, 4, 200., 0, 1, [[-4.99998240e+03, -1.20144922e-01, 0.5], [4.99819692e+03, 5.06745590e+02, 0.9999],
                              [-9.99996480e+02, -2.40289843e-02, 0.5], [9.99090798e+02, 5.01259730e+02, 0.9999]]),
    ("3d", 1, 1, 10., 1, 1, [[-5.06900817e+03, -4.04228110e-01, 0.5], [5.28103619e+03, 5.34857127e+02, 0.9999],
                             [-1.01380163e+03, -8.08456219e-02, 0.5], [1.05567231e+03, 5.29507814e+02, 0.9999]]),
    # tests/test_31_lsm2_analytical_advection_3d.py:122-128, 206
    ("3d", 2, 5, 10., 0, 1, [[-5.07600851e+03, -5.30137280e-01, 0.5], [5.29211338e+03, 5.36375306e+02, 0.9999],
                             [-1.01721737e+03, -1.44511801e-01, 0.5], [1.06172324e+03, 5.32820087e+02, 0.9999]]),
    ("3d", 2, 5, 10., 1, 2, [[-5.07600851e+03, -5.30137280e-01, 0.5], [5.29211338e+03, 5.36375306e+02, 0.9999],
                             [-1.01721737e+03, -1.44511801e-01, 0.5], [1.06172324e+03, 5.32820087e+02, 0.9999]]),
    # tests/test_22_lsm1_analytical_advection_p3d.py:116
    ("p3d", 1, 1, 10., 0, 1, [[4.96199336e+03, -1.90971737e+03, 0.5], [2.46890626e+03, -4.83763012e+03, 0.9999],
                              [9.92398673e+02, -3.81943473e+02, 0.5], [8.96515570e+02, -8.10286955e+02, 0.9999]]),
    # tests/test_32_lsm2_analytical_advection_p3d.py:118, 191
    ("p3d", 2, 5, 10., 0, 1, [[4.97261675e+03, -1.91332310e+03, 0.5], [2.47449397e+03, -4.84746475e+03, 0.9999],
                              [9.98973944e+02, -3.84159272e+02, 0.5], [9.02721202e+02, -8.15390229e+02, 0.9999]]),
    ("p3d", 2, 5, 10., 1, 2, [[4.97261675e+03, -1.91332310e+03, 0.5], [2.47449397e+03, -4.84746475e+03, 0.9999],
                              [9.98973944e+02, -3.84159272e+02, 0.5], [9.02721202e+02, -8.15390229e+02, 0.9999]]),
]
GOLDEN_3D_IDS = ["3d_lsm1_euler", "3d_lsm1_rk2", "3d_lsm1_rk4", "3d_lsm1_em", "3d_lsm2_di1", "3d_lsm2_di2",
                 "p3d_lsm1_euler", "p3d_lsm2_di1", "p3d_lsm2_di2"]


def golden_3d_setup(kind, model, ts, dt, dm, opt):
    case = reference_square_3d_case() if kind == "3d" else reference_square_p3d_case()
    params = dict(final_time=3600., time_step=dt, time_scheme=ts, frozen_eulerian_fields=True, model=model,
                  diffusion_model=dm, horizontal_diffusivity=0., vertical_diffusivity=0.)
    if model == 2:
        params.update(particle_velocity_init=1, diffusion_lsm2_option=opt)
    return case, params


@pytest.mark.parametrize("kind,model,ts,dt,dm,opt,expected", GOLDEN_3D, ids=GOLDEN_3D_IDS)
def test_3d_and_pseudo3d_golden(kind, model, ts, dt, dm, opt, expected):
    case, params = golden_3d_setup(kind, model, ts, dt, dm, opt)
    out = case.run_oracle(R, params)
    np.testing.assert_allclose(out.position[out.inactive == 0], np.array(expected), rtol=1e-7)


REBOUND_EXPECTED = np.array([
    # tests/test_13_core_rebound_bnd_conditions.py:213-224
    [2900.024, 7500.], [2900.024, 7700.012], [2900.024, 2700.012], [2900.024, -2299.988],
    [2900.024, -7299.988], [2900.024, -7700.012], [2900.024, -2700.012],
    [2900.024, 2299.988], [2900.024, 7299.988], [2900.024, 7900.024]])


def test_model0_specular_rebound_golden():
    pos = np.full((11, 2), 7500.)
    vel = 10. * np.array([[1., 0.1 * k] for k in range(11)])
    case = reference_square_case(pos, vel, pool=12)
    params = dict(final_time=5000., time_step=50., time_scheme=1, frozen_eulerian_fields=True,
                  model=0, particle_velocity_init=0, boundary_conditions=True,
                  max_wall_rebounds=4, rebound_damping_coef=0.)
    out = case.run_oracle(R, params)
    xp = out.position[np.argsort(out.inactive, kind="stable")][:int((out.inactive == 0).sum())]
    np.testing.assert_allclose(xp, REBOUND_EXPECTED, rtol=1e-7)


def _seg(a, b, c, d):
    f = [np.ascontiguousarray(v, dtype=float) for v in (a, b, c, d)]
    return R.lib.ref_intersec2d(*[R.P(v, "d") for v in f])


def test_segment_intersection_vectors():
    # tests/test_01_geom_intersections.py:102-114
    A = [[1., 1.], [1., 1.], [1., 2.], [5., 2.], [4., 1.]]
    B = [[4., 4.], [4., 4.], [3., 4.], [7., 4.], [6., 3.]]
    Cc = [[2., 2.], [3., 1.], [1., 4.], [6., 5.], [6., 5.]]
    D = [[5., 5.], [6., 4.], [3., 1.], [8., 3.], [8., 3.]]
    assert [_seg(A[i], B[i], Cc[i], D[i]) for i in range(5)] == [1, 0, 1, 1, 0]


@pytest.mark.parametrize("damp,expected", [(0., [0., 4.]), (1., [1.5, 2.5])])
def test_symmetric_point_vectors(damp, expected):
    # tests/test_01_geom_intersections.py:14-27
    a, b, d = (np.array(v, dtype=float) for v in ([1., 2.], [3., 4.], [3., 1.]))
    out = np.zeros(2)
    R.lib.ref_symmetric_point2d(R.P(a, "d"), R.P(b, "d"), R.P(d, "d"), damp, R.P(out, "d"))
    assert list(out) == expected


def test_lower_layer_index_vectors():
    # tests/test_03_geom_zlocalizer.py:15-24,58-67
    assert R.lib.ref_lower_layer_index(11, 10., 0., 10. * 0.55) == 5
    assert R.lib.ref_lower_layer_index(11, 10., 0., 10. * 0.74) == 7
    # Cython lowers double // to floor(a / b): with hl = 0.1, 1.0 / hl rounds to
    # 10.0 (Python's 1.0 // 0.1 would give 9.0), so the index is 10 = nl - 2
    assert 1.0 // ((1.1 - 0.) / 11) == 9.0
    assert R.lib.ref_lower_layer_index(12, 1.1, 0., 1.0) == 10
    assert R.lib.ref_lower_layer_index(1, 1., 0., 0.5) == 0
    assert R.lib.ref_lower_layer_index(5, 0., 1., 0.5) == 0   # dry column
    assert R.lib.ref_lower_layer_index(5, 4., 0., -1.) == 0
    assert R.lib.ref_lower_layer_index(5, 4., 0., 9.) == 3


def test_localization_agrees_with_brute_force_and_trifinder():
    """Method-0 search (KD-tree rule) returns a containing triangle, the
    brute-force one off shared edges, and agrees with matplotlib's trifinder
    on interior points (tests/test_02_geom_xylocalizer.py:258-315)."""
    from matplotlib.tri import Triangulation
    case = reference_square_case(P4)
    m = R.RefMesh(case.x, case.y, case.tri)
    tf = Triangulation(case.x, case.y, case.tri).get_trifinder()
    rng = np.random.default_rng(0)
    xs = rng.uniform(case.x.min(), case.x.max(), 300)
    ys = rng.uniform(case.y.min(), case.y.max(), 300)
    for x, y in zip(xs, ys):
        k = m.kdtree(x, y)
        kb = m.brute(x, y)
        assert (k == -1) == (kb == -1)
        if k != -1:
            assert R.lib.ref_point_in_tri(C.byref(m.s), k, x, y) == k
            assert k == int(tf(x, y))


def test_walk_reaches_the_containing_triangle():
    case = reference_square_case(P4)
    m = R.RefMesh(case.x, case.y, case.tri)
    rng = np.random.default_rng(1)
    for _ in range(200):
        k0 = int(rng.integers(m.nt))
        t = m.triangles[k0]
        p0 = (m.x[t].mean(), m.y[t].mean())
        p1 = (p0[0] + rng.normal() * 300., p0[1] + rng.normal() * 300.)
        k, b = m.localize(k0, p0[0], p0[1], p1[0], p1[1], 3)
        if b == -1:
            assert k == m.brute(*p1) or R.lib.ref_point_in_tri(C.byref(m.s), k, *p1) == k
        else:
            assert m.edges_labels[b] >= 0 and m.edgetri[b, 1] == -1


def test_diffusivity_gradient_exact_on_linear_nodal_K():
    """K nodal linear in x, y is reproduced exactly by P1 with its gradient
    (tests/test_16_core_diffusivity_gradient_drift.py:48-102)."""
    from tests.cases import structured_square
    x, y, tri = structured_square(11, 10.)
    m = R.RefMesh(x, y, tri)
    a, bx, by = 0.3, 0.05, -0.02
    K = a + bx * x + by * y
    eps = np.full(x.size, 1e-3)
    tke = np.sqrt(K * eps / (0.09 / 0.72))
    n = x.size
    f = R.RefFields(m, 2, 1, np.zeros(1), np.ones((1, n)), np.zeros((1, n)), np.zeros((1, n)),
                    opt_fields=np.stack([tke[None], eps[None]]),
                    opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    out = [C.c_double() for _ in range(4)]
    for (px, py) in [(1.3, 2.7), (5.5, 5.1), (9.9, 0.2)]:
        k = m.brute(px, py)
        R.lib.ref_diffusivity_gradient(C.byref(f.s), C.byref(m.s), k, 0, px, py, 0., 0.72,
                                       *[C.byref(o) for o in out])
        assert out[0].value == pytest.approx(a + bx * px + by * py, rel=1e-12)
        assert out[1].value == pytest.approx(bx, rel=1e-9)
        assert out[2].value == pytest.approx(by, rel=1e-9)


@pytest.mark.parametrize("a1,dt", [(0.0, 2.), (1e-3, 2.), (0.04, 2.), (0.3, 2.), (5., 2.)])
def test_lsm2_direct_coefficients_vs_quadrature(a1, dt):
    """Frozen-step OU coefficients of du = -a1 u dt + dW, dx = u dt: the
    transition e^{A dt} (scipy expm) and the covariance integral
    Q = int_0^dt e^{As} B B^T e^{A^T s} ds (adaptive quadrature)."""
    from scipy.integrate import quad
    from scipy.linalg import expm
    vals = [C.c_double() for _ in range(7)]
    assert R.lib.ref_lsm2_direct_coefficients(a1, dt, *[C.byref(v) for v in vals]) == 0
    beta, D1, D2, fx, quu, qxu, qxx = [v.value for v in vals]
    Phi = expm(np.array([[-a1, 0.], [1., 0.]]) * dt)
    assert beta == pytest.approx(Phi[0, 0], rel=1e-12)
    assert D1 == pytest.approx(Phi[1, 0], rel=1e-12)
    g = (lambda s: np.exp(-a1 * s)) if a1 > 0 else (lambda s: 1.)
    h = (lambda s: -np.expm1(-a1 * s) / a1) if a1 > 0 else (lambda s: s)
    o = dict(epsabs=0., epsrel=1e-13, limit=200)
    assert quu == pytest.approx(quad(lambda s: g(s) ** 2, 0, dt, **o)[0], rel=1e-11)
    assert qxu == pytest.approx(quad(lambda s: g(s) * h(s), 0, dt, **o)[0], rel=1e-11)
    assert qxx == pytest.approx(quad(lambda s: h(s) ** 2, 0, dt, **o)[0], rel=1e-11)
    assert D2 == pytest.approx(dt - D1, rel=1e-12, abs=1e-15)
    # force term: int_0^dt (1 - e^{-a1 s})/a1 ds
    assert fx == pytest.approx(quad(h, 0, dt, **o)[0], rel=1e-11)
