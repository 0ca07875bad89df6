"""The step's documented departures from the reference's serial draws, and
the device's replacement searches, checked against the oracle.

* Stranding (compute_bc.pyx:100-121) and the rough-wall normal perturbation
  (compute_bc.pyx:199-208) draw, in the reference, from the serial stream
  in the middle of the pool loop; the device draws them from per-row
  sub-streams.  The outcomes can only agree in distribution: the stranded
  fraction and the distribution of rebound angles must match the oracle's
  (and their closed forms) within 99% intervals.
* The fallback search after a failed walk (xy_localize method 3, KD-tree
  scan of xylocalizer.pyx:200-286; a bucket grid on the device) and the
  initial search of rows without a trifinder (solver.pyx:262-268) must give
  the oracle's element ids exactly.
"""
from types import SimpleNamespace

import numpy as np
import pytest
from scipy import stats

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import square_case
from tests.test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu

Z99 = 2.5758


def wall_release(n, seed, vy, u=0.):
    """Rows just below the top wall of the 100 m square, in a uniform flow
    (u, vy) that carries every one of them across the wall in the first step."""
    case = square_case(n=51, npart=1, u=u, v=vy)
    rng = np.random.default_rng(seed)
    case.position = np.stack([rng.uniform(20., 80., n), rng.uniform(96., 99.5, n)], 1)
    return case


@pytest.mark.parametrize("p", [0.3, 0.7])
def test_stranded_fraction(p):
    n = 20_000
    case = wall_release(n, 41, vy=2.)
    params = dict(final_time=3., time_step=3., frozen_eulerian_fields=True, model=1, diffusion_model=0,
                  stranding_probability=p)
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    kg, kr = int(g.inactive.sum()), int(r.inactive.sum())
    # each row hits the wall once and strands with probability p
    for k in (kg, kr):
        assert abs(k / n - p) <= Z99 * np.sqrt(p * (1. - p) / n), (k, n, p)
    assert abs(kg - kr) / n <= Z99 * np.sqrt(2. * p * (1. - p) / n), (kg, kr)
    # the rows that stranded in neither run (independent draws: a share of
    # (1-p)^2) rebounded identically, as a plain rebound draws nothing
    keep = (g.inactive == 0) & (r.inactive == 0)
    assert keep.sum() > 0.5 * (1. - p) ** 2 * n
    np.testing.assert_allclose(g.position[keep], r.position[keep], rtol=1e-12)


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_rough_wall_rebound_angles(rng_method):
    """Test model (x += up dt), velocity (0, 2) against the top wall: the
    reflected velocity turns by twice the normal's perturbation theta =
    normal * roughness, clipped at 85 degrees (compute_bc.pyx:199-208)."""
    n = 6000
    rough = 15.
    case = wall_release(n, 43, vy=0.)
    case.velocity = np.tile([0., 2.], (n, 1))
    params = dict(final_time=3., time_step=3., frozen_eulerian_fields=True, model=0, particle_velocity_init=0,
                  boundary_conditions_type=2, wall_roughness_angle=rough, rng_method=rng_method,
                  rebound_damping_coef=0.)
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    ang = []
    for out in (g, r):
        act = out.inactive == 0
        v = out.velocity[act]
        assert act.sum() > 0.95 * n
        np.testing.assert_allclose(np.hypot(v[:, 0], v[:, 1]), 2., rtol=1e-12)   # undamped: speed kept
        ang.append(np.arctan2(v[:, 0], -v[:, 1]))   # 0 for a specular rebound
    # same distribution (two-sample KS at 1%), and twice the perturbation's spread
    assert stats.ks_2samp(ang[0], ang[1]).pvalue > 0.01
    # every sampler has unit variance; clipping at 85 degrees (5.7 sigma) removes nothing measurable
    sd_true = 2. * np.deg2rad(rough)
    for a in ang:
        s = a.std()
        assert abs(s / sd_true - 1.) <= Z99 * np.sqrt(0.5 / len(a)) * (1.5 if rng_method == 1 else 1.), (s, sd_true)


def test_fallback_search_after_long_walks():
    """Rows crossing more than 50 elements in a step: the walk gives up and the
    fallback search locates them (device bucket grid against the reference's
    KD-tree scan).  Elements have a diameter of sqrt(2) m and every path is
    78 m long, so each crosses more than 50 of them."""
    n = 3000
    case = square_case(n=101, npart=1, u=60., v=50.)
    rng = np.random.default_rng(45)
    case.position = np.stack([rng.uniform(3., 35., n), rng.uniform(3., 45., n)], 1)
    params = dict(final_time=1., time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=0)
    g = case.run_gpu(cy, params)
    r = case.run_oracle(R, params)
    assert np.all(g.inactive == 0)
    assert_parity(g, r)


@pytest.mark.parametrize("on_vertices", [False, True])
def test_initial_search_without_trifinder(on_vertices):
    """Rows located on the device (no trifinder, the reference's KD-tree path
    of solver.pyx:262-268): random interior points, and points on mesh
    vertices and edge midpoints, where several elements contain the point and
    the tie order decides (nearest centroid, then lower index)."""
    case = square_case(n=41, npart=1, u=0.03, v=0.02)
    case.use_trifinder = False
    rng = np.random.default_rng(47)
    if on_vertices:
        xs = np.linspace(0., 100., 41)
        vx, vy = np.meshgrid(xs[2:-2], xs[2:-2])
        mid = 0.5 * (xs[2:-3] + xs[3:-2])
        pts = np.concatenate([np.stack([vx.ravel(), vy.ravel()], 1),
                              np.stack([np.repeat(mid, 3)[:400], np.tile(xs[5:8], len(mid))[:400]], 1)])
    else:
        pts = np.stack([rng.uniform(1., 99., 4000), rng.uniform(1., 99., 4000)], 1)
    case.position = pts
    params = dict(final_time=20., time_step=2., frozen_eulerian_fields=True, model=1, diffusion_model=1,
                  horizontal_diffusivity=0.05)
    s = case.build_gpu(cy, params)
    ps = s.particle_set
    tri0 = ps.tri.copy()
    # the oracle's restatement of the KD-tree scan (ties: nearest centroid, then lower index)
    mesh = R.RefMesh(case.x, case.y, case.tri)
    ref0 = np.array([mesh.kdtree(x, y) for x, y in pts], np.int32)
    np.testing.assert_array_equal(tri0, ref0)
    s.solve()
    g = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                        inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy())
    assert_parity(g, case.run_oracle(R, params))
