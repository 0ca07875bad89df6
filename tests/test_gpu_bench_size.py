"""Parity at the measured workloads' own mesh sizes, against the oracle.

* The plain 2D workload: Delaunay mesh of 1M jittered points (~2M elements),
  two field frames, k-epsilon K with its drift, Ziggurat normals; 156,250
  particles (every 64th of the 10M release) for 24 steps with the element
  re-sort every 8 steps, so three re-sorts happen.
* The 3D sigma-layer channel (~500k elements, 10 planes) with settling
  sediment (LSM-2 direct integrator, k-epsilon K(z)) and with LSM-1.

Neither workload has a trifinder: the device locates the releases itself
and the oracle follows the reference's KD-tree search (solver.pyx:262-268,
xylocalizer.pyx:200-286).  Identical element ids, layer ids, flags, tags and
random-stream state; positions and velocities within 1e-9 relative."""
from types import SimpleNamespace

import numpy as np
import pytest

import cylag_amd as cy
from oracle import refpy as R
from tests.cases import SEED, c2_workload, c3_release, c3_workload
from tests.test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu


def run_pair(x, y, tri, fields, pos, params, nsteps, sort_interval, pset_kw=None, diameter=None):
    dim = pos.shape[1]
    mesh = cy.TriangularMesh(x, y, tri)
    fs = cy.EulerianFieldSet(mesh, **fields)
    ps = cy.LagrangianParticleSet(pos.copy(), dim=dim, **(pset_kw or {}))
    if diameter is not None:
        ps.diameter = diameter
        ps.volume = np.pi * diameter ** 3 / 6.
    cy.pcg32_seed(*SEED)
    s = cy.Solver(fs, ps, params)
    s.ctx.lib.cyl_set_tuning(s.ctx.h, sort_interval, 0, 0, 0)
    s._run(nsteps)
    fused = s.ctx.last_path()[0]
    g = SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(), tri=ps.tri.copy(),
                        inactive=ps.inactive.copy(), lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy())
    s.ctx.close()
    rm = R.RefMesh(x, y, tri)
    rf = R.RefFields(rm, **fields)
    rp = R.RefParticles(pos.copy(), dim=dim, diameter=diameter, **(pset_kw or {}))
    rng = R.RefRng(*SEED)
    o = R.RefSolver(rf, rp, params, trifinder=rm.kdtree_locator(), rng=rng)
    o.forward(nsteps)
    r = SimpleNamespace(position=rp.position.copy(), velocity=rp.velocity.copy(), tri=rp.tri.copy(),
                        inactive=rp.inactive.copy(), lowerlayer=rp.lowerlayer.copy(), tag=rp.tag.copy(), rng=rng)
    return g, r, fused


@pytest.fixture(scope="module")
def c2():
    x, y, tri, fields, pos, params = c2_workload(npart=10_000_000)
    return x, y, tri, fields, pos[::64].copy(), params


@pytest.mark.parametrize("rng_method", [2, 3])
def test_plain_workload_mesh_resorted(c2, rng_method):
    x, y, tri, fields, pos, params = c2
    assert tri.shape[0] > 1_990_000 and pos.shape[0] == 156_250
    g, r, fused = run_pair(x, y, tri, fields, pos, dict(params, rng_method=rng_method), 24, 8)
    assert fused == 24
    assert_parity(g, r)


@pytest.fixture(scope="module")
def c3():
    x, y, tri, h, fields = c3_workload()
    pos, d = c3_release(x, y, h, 60_000)
    return x, y, tri, fields, pos, d


@pytest.mark.parametrize("model", [1, 2])
def test_3d_channel_settling(c3, model):
    x, y, tri, fields, pos, d = c3
    assert tri.shape[0] > 490_000
    params = dict(final_time=1e9, time_step=2., frozen_eulerian_fields=True, diffusion_model=3)
    if model == 1:
        params.update(model=1)
    else:
        params.update(model=2, time_scheme=5, diffusion_lsm2_option=1, particle_velocity_init=1)
    kw = dict(particle_density=2650., buoyancy_velocity_model=1)
    g, r, fused = run_pair(x, y, tri, fields, pos, params, 20, 8, pset_kw=kw, diameter=d)
    assert fused == 20
    assert_parity(g, r)
