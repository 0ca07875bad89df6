"""Served requests: each request's result from the serving loop (admitted
while others run, sharing the device batch, retired when done) equals a plain
Solver run of that request alone and the oracle's serial run of it -- element
ids, flags, tags and random stream identical, positions within 1e-9
relative."""
import numpy as np
import pytest

import cylag_amd as cy
from tests.cases import settling_3d_case, vortex_case
from tests.test_gpu_batch import compare, make_pset, pick, run_alone, run_oracle, streams

pytestmark = pytest.mark.gpu

REQS = [  # (parameters, particles, steps, arrival s)
    (dict(model=1, diffusion_model=2), 1500, 40, 0.),
    (dict(model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=1, particle_velocity_init=1), 800, 25, 0.),
    (dict(model=1, diffusion_model=1, horizontal_diffusivity=0.4), 1200, 30, 0.02),
    (dict(model=1, time_scheme=4, diffusion_model=0), 600, 12, 0.05),
    (dict(model=2, time_scheme=5, diffusion_model=2, diffusion_lsm2_option=2, particle_velocity_init=1), 700, 18, 0.05),
    (dict(model=1, diffusion_model=2), 2500, 22, 0.1),
]


@pytest.mark.parametrize("rng_method", [1, 2, 3])
def test_served_equals_alone(rng_method):
    case = vortex_case(npart=5000)
    base = dict(final_time=10000., time_step=5., rng_method=rng_method, frozen_eulerian_fields=True)
    st = streams(len(REQS))
    reqs = [cy.Request(make_pset(case, n, 400 + i), dict(base, **p), st[i], k, arrival=a)
            for i, (p, n, k, a) in enumerate(REQS)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    # a clock driven by the loop's own steps (10 ms per step), so that the
    # admission order does not depend on the GPU's speed
    ticks = [0.]

    def clock():
        ticks[0] += 1e-4
        return ticks[0] + 0.01 * loop.steps

    loop = cy.ServeLoop(fs, capacity=4000, tick_steps=4, sort_interval=3, clock=clock)
    done = loop.serve(reqs)
    loop.close()
    assert len(done) == len(REQS) and not loop.unfinished
    assert loop.max_running >= 3   # requests overlapped on the device
    for i, (p, n, k, a) in enumerate(REQS):
        r = reqs[i]
        alone = run_alone(case, make_pset(case, n, 400 + i), dict(base, **p), st[i], k)
        compare(r.particle_set, alone, r.final_rng_state)
        compare(r.particle_set, run_oracle(case, pick(case, n, 400 + i), dict(base, **p), st[i], k),
                r.final_rng_state)
    st_ = cy.latency_stats(done)
    assert st_["completed"] == len(REQS)


REQS_3D = [  # (model, particles, steps, arrival s): point releases of settling sediment
    (1, 900, 40, 0.),
    (2, 1400, 25, 0.),
    (2, 300, 30, 0.02),
    (1, 2000, 15, 0.03),
    (1, 500, 35, 0.05),
    (2, 700, 20, 0.08),
]


@pytest.mark.parametrize("rng_method", [1, 2])
def test_served_3d_settling_against_oracle(rng_method):
    """The served workload's kind of request on a small sigma-layer mesh: point
    releases at random depths of sediment with per-row settling diameters,
    LSM-1 or 3D LSM-2 (k-epsilon K(z), frozen fields), overlapping on the
    device; each request against the oracle."""
    case = settling_3d_case(npart=10, spread="lognormal")
    rng = np.random.default_rng(31)
    base = dict(final_time=1e9, time_step=10., rng_method=rng_method, frozen_eulerian_fields=True,
                diffusion_model=3)
    mp = {1: dict(model=1), 2: dict(model=2, time_scheme=5, diffusion_lsm2_option=1, particle_velocity_init=1)}
    st = streams(len(REQS_3D))
    rows = []
    for i, (m, n, k, a) in enumerate(REQS_3D):
        src = case.position[i]
        rows.append((np.repeat(src[None], n, 0), np.exp(np.log(3.3e-5) + 0.5 * rng.standard_normal(n))))
    kw = dict(particle_density=2650., buoyancy_velocity_model=1)

    def pset(i):
        pos, d = rows[i]
        ps = cy.LagrangianParticleSet(pos.copy(), dim=3, **kw)
        ps.diameter = d
        ps.volume = np.pi * d ** 3 / 6.
        return ps

    reqs = [cy.Request(pset(i), dict(base, **mp[m]), st[i], k, arrival=a) for i, (m, n, k, a) in enumerate(REQS_3D)]
    mesh = cy.TriangularMesh(case.x, case.y, case.tri)
    fs = cy.EulerianFieldSet(mesh, trifinder=case.trifinder(), **case.fields)
    ticks = [0.]

    def clock():
        ticks[0] += 1e-4
        return ticks[0] + 0.01 * loop.steps

    loop = cy.ServeLoop(fs, capacity=3500, tick_steps=5, sort_interval=4, clock=clock)
    done = loop.serve(reqs)
    loop.close()
    assert len(done) == len(REQS_3D) and not loop.unfinished
    assert loop.max_running >= 3
    for i, (m, n, k, a) in enumerate(REQS_3D):
        r = reqs[i]
        ref = run_oracle(case, rows[i][0], dict(base, **mp[m]), st[i], k, pset_kw=kw, diameter=rows[i][1])
        compare(r.particle_set, ref, r.final_rng_state)
        assert np.any(r.particle_set.lowerlayer != r.particle_set.lowerlayer[0])
