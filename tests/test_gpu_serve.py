"""Served requests: each request's result from the serving loop (admitted
while others run, sharing the device batch, retired when done) equals a plain
Solver run of that request alone -- element ids, flags, tags and random stream
identical, positions within 1e-9 relative."""
import numpy as np
import pytest

import cylag_amd as cy
from tests.cases import vortex_case
from tests.test_gpu_batch import compare, make_pset, run_alone, streams

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
    st_ = cy.latency_stats(done)
    assert st_["completed"] == len(REQS)
