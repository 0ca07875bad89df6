"""The arrival path end to end on the device: a TELEMAC flume read from the
reference's data, a polygon source releasing 4 particles every 5 s, the TXT
output every step, read back -- the golden values of the reference's
tests/test_52_extra_sources.py:27-79."""
import os

import numpy as np
import pytest

import cylag_amd as cy

pytestmark = pytest.mark.gpu

DATA = os.path.join(os.path.dirname(__file__), "golden", "reference_data")


def test_flume_source_and_txt_output(tmp_path):
    fset = cy.EulerianFieldSet.from_telemac2d(os.path.join(DATA, "r2d_flume.slf"),
                                              bnd_file=os.path.join(DATA, "geo_flume.cli"))
    pset = cy.LagrangianParticleSet(None, dim=2, initial_pool_size=1, min_pool_size_update=1)
    poly = np.array([[5.1, 0.5], [6.1, 0.5], [6.1, 1.5], [5.1, 1.5]])
    source = cy.ParticleSource(cy.schedules(time_delta=5.), poly, xy_method=1, xy_npart=4)
    parameters = {
        'final_time': 20., 'time_step': 0.1, 'frozen_eulerian_fields': True, 'particle_velocity_init': 1,
        'model': 1, 'time_scheme': 1, 'diffusion_model': 0, 'horizontal_diffusivity': 3.e-5,
        'vertical_diffusivity': 0., 'boundary_conditions': True, 'listing': False,
        'output_file': True, 'output_printout_period': 1, 'output_file_format': 'txt',
        'output_file_name': 'particles_2d', 'output_rep': str(tmp_path), 'bnd_statistics': False,
    }
    solver = cy.Solver(fset, pset, parameters, [source])
    solver.solve()
    xp, up, us = solver.particle_set.get_state()
    xpa = np.array([[10.35, 0.75], [10.85, 0.75], [10.35, 1.25], [10.85, 1.25], [5.45, 0.75], [5.95, 0.75],
                    [5.45, 1.25], [5.95, 1.25], [15.25, 0.75], [15.75, 0.75], [15.25, 1.25], [15.75, 1.25]])
    upa = np.tile([1., 0.], (12, 1))
    np.testing.assert_allclose(xp, xpa, rtol=1e-7)
    np.testing.assert_allclose(up, upa, rtol=1e-7)
    part = cy.ParticlesIO.from_cylag_txt(os.path.join(str(tmp_path), "particles_2d.txt"))
    np.testing.assert_allclose(part.xp[100], [15.35, 15.75, 15.25, 15.75, 10.25, 10.75, 10.25, 10.75], rtol=1e-7)
    np.testing.assert_allclose(part.xp[200], [10.35, 10.85, 10.35, 10.85, 5.45, 5.95, 5.45, 5.95, 15.25, 15.75,
                                              15.25, 15.75], rtol=1e-7)
