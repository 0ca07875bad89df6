"""Host side of the arrival path and the data formats around the step:
schedulers (the reference's golden trigger times, tests/test_51), initial
positions in polygons, the SELAFIN reader (against the triangulation the
reference's reader exported, tests/data/test_triangulation_mesh*.txt), and the
particle TXT writer / reader (round trip, and the reference's own result file)."""
import os

import numpy as np
import pytest

import cylag_amd as cy
from cylag_amd.particles_io import ParticlesIO, prepare_output, write_particles

DATA = os.path.join(os.path.dirname(__file__), "golden", "reference_data")


def _trigger_times(scheduler):
    st = cy.SimuTime(False, final_time=5., time_step=5. / 35.)
    out = []
    while not st.is_finished:
        if scheduler.now(st):
            out.append(st.time)
        st.increment()
    return out


# tests/test_51_extra_scheduler.py:41-121
SCHEDULER_GOLDEN = [
    (dict(time_delta=1.5), [0.0, 1.4285714285714282, 2.999999999999999, 4.428571428571428]),
    (dict(iteration_delta=5), [0.0, 0.7142857142857142, 1.4285714285714282, 2.1428571428571423,
                               2.8571428571428563, 3.5714285714285703, 4.285714285714285]),
    (dict(times=[0.5, 1., 2.5, 3.]), [0.42857142857142855, 0.9999999999999998, 2.428571428571428,
                                      2.999999999999999]),
    (dict(iterations=[10, 12, 14, 16]), [1.4285714285714282, 1.7142857142857137, 1.9999999999999993,
                                         2.285714285714285]),
    (dict(count=4), [0.0, 1.571428571428571, 3.2857142857142847, 4.857142857142858]),
]


@pytest.mark.parametrize("kw,expected", SCHEDULER_GOLDEN, ids=[list(k)[0] for k, _ in SCHEDULER_GOLDEN])
def test_scheduler_golden(kw, expected):
    np.testing.assert_equal(_trigger_times(cy.schedules(**kw)), expected)


def test_scheduler_always_never_and_errors():
    assert len(_trigger_times(cy.schedules(always=True))) == 35
    assert _trigger_times(cy.schedules(never=True)) == []
    with pytest.raises(ValueError):
        cy.schedules()
    with pytest.raises(ValueError):
        cy.schedules(count=2, always=True)


def test_polygon_area_and_circle():
    assert cy.polygon_area([0., 1., 1., 0.], [0., 0., 1., 1.]) == 1.
    c = cy.circle(0., 0., 2., 720)
    np.testing.assert_allclose(np.hypot(c[:, 0], c[:, 1]), 2.)
    # inscribed regular polygon: (n/2) r^2 sin(2 pi / n)
    assert abs(cy.polygon_area(c[:, 0], c[:, 1]) - 0.5 * 720 * 4. * np.sin(2 * np.pi / 720)) < 1e-9


def test_init_positions_2d_grid_centres():
    sq = np.array(([0., 0.], [1., 0.], [1., 1.], [0., 1.]))
    pos = cy.init_positions_2d(poly=sq, poly_area=1., xy_method=2, xy_density=100)
    assert pos.shape == (100, 2)
    g = (np.arange(10) + 0.5) / 10.
    np.testing.assert_allclose(np.sort(pos[:, 0]), np.repeat(g, 10))
    # a triangle keeps exactly the centres inside it (none lies on an edge here)
    tri = np.array(([0., 0.], [1., 0.], [0.3, 1.]))
    pt = cy.init_positions_2d(poly=tri, grid_res=(10, 10))
    c = (np.arange(10) + 0.5) / 10.
    X, Y = np.meshgrid(c, c)
    s1 = Y                      # above the base
    s2 = 1. * X - 0.3 * Y       # right of the left edge (0,0)-(0.3,1)
    s3 = (1. - X) - 0.7 * Y     # left of the right edge (1,0)-(0.3,1)
    assert min(np.abs(s2).min(), np.abs(s3).min()) > 1e-9
    inside = (s1 > 0) & (s2 > 0) & (s3 > 0)
    assert len(pt) == inside.sum()
    assert set(map(tuple, np.round(pt, 12))) == set(zip(np.round(X[inside], 12), np.round(Y[inside], 12)))
    # xy_method 1: about xy_npart points for the polygon's area
    p1 = cy.init_positions_2d(poly=sq, poly_area=1., xy_method=1, xy_npart=16)
    assert p1.shape == (16, 2)


def test_init_positions_3d_layers():
    sq = np.array(([0., 0.], [1., 0.], [1., 1.], [0., 1.]))
    pos = cy.init_positions_3d(poly=sq, poly_area=1., grid_res=[2, 2, 2], z_method=2, z_density=1.,
                               zmin=0., zmax=2.)
    assert pos.shape == (8, 3)
    np.testing.assert_allclose(np.unique(pos[:, 2]), [0.5, 1.5])
    with pytest.raises(ValueError):
        cy.init_positions_2d(poly=sq[:2])


def test_source_releases_on_schedule():
    src = cy.ParticleSource(cy.schedules(time_delta=5.), np.array([[5.1, 0.5], [6.1, 0.5], [6.1, 1.5],
                                                                   [5.1, 1.5]]), xy_method=1, xy_npart=4)
    st = cy.SimuTime(False, final_time=20., time_step=0.1)
    n = 0
    while not st.is_finished:
        if src.release_condition(st, None):
            n += 1
            pos = src.init_positions(None, 2)
            np.testing.assert_allclose(np.sort(pos[:, 0]), [5.35, 5.35, 5.85, 5.85])
        st.increment()
    assert n == 5  # t = 0, 5, 10, 15 and the last step's start just below 20


def test_selafin_reader_matches_reference_triangulation():
    r = cy.SelafinFile(os.path.join(DATA, "test_triangulation.slf"))
    tri = r.tri
    np.testing.assert_array_equal(tri.triangles, np.loadtxt(os.path.join(DATA, "test_triangulation_meshtri.txt")))
    # the exported coordinates carry three decimals
    assert np.abs(tri.x - np.loadtxt(os.path.join(DATA, "test_triangulation_meshx.txt"))).max() <= 5e-4 + 1e-9
    assert np.abs(tri.y - np.loadtxt(os.path.join(DATA, "test_triangulation_meshy.txt"))).max() <= 5e-4 + 1e-9


def test_from_telemac2d_flume():
    fs = cy.EulerianFieldSet.from_telemac2d(os.path.join(DATA, "r2d_flume.slf"),
                                            bnd_file=os.path.join(DATA, "geo_flume.cli"))
    m = fs.triangular_mesh
    assert (m.nv, m.nt, m.nopen) == (729, 1280, 2)
    np.testing.assert_array_equal(fs.times, [0.])
    np.testing.assert_allclose(fs.velocity_x, 1.)
    np.testing.assert_allclose(fs.velocity_y, 0.)


def test_reference_particle_file():
    part = ParticlesIO.from_cylag_txt(os.path.join(DATA, "particles_2d.txt"))
    assert part.ntimes == 23
    assert part.iterations[0] == 0 and part.npart[0] == 0 and part.npart[1] == 9
    assert part.times[1] == 87200.
    assert part.xp[1][0] == 2.1301452107260420e+05 and part.yp[1][0] == 1.7964914899277448e+05
    assert part.up[1][0] == 1.41179050e+00


def test_txt_writer_round_trip(tmp_path):
    pos = np.array([[1.25, 2.5], [3.0, 4.125], [5.5, 6.0]])
    ps = cy.LagrangianParticleSet(pos, dim=2)
    ps.velocity = np.array([[0.1, 0.2], [0.3, 0.4], [0.5, 0.6]])
    ps.inactive = np.array([0, 1, 0], dtype=np.int32)
    ps.npart_active = 2
    params = cy.Parameters(dict(output_rep=str(tmp_path), output_file_name="p", output_file=True))
    prepare_output(ps, params.output_rep, params.output_file_format, params.output_file_name, params.model)
    write_particles(ps, 0., 0, params)
    write_particles(ps, 12.5, 10, params)
    part = ParticlesIO.from_cylag_txt(os.path.join(str(tmp_path), "p.txt"))
    assert part.ntimes == 2 and list(part.iterations) == [0, 10] and part.times[1] == 12.5
    np.testing.assert_array_equal(part.tags[1], [0, 2])
    np.testing.assert_array_equal(part.xp[1], [1.25, 5.5])
    np.testing.assert_array_equal(part.zp[1], [0., 0.])
    np.testing.assert_array_equal(part.vp[0], [0.2, 0.6])
    np.testing.assert_array_equal(part.get_trajectory(2), [[5.5, 6.0, 0.], [5.5, 6.0, 0.]])


def test_vtk_writer_layout(tmp_path):
    """writer_vtk.pyx:7-304: legacy-VTK vertex cells of the active rows, with
    pool index, tag, elevation (3D), velocity, fluid velocity seen (LSM-2/3)."""
    pos = np.array([[1.25, 2.5, -0.5], [3.0, 4.125, -1.0], [5.5, 6.0, -2.0]])
    ps = cy.LagrangianParticleSet(pos, dim=3)
    ps.velocity = np.array([[0.1, 0.2, 0.3], [0.3, 0.4, 0.5], [0.5, 0.6, 0.7]])
    ps.inactive = np.array([0, 1, 0], dtype=np.int32)
    ps.npart_active = 2
    params = cy.Parameters(dict(output_rep=str(tmp_path), output_file_name="p", output_file=True,
                                output_file_format='all', model=2))
    prepare_output(ps, params.output_rep, params.output_file_format, params.output_file_name, params.model)
    write_particles(ps, 12.5, 10, params)
    lines = (tmp_path / "p_10.vtk").read_text().splitlines()
    assert lines[:10] == ['# vtk DataFile Version 2.0', 'particle positions (Points 3D)', 'ASCII',
                          'DATASET UNSTRUCTURED_GRID', 'FIELD FIELDDATA 2', 'TIME 1 1 DOUBLE', '12.5',
                          'CYCLE 1 1 INT', '10', 'POINTS 2 FLOAT']
    assert lines[10:12] == ['1.25 2.5 -0.5', '5.5 6.0 -2.0']
    assert lines[12:17] == ['CELLS 2 4', '1 0', '1 1', 'CELL_TYPES 2', '1']
    i = lines.index('SCALARS Index INT')
    assert lines[i + 2:i + 4] == ['0', '2']
    i = lines.index('SCALARS tag INT')
    assert lines[i + 2:i + 4] == ['0', '2']
    i = lines.index('SCALARS elevation FLOAT')
    assert lines[i + 2:i + 4] == ['-0.5', '-2.0']
    i = lines.index('VECTORS velocity FLOAT')
    assert lines[i + 1:i + 3] == ['0.1 0.2 0.3', '0.5 0.6 0.7']
    assert 'VECTORS fluid_velocity_seen FLOAT' in lines and 'SCALARS diameter FLOAT' not in lines
    assert (tmp_path / "p.txt").exists()  # 'all' writes both


def test_stranded_file_layout(tmp_path):
    """write_particles.pyx:339-388: header, then x, y, z per stranded row."""
    from cylag_amd.particles_io import write_stranded_particles_txt
    buf = np.array([[1.5, 2.25, 0.], [3., 4., 0.], [9., 9., 9.]])
    write_stranded_particles_txt(str(tmp_path), "stranded", buf, 2)
    lines = (tmp_path / "stranded.txt").read_text().splitlines()
    assert lines == ['cylag stranded particles file', 'Variables: Xp, Yp, Zp',
                     '1.5000000000000000e+00, 2.2500000000000000e+00, 0.0000000000000000e+00',
                     '3.0000000000000000e+00, 4.0000000000000000e+00, 0.0000000000000000e+00']
    write_stranded_particles_txt(str(tmp_path), "none", buf, 0)
    assert not (tmp_path / "none.txt").exists()
