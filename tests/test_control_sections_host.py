"""Control sections on the CPU: the oracle's per-step counts against closed
forms, the ControlSection segment rule and the statistics file format."""
import numpy as np

import cylag_amd as cy
from cylag_amd.control_shapes import section_segments
from oracle import refpy as R
from tests.cases import open_channel_case


def uniform_channel(xs, ys, u=0.8):
    """Open channel with no diffusion: every row moves by u * dt per step."""
    case = open_channel_case(u=u)
    case.position = np.stack(np.meshgrid(xs, ys), -1).reshape(-1, 2)
    case.pool = case.position.shape[0] + 4
    return case


def test_counts_uniform_flow_closed_form():
    """A row starting at x0 crosses the line x = X during step k when
    x0 + (k - 1) u dt < X < x0 + k u dt; rows leaving at x = 100 are not
    counted (the line lies upstream)."""
    xs = np.linspace(5.05, 60.05, 23)
    ys = np.linspace(10., 90., 9)
    case = uniform_channel(xs, ys)
    X, u, dt, nt = 47.37, 0.8, 1., 60
    secs = [np.array([[X, -5.], [X, 105.]])]
    params = dict(final_time=float(nt), time_step=dt, frozen_eulerian_fields=True, model=1,
                  diffusion_model=0, bnd_statistics=True)
    r = case.run_oracle(R, params, control_sections=secs)
    got = np.array(r.solver.csec_history)[:, 0]
    k = np.arange(1, nt + 1)[:, None]
    x0 = np.repeat(xs[None, :], 1, 0)
    expect = ((x0 + (k - 1) * u * dt < X) & (x0 + k * u * dt > X)).sum(1) * ys.size
    assert expect.sum() > 0
    np.testing.assert_array_equal(got, expect)


def test_deleted_rows_count_at_the_origin():
    """delete_single zeroes position and last_position, so a section through
    the origin counts each step's deletions; one that misses it counts none."""
    xs = np.linspace(60.05, 99.05, 14)
    case = uniform_channel(xs, np.linspace(10., 90., 5))
    secs = [np.array([[-1., -1.], [1., 1.]]), np.array([[0.5, -1.], [2., 1.]])]
    params = dict(final_time=60., time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=0,
                  bnd_statistics=True)
    r = case.run_oracle(R, params, control_sections=secs)
    cs = np.array(r.solver.csec_history)
    dels = np.array(r.solver.bnd_history).sum(1)
    assert dels.sum() == case.position.shape[0]
    np.testing.assert_array_equal(cs[:, 0], dels)
    assert cs[:, 1].sum() == 0


def test_only_the_first_segment_is_tested():
    """ControlSection counts poly.shape[1] - 1 segments, i.e. one for an
    (npoint, 2) polygon: a crossing of the second segment alone is not seen."""
    c = cy.ControlSection(np.array([[30., -5.], [30., 40.], [80., 40.], [80., 105.]]))
    assert c.nseg == 1
    np.testing.assert_array_equal(c.segments(), [[30., -5., 30., 40.]])
    case = uniform_channel(np.linspace(5.05, 60.05, 12), np.array([70., 80.]))
    params = dict(final_time=40., time_step=1., frozen_eulerian_fields=True, model=1, diffusion_model=0)
    r = case.run_oracle(R, params, control_sections=[c.poly])
    assert np.array(r.solver.csec_history).sum() == 0  # rows at y = 70, 80 miss [30,-5]-[30,40]
    nseg_of, seg = section_segments([c, c])
    np.testing.assert_array_equal(nseg_of, [1, 1])
    assert seg.shape == (8,)


def test_statistics_files_round_trip(tmp_path):
    """Header lines and "time, count" rows as write_statistics.pyx writes them,
    read back the way the reference's statistics example reads them."""
    cy.prepare_secstats_output(True, 2, str(tmp_path), "bnd_stats", True, 1, "cs_stats")
    times = np.array([10., 20., 30.])
    data = np.array([[1, 4], [0, 5], [7, 0]], dtype=np.int32)
    cy.write_secstats_buffer(times, data, 3, 2, str(tmp_path), "bnd_stats")
    cy.write_secstats_buffer(times, data, 2, 2, str(tmp_path), "bnd_stats")
    lines = (tmp_path / "bnd_stats_1.txt").read_text().splitlines()
    assert lines[:2] == ["cylag boundary statistics file", "Variables: Time, Npart"]
    assert lines[2] == "10.00000000, {:16d}  ".format(1)
    d = np.loadtxt(tmp_path / "bnd_stats_2.txt", delimiter=',', skiprows=2)
    np.testing.assert_array_equal(d[:, 1], [4, 5, 0, 4, 5])
    assert (tmp_path / "cs_stats_1.txt").read_text().startswith("cylag control section file\n")
