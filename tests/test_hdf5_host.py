"""HDF5 mesh and field-set files without h5py: the reader against the mesh
file the reference's own test wrote from the TELEMAC flume
(tests/test_04_geom_mesh_io.py:38-50, tests/data/geo_flume.h5), which also
pins every connectivity table of the mesh built here from the same TELEMAC
and boundary files; write/read round trips of meshes and field sets."""
import numpy as np
import pytest

import cylag_amd as cy
from cylag_amd.hdf5 import H5Reader, write_h5
from tests.cases import vortex_3d_case, vortex_case

DATA = "tests/golden/reference_data/"


def flume_fset():
    return cy.EulerianFieldSet.from_telemac2d(DATA + "r2d_flume.slf", bnd_file=DATA + "geo_flume.cli")


def test_reference_mesh_file_matches_tables_built_here():
    f = H5Reader(DATA + "geo_flume.h5")
    m = flume_fset().triangular_mesh
    expected = {"x": np.float64, "y": np.float64, "triangles": np.int32, "edges": np.int32,
                "vertex_adjacency": np.int32, "edgetri_map": np.int32, "triedge_map": np.int32,
                "boundary_points": np.int32, "boundary_edges": np.int32, "vertex_labels": np.int32,
                "edges_labels": np.int32, "signed_area": np.float64, "gradx": np.float64,
                "grady": np.float64, "det": np.float64, "triangle_centroids": np.float64}
    assert sorted(f.keys()) == sorted(list(expected) + ["nopen"])
    for k, dt in expected.items():
        a = f[k]
        assert a.dtype == dt, k
        np.testing.assert_array_equal(a, np.asarray(getattr(m, k)), err_msg=k)
    assert f["nopen"].tolist() == [m.nopen] == [2]
    assert f["vertex_adjacency"].shape == (m.nv, cy.MAX_VERTEX_ADJACENCY)


def test_mesh_from_reference_file():
    m = cy.TriangularMesh.from_hdf5(DATA + "geo_flume.h5")
    ref = flume_fset().triangular_mesh
    assert (m.nv, m.nt, m.ne, m.nopen, m.nvb) == (ref.nv, ref.nt, ref.ne, ref.nopen, ref.nvb)
    for k in ("edges", "edgetri_map", "triedge_map", "edges_labels", "vertex_labels", "det"):
        np.testing.assert_array_equal(getattr(m, k), getattr(ref, k))


def test_mesh_round_trip(tmp_path):
    ref = flume_fset().triangular_mesh
    p = str(tmp_path / "mesh.h5")
    ref.save_mesh_hdf5(p)
    f = H5Reader(p)
    assert "triangle_adjacency" not in f  # not built here, so not written
    m = cy.TriangularMesh.from_hdf5(p)
    for k in cy.TriangularMesh.HDF5_TABLES:
        if getattr(ref, k) is not None:
            np.testing.assert_array_equal(getattr(m, k), getattr(ref, k), err_msg=k)
    assert m.nopen == ref.nopen


@pytest.mark.parametrize("case_fn", [vortex_case, vortex_3d_case], ids=["2d", "3d"])
def test_field_set_round_trip(tmp_path, case_fn):
    case = case_fn()
    fs = cy.EulerianFieldSet(cy.TriangularMesh(case.x, case.y, case.tri), **case.fields)
    fp, mp = str(tmp_path / "fset.h5"), str(tmp_path / "mesh.h5")
    fs.save_hdf5(fset_file=fp, mesh_file=mp)
    g = cy.EulerianFieldSet.from_hdf5(fset_file=fp, mesh_file=mp, init_trifinder=False)
    assert (g.dim, g.nlayers, g.nopt) == (fs.dim, fs.nlayers, fs.nopt)
    for k in ("times", "elevation", "velocity_x", "velocity_y", "opt_fields"):
        np.testing.assert_array_equal(getattr(g, k), getattr(fs, k), err_msg=k)
    if fs.dim == 3:
        np.testing.assert_array_equal(g.velocity_z, fs.velocity_z)
    assert list(g.opt_fields_names or []) == list(fs.opt_fields_names or [])
    assert (g.fid_turbeng, g.fid_dissip) == (fs.fid_turbeng, fs.fid_dissip)
    # names given by the caller win over the stored ones
    if fs.nopt:
        names = ["a%d" % i for i in range(fs.nopt)]
        assert cy.EulerianFieldSet.from_hdf5(fp, mp, optional_fields=names,
                                             init_trifinder=False).opt_fields_names == names


def test_writer_many_members_and_types(tmp_path):
    """More members than one symbol table node holds, every numeric kind the
    writer takes, empty arrays and strings."""
    rng = np.random.default_rng(3)
    d = {"m%02d" % i: rng.standard_normal((i + 1, 2)) for i in range(20)}
    d.update(i8=np.arange(-3, 3, dtype=np.int8), u16=np.arange(7, dtype=np.uint16),
             i64=np.array([-(2 ** 62), 5]), f4=np.linspace(0, 1, 5, dtype=np.float32),
             empty=np.zeros((0, 3)), names=["TURBULENT ENERG.", "DISSIPATION", ""],
             big_endian=np.arange(4, dtype=">f8"))
    p = str(tmp_path / "many.h5")
    write_h5(p, d)
    f = H5Reader(p)
    assert sorted(f.keys()) == sorted(d)
    for k, v in d.items():
        a = f[k]
        if k == "names":
            assert [x.decode() for x in a] == v
        else:
            np.testing.assert_array_equal(a, np.asarray(v))
            assert a.shape == np.asarray(v).shape


def test_reader_rejects_other_files(tmp_path):
    p = tmp_path / "x.h5"
    p.write_bytes(b"not an hdf5 file at all")
    with pytest.raises(ValueError):
        H5Reader(str(p))
    with pytest.raises(KeyError):
        H5Reader(DATA + "geo_flume.h5")["no_such_dataset"]


def test_source_from_field_dual_cells():
    """ParticleSourceFromField.init_positions walks each vertex's triangles
    (the vertex adjacency table): dual-cell areas against a direct sum over
    the triangles (particle_source.pyx:200-239)."""
    fs = flume_fset()
    m = fs.triangular_mesh
    area = np.zeros(m.nv)
    np.add.at(area, m.triangles.reshape(-1), np.repeat(np.abs(m.signed_area) / 3., 3))
    dens = np.asarray(fs.field_by_id(0)[fs.record])
    src = cy.ParticleSourceFromField(cy.AlwaysScheduler(), 0, 0, ">", -1e30, virtual_mass=0.05)
    src.loc = np.arange(m.nv)
    cy.random_utils.set_state(42, 54)
    pos = src.init_positions(fs, dim=2)
    inner = [k for k in range(m.nv) if m.vertex_labels[k] == cy.INTERN_REF and dens[k] != 0.]
    counts = [max(1, int(dens[k] / 0.05 * area[k])) for k in inner]
    assert pos.shape == (sum(counts), 2)
    s = 0
    for k, n in zip(inner, counts):
        r = np.hypot(pos[s:s + n, 0] - m.x[k], pos[s:s + n, 1] - m.y[k])
        assert r.max() <= np.sqrt(area[k] / np.pi) * (1 + 1e-12)
        s += n


def _heap_free_list(f):
    """(data segment size, free-list head, next of the head block) of the root group's local heap."""
    b = f.buf
    root = int.from_bytes(b[24 + 4 * 8 + 8:24 + 4 * 8 + 16], "little")
    for t, d, _ in f._messages(root):
        if t == 0x11:
            hp = int.from_bytes(b[d + 8:d + 16], "little")
            size, head = int.from_bytes(b[hp + 8:hp + 16], "little"), int.from_bytes(b[hp + 16:hp + 24], "little")
            data = int.from_bytes(b[hp + 24:hp + 32], "little")
            return size, head, int.from_bytes(b[data + head:data + head + 8], "little")
    raise AssertionError("no symbol table")


def test_writer_heap_free_list_as_library(tmp_path):
    """The HDF5 library rejects a local heap whose free-list head is not
    inside the data segment, and ends the list with offset 1: the reference's
    file (written by the library) and ours agree on both."""
    size, head, nxt = _heap_free_list(H5Reader(DATA + "geo_flume.h5"))
    assert head < size and nxt == 1
    p = str(tmp_path / "mesh.h5")
    flume_fset().triangular_mesh.save_mesh_hdf5(p)
    size, head, nxt = _heap_free_list(H5Reader(p))
    assert head + 16 <= size and nxt == 1
