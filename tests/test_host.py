"""Host-side product logic (no GPU): mesh tables in the reference's order,
parameters, clock, the C ABI library's exports and its host RNG functions."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cylag_amd as cy
from cylag_amd import _hip
from oracle import refpy as R
from tests.cases import GOLDEN, jittered_delaunay, structured_square

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _meshes():
    x = np.loadtxt(os.path.join(GOLDEN, "square_meshx.txt"))
    y = np.loadtxt(os.path.join(GOLDEN, "square_meshy.txt"))
    t = np.loadtxt(os.path.join(GOLDEN, "square_meshtri.txt"), dtype=np.int32)
    yield "reference_square", x, y, t, None
    x = np.loadtxt(os.path.join(GOLDEN, "test_triangulation_meshx.txt"))
    y = np.loadtxt(os.path.join(GOLDEN, "test_triangulation_meshy.txt"))
    t = np.loadtxt(os.path.join(GOLDEN, "test_triangulation_meshtri.txt"), dtype=np.int32)
    yield "test_triangulation", x, y, t, None
    x, y, t = structured_square(21)
    yield "structured", x, y, t, None
    x, y, t, mpl = jittered_delaunay(30, 12, 3000., 600.)
    yield "delaunay", x, y, t, None
    yield "delaunay_mpl_edges", x, y, t, mpl.edges


@pytest.mark.parametrize("name,x,y,t,edges", list(_meshes()), ids=lambda v: v if isinstance(v, str) else "")
def test_mesh_tables_match_oracle_builder(name, x, y, t, edges):
    m = cy.TriangularMesh(x, y, t, edges=edges)
    r = R.RefMesh(x, y, t, edges=edges)
    assert np.array_equal(m.edges, r.edges)
    assert np.array_equal(m.edgetri_map, r.edgetri)
    assert np.array_equal(m.triedge_map, r.triedge)
    assert np.array_equal(m.edges_labels, r.edges_labels)
    assert np.array_equal(m.det, r.det) and np.array_equal(m.gradx, r.gradx)
    assert np.array_equal(m.grady, r.grady)
    assert np.array_equal(m.triangle_centroids, r.centroids)


def test_mesh_labels_from_cli():
    x = np.loadtxt(os.path.join(GOLDEN, "square_meshx.txt"))
    y = np.loadtxt(os.path.join(GOLDEN, "square_meshy.txt"))
    t = np.loadtxt(os.path.join(GOLDEN, "square_meshtri.txt"), dtype=np.int32)
    bp, nb = cy.read_cli(os.path.join(GOLDEN, "square_geo.cli"))
    bpr, nbr = R.read_cli(os.path.join(GOLDEN, "square_geo.cli"))
    assert nb == nbr and np.array_equal(bp, bpr)
    m = cy.TriangularMesh(x, y, t, boundary_points=bp)
    r = R.RefMesh(x, y, t, boundary_points=bpr)
    assert np.array_equal(m.edges_labels, r.edges_labels)


def test_degenerate_triangle_raises():
    # tests/test_00_geom_mesh.py:376-391
    x = np.array([0., 1., 2., 0.])
    y = np.array([0., 0., 0., 1.])
    t = np.array([[0, 1, 2], [0, 1, 3]], dtype=np.int32)
    with pytest.raises(ValueError, match="Degenerate triangle"):
        cy.TriangularMesh(x, y, t)


def test_vertex_adjacency_overflow_raises():
    # tests/test_00_geom_mesh.py:446-458: a fan of 16 triangles around one vertex
    ang = np.linspace(0, 2 * np.pi, 17)[:-1]
    x = np.concatenate([[0.], np.cos(ang)])
    y = np.concatenate([[0.], np.sin(ang)])
    t = np.array([[0, 1 + i, 1 + (i + 1) % 16] for i in range(16)], dtype=np.int32)
    with pytest.raises(ValueError, match="MAX_VERTEX_ADJACENCY"):
        cy.TriangularMesh(x, y, t)


def test_parameters_semantics():
    p = cy.Parameters({'model': 1, 'diffusion_model': 1})
    assert p.time_scheme == 1
    assert cy.Parameters({'model': 1}).time_scheme == 4
    assert cy.Parameters({'model': 2}).time_scheme == 5
    assert cy.Parameters({'frozen_eulerian_fields': True, 'initial_time': 5.}).initial_time == 0.
    with pytest.raises(TypeError):
        cy.Parameters({'final_time': 3600})   # int where the default is a float
    with pytest.raises(TypeError):
        cy.Parameters({'no_such_key': 1})
    with pytest.raises(ValueError):
        cy.Parameters({'rng_method': 4}).check_values()
    with pytest.raises(ValueError):
        cy.Parameters({'model': 1, 'time_scheme': 5}).check_values()


def test_simutime():
    st = cy.SimuTime(False, 0., 10., 0.1)
    assert st.nt == 100
    for _ in range(99):
        st.increment()
    with pytest.raises(ValueError):
        cy.SimuTime(False, 1., 1., 0.1)


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "cylag_hip.h")).read()
    return sorted(set(re.findall(r"\b(cyl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_hip.LIB_PATH)
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_hip.SIGNATURES)


def test_host_rng_matches_oracle():
    cy.pcg32_seed(42, 54)
    assert [cy.random_utils.pcg32_next() for _ in range(6)] == [
        0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    for method in (1, 2, 3):
        cy.pcg32_seed(9, 77)
        g = R.RefRng(9, 77)
        for _ in range(50):
            assert cy.generate_random_pair(method) == g.pair(method)
        assert cy.random_utils.get_state() == g.state()
    cy.pcg32_seed(42, 0xda3e06d11bb2b203)


def test_crmath_tables_match_generator():
    """The double-double constants of the device exp/log are the generator's output."""
    import subprocess
    import sys
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "cylag_amd", "csrc")
    out = subprocess.run([sys.executable, os.path.join(here, "gen_crmath_tables.py")], capture_output=True,
                         text=True, check=True).stdout
    with open(os.path.join(here, "cyl_crmath_tables.h")) as f:
        assert f.read() == out
