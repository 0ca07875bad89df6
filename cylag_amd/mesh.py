"""Triangular mesh and its connectivity tables (cylag/geom/triangular_mesh.pyx).

The tables are built with vectorized numpy but in the reference's exact
order, because walk order and boundary tie-breaking depend on it:
  * edges: sorted unique (min, max) vertex pairs (triangular_mesh.pyx:288-309);
  * triedge_map[k]: the three edge ids of triangle k in increasing order
    (the edge loop of triangular_mesh.pyx:445-488 fills them that way);
  * edgetri_map[e]: the triangles sharing edge e in increasing order, -1 if
    boundary (same loop, sorted merge of vertex adjacency lists);
  * edges_labels: -1 internal, 0 wall, >= 1 open boundary (:490-558);
  * det/gradx/grady: P1 basis coefficients with alpha = 0.5/area (:560-600).
"""
import numpy as np

from .constants import BND_WALL_REF, EPSILON, INTERN_REF, MAX_NUMBER_OF_OPEN_BND, MAX_VERTEX_ADJACENCY


class TriangularMesh:
    def __init__(self, x, y, triangles, edges=None, vertex_adjacency=None,
                 triangle_adjacency=None, triangle_neighbors=None, edgetri_map=None,
                 triedge_map=None, boundary_points=None, boundary_edges=None, vertex_labels=None,
                 edges_labels=None, signed_area=None, gradx=None, grady=None, det=None,
                 triangle_centroids=None, nopen=0):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int32)
        self.nv = self.x.shape[0]
        self.nt = self.triangles.shape[0]
        self.edges = (self._init_edges() if edges is None
                      else np.ascontiguousarray(edges, dtype=np.int32))
        self.ne = self.edges.shape[0]
        if vertex_adjacency is None:
            self.vertex_adjacency = self._init_vertex_adjacency()
            self.triangle_adjacency = None
            self.triangle_neighbors = None
            self._init_triangles_edges_connectivity()
        else:
            self.vertex_adjacency = vertex_adjacency
            self.triangle_adjacency = triangle_adjacency
            self.triangle_neighbors = triangle_neighbors
            assert edgetri_map is not None and triedge_map is not None
            self.edgetri_map = np.ascontiguousarray(edgetri_map, dtype=np.int32)
            self.triedge_map = np.ascontiguousarray(triedge_map, dtype=np.int32)
        self.nopen = nopen
        self.boundary_points = boundary_points
        self.boundary_edges = boundary_edges
        if vertex_labels is None or edges_labels is None:
            self.nvb = 0
            self._init_boundaries()
        else:
            self.vertex_labels = np.ascontiguousarray(vertex_labels, dtype=np.int32)
            self.edges_labels = np.ascontiguousarray(edges_labels, dtype=np.int32)
            self.nvb = np.shape(self.boundary_points)[0]
        if signed_area is None:
            self._init_mesh_properties()
        else:
            self.signed_area = np.ascontiguousarray(signed_area, dtype=np.float64)
            self.gradx = np.ascontiguousarray(gradx, dtype=np.float64)
            self.grady = np.ascontiguousarray(grady, dtype=np.float64)
            self.det = np.ascontiguousarray(det, dtype=np.float64)
        t = self.triangles
        if triangle_centroids is not None:
            self.triangle_centroids = np.ascontiguousarray(triangle_centroids, dtype=np.float64)
        else:
            self.triangle_centroids = np.stack(
                [(self.x[t[:, 0]] + self.x[t[:, 1]] + self.x[t[:, 2]]) / 3.0,
                 (self.y[t[:, 0]] + self.y[t[:, 1]] + self.y[t[:, 2]]) / 3.0], axis=1)
        # the centroid KD-tree of the reference is replaced by the device's bucket grid
        self.kdtree = None

    @staticmethod
    def from_matplotlib(tri, boundary_points=None, boundary_edges=None, nopen=0):
        """Mesh from a matplotlib Triangulation, keeping its edge numbering
        (triangular_mesh.pyx:181-202)."""
        return TriangularMesh(tri.x, tri.y, tri.triangles, edges=tri.edges,
                              boundary_points=boundary_points, boundary_edges=boundary_edges,
                              nopen=nopen)

    # the tables a mesh file holds, in the reference's names
    HDF5_TABLES = ("x", "y", "triangles", "edges", "vertex_adjacency", "triangle_adjacency",
                   "triangle_neighbors", "edgetri_map", "triedge_map", "boundary_points", "boundary_edges",
                   "vertex_labels", "edges_labels", "signed_area", "gradx", "grady", "det",
                   "triangle_centroids")

    @staticmethod
    def from_hdf5(file_name="cylag_mesh_file.h5"):
        """Mesh from a CyLag mesh file: the tables are taken as stored, not
        rebuilt (triangular_mesh.pyx:204-244)."""
        from .hdf5 import H5Reader
        f = H5Reader(file_name)
        t = {k: (f[k] if k in f else None) for k in TriangularMesh.HDF5_TABLES}
        return TriangularMesh(nopen=int(f["nopen"][0]), **t)

    def save_mesh_hdf5(self, file_name="cylag_mesh_file.h5"):
        """Write the mesh tables to a CyLag mesh file (triangular_mesh.pyx:246-286);
        the triangle adjacency tables only when present."""
        from .hdf5 import write_h5
        d = {k: np.asarray(getattr(self, k)) for k in self.HDF5_TABLES if getattr(self, k, None) is not None}
        d["nopen"] = np.array([self.nopen], dtype=np.int64)
        write_h5(file_name, d)

    def _sides(self):
        t = self.triangles
        a = t.reshape(-1)
        b = t[:, [1, 2, 0]].reshape(-1)
        return np.minimum(a, b), np.maximum(a, b)

    def _init_edges(self):
        lo, hi = self._sides()
        return np.ascontiguousarray(np.unique(np.stack([lo, hi], axis=1), axis=0), dtype=np.int32)

    def _init_vertex_adjacency(self):
        """(nv, MAX_VERTEX_ADJACENCY) table of the triangles of each vertex;
        the reference raises beyond MAX_VERTEX_ADJACENCY (triangular_mesh.pyx:311-367)."""
        counts = np.bincount(self.triangles.reshape(-1), minlength=self.nv)
        if counts.size and counts.max() > MAX_VERTEX_ADJACENCY:
            over = np.where(counts > MAX_VERTEX_ADJACENCY)[0]
            raise ValueError(
                "MAX_VERTEX_ADJACENCY overflow: {} vertices exceed limit.\n"
                "Current MAX_VERTEX_ADJACENCY: {}\nRequired MAX_VERTEX_ADJACENCY: {}\n"
                "Affected vertices ({} of {} shown): {}\n".format(
                    len(over), MAX_VERTEX_ADJACENCY, int(counts.max()), min(5, len(over)),
                    len(over), over[:5].tolist()))
        # triangles of each vertex in increasing order (the reference's loop
        # over triangles appends them that way), -1 past the last
        flat = self.triangles.reshape(-1).astype(np.int64)
        order = np.argsort(flat, kind="stable")
        vert = flat[order]
        first = np.concatenate([[0], np.cumsum(counts)[:-1]])
        slot = np.arange(flat.size) - first[vert]
        table = np.full((self.nv, MAX_VERTEX_ADJACENCY), -1, dtype=np.int32)
        table[vert, slot] = (order // 3).astype(np.int32)
        return table

    def _init_triangles_edges_connectivity(self):
        nt, ne, nv = self.nt, self.ne, self.nv
        lo, hi = self._sides()
        e0 = np.minimum(self.edges[:, 0], self.edges[:, 1]).astype(np.int64)
        e1 = np.maximum(self.edges[:, 0], self.edges[:, 1]).astype(np.int64)
        ekey = e0 * nv + e1   # edges may come in either vertex order (matplotlib)
        order = np.argsort(ekey, kind="stable")
        skey = lo.astype(np.int64) * nv + hi
        pos = np.searchsorted(ekey[order], skey)
        pos = np.minimum(pos, ne - 1)
        found = ekey[order][pos] == skey
        side_edge = np.where(found, order[pos], -1).astype(np.int64)
        tri_of_side = np.repeat(np.arange(nt, dtype=np.int64), 3)
        # triedge_map: the triangle's edge ids in increasing order, -1 padded
        te = side_edge.reshape(nt, 3).copy()
        te[te < 0] = np.iinfo(np.int64).max
        te.sort(axis=1)
        te[te == np.iinfo(np.int64).max] = -1
        self.triedge_map = np.ascontiguousarray(te, dtype=np.int32)
        # edgetri_map: sharing triangles in increasing order
        ok = side_edge >= 0
        se, st = side_edge[ok], tri_of_side[ok]
        o = np.lexsort((st, se))
        se, st = se[o], st[o]
        first = np.ones(se.size, bool)
        first[1:] = se[1:] != se[:-1]
        et = -np.ones((ne, 2), np.int32)
        et[se[first], 0] = st[first]
        sec = ~first
        et[se[sec], 1] = st[sec]
        self.edgetri_map = et

    def _init_boundaries(self):
        """Boundary points/edges and labels (triangular_mesh.pyx:490-558)."""
        bnd = self.edgetri_map[:, 1] == -1
        if self.boundary_points is None:
            be = self.edges[bnd]
            pts = np.stack([be.reshape(-1), np.full(be.size, BND_WALL_REF)], axis=1)
            self.boundary_points = np.asarray(np.unique(pts, axis=0), dtype=np.int32)
            self.boundary_edges = np.asarray(
                np.concatenate([be, np.full((be.shape[0], 1), BND_WALL_REF)], axis=1), dtype=np.int32)
            self.nvb = self.boundary_points.shape[0]
        else:
            bp = np.asarray(self.boundary_points, dtype=np.int32)
            self.boundary_points = bp
            self.nvb = bp.shape[0]
            nxt = np.roll(np.arange(self.nvb), -1)
            lab = np.where(bp[:, 1] == bp[nxt, 1], bp[:, 1], BND_WALL_REF)
            self.boundary_edges = np.stack([bp[:, 0], bp[nxt, 0], lab], axis=1).astype(np.int32)
        vl = np.full(self.nv, INTERN_REF, dtype=np.int32)
        bp = np.asarray(self.boundary_points, dtype=np.int32)
        vl[bp[:, 0]] = bp[:, 1]   # later rows win, as in the reference's loop
        self.vertex_labels = vl
        el = np.full(self.ne, INTERN_REF, dtype=np.int32)
        i0, i1 = self.edges[bnd, 0], self.edges[bnd, 1]
        el[bnd] = np.where(vl[i0] == vl[i1], vl[i0], BND_WALL_REF)
        self.edges_labels = el

    def _init_mesh_properties(self):
        t = self.triangles
        xa, ya = self.x[t[:, 0]], self.y[t[:, 0]]
        xb, yb = self.x[t[:, 1]], self.y[t[:, 1]]
        xc, yc = self.x[t[:, 2]], self.y[t[:, 2]]
        area = 0.5 * ((xb - xa) * (yc - ya) - (yb - ya) * (xc - xa))
        bad = np.abs(area) < EPSILON
        if np.any(bad):
            k = int(np.argmax(bad))
            raise ValueError("Degenerate triangle detected at index {}: vertices ({}, {}, {}), "
                             "area = {:.2e}. Check for duplicate or collinear vertices.".format(
                                 k, t[k, 0], t[k, 1], t[k, 2], area[k]))
        alpha = 0.5 / area
        self.signed_area = area
        self.gradx = np.ascontiguousarray(np.stack([(yb - yc) * alpha, (yc - ya) * alpha,
                                                    (ya - yb) * alpha], axis=1))
        self.grady = np.ascontiguousarray(np.stack([(xc - xb) * alpha, (xa - xc) * alpha,
                                                    (xb - xa) * alpha], axis=1))
        self.det = np.ascontiguousarray(np.stack([(xb * yc - xc * yb) * alpha,
                                                  (xc * ya - xa * yc) * alpha,
                                                  (xa * yb - xb * ya) * alpha], axis=1))


def read_cli(file_name, vertex_id_column=11):
    """Telemac .cli boundary file -> (boundary_points, number of open boundaries)
    (cylag/geom/read_cli.pyx:24-98)."""
    points, new_boundary, count = [], False, 0
    bnd_ids = np.arange(1, MAX_NUMBER_OF_OPEN_BND + 1, 1)
    with open(file_name, "r") as f:
        for line in f:
            if count > MAX_NUMBER_OF_OPEN_BND:
                raise ValueError("Exceeding maximum number of open boundary (= {})".format(
                    MAX_NUMBER_OF_OPEN_BND))
            w = line.split()
            if not w:
                continue
            if int(w[0]) == 2 and int(w[1]) == 2 and int(w[2]) == 2:
                label, new_boundary = BND_WALL_REF, False
            else:
                if not new_boundary:
                    count += 1
                label, new_boundary = bnd_ids[count - 1], True
            points.append([int(w[vertex_id_column]) - 1, label])
    bp = np.asarray(points).astype(np.int32)
    nb = bp.shape[0]
    if bp[0, 1] != BND_WALL_REF and bp[nb - 1, 1] != BND_WALL_REF:
        last = bp[nb - 1, 1]
        bp[bp[:, 1] == last, 1] = bp[0, 1]
    return bp, count
