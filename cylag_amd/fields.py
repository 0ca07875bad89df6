"""Eulerian field set: mesh, time records and node fields
(cylag/core/eulerian_field_set.pyx:125-315, 606-622, 1770-1885).

Fields are (ntimes, npoints*nlayers), layer-major; optional fields are
(nopt, ntimes, npoints*nlayers) and resolved by name.  The step uploads every
record to the GPU once and time-interpolates on the device.
"""
import numpy as np

from .constants import EPSILON

_TURB = ("TURBULENT ENERG.", "TURBULENT ENERGY", "ENERGIE TURBUL.")


class EulerianFieldSet:
    def __init__(self, triangular_mesh, dim=2, nlayers=1, times=None, elevation=None,
                 velocity_x=None, velocity_y=None, velocity_z=None, opt_fields=None,
                 opt_fields_names=None, trifinder=None, friction_model=1, friction_coefficient=40.):
        self.triangular_mesh = triangular_mesh
        self.npoints = triangular_mesh.nv
        self.nlayers = nlayers
        self.dim = dim
        self.trifinder = trifinder
        self.times = np.ascontiguousarray(times, dtype=np.float64)
        self.ntimes = len(self.times)
        self.record = 0
        self.time_interp_factor = 1.
        if self.ntimes == 1:
            self.frozen_hydro = True
        else:
            self.frozen_hydro = False
            for i in range(self.ntimes - 1):
                dt = self.times[i + 1] - self.times[i]
                if dt <= EPSILON:
                    raise ValueError(
                        "Invalid times array: times must be strictly increasing. "
                        "Found times[{}]={:.6e} and times[{}]={:.6e} with dt={:.6e} "
                        "(must be > {:.6e}). Check input data for duplicate or reversed time "
                        "values.".format(i, self.times[i], i + 1, self.times[i + 1], dt, EPSILON))
        self.check_dim(elevation)
        self.check_dim(velocity_x)
        self.check_dim(velocity_y)
        if self.dim == 3:
            self.check_dim(velocity_z)
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.elevation = c(elevation)
        self.velocity_x = c(velocity_x)
        self.velocity_y = c(velocity_y)
        self.velocity_z = c(velocity_z)
        if self.dim == 3:
            self.check_regular_layers()
        self.field_names = ['ELEVATION', 'VELOCITY U', 'VELOCITY V', 'VELOCITY W']
        if opt_fields is not None:
            assert opt_fields_names is not None
            self.nopt = len(opt_fields_names)
        else:
            self.nopt = 0
        self.opt_fields = c(opt_fields)
        self.opt_fields_names = opt_fields_names
        self.fid_zs, self.fid_ux, self.fid_uy, self.fid_uz = 0, 1, 2, 3
        self.fid_turbeng = self.fid_dissip = self.fid_zb = self.fid_windx = self.fid_windy = -1
        if opt_fields_names is not None:
            names = list(opt_fields_names)
            pick = lambda cands: next((4 + names.index(n) for n in cands if n in names), -1)
            self.fid_turbeng = pick(_TURB)
            self.fid_dissip = pick(("DISSIPATION",))
            self.fid_zb = pick(("BOTTOM", "FOND"))
            self.fid_windx = pick(("WIND X", "VENT X"))
            self.fid_windy = pick(("WIND Y", "VENT Y"))
        self.friction_model = friction_model
        self.friction_coefficient = friction_coefficient

    @staticmethod
    def from_telemac2d(file_name, main_var_names=['FREE SURFACE', 'VELOCITY U', 'VELOCITY V'], bnd_file=None,
                       optional_fields=None, init_trifinder=True, friction_model=1, friction_coefficient=40.):
        """Field set of a TELEMAC-2D result file, times relative to its first
        record (eulerian_field_set.pyx:317-398)."""
        return EulerianFieldSet._from_telemac(file_name, 2, main_var_names, bnd_file, optional_fields,
                                              init_trifinder, friction_model, friction_coefficient)

    @staticmethod
    def from_telemac3d(file_name, main_var_names=['ELEVATION Z', 'VELOCITY U', 'VELOCITY V', 'VELOCITY W'],
                       bnd_file=None, optional_fields=None, init_trifinder=True, friction_model=1,
                       friction_coefficient=40.):
        """Field set of a TELEMAC-3D result file (plane-major fields, the
        bottom plane's triangulation; eulerian_field_set.pyx:400-483)."""
        return EulerianFieldSet._from_telemac(file_name, 3, main_var_names, bnd_file, optional_fields,
                                              init_trifinder, friction_model, friction_coefficient)

    @staticmethod
    def from_hdf5(fset_file="cylag_fset_file.h5", mesh_file="cylag_mesh_file.h5", optional_fields=None,
                  init_trifinder=True, friction_model=1, friction_coefficient=40.):
        """Field set from a CyLag field-set file and mesh file
        (eulerian_field_set.pyx:486-566)."""
        from .hdf5 import H5Reader
        from .mesh import TriangularMesh
        mesh = TriangularMesh.from_hdf5(file_name=mesh_file)
        trifinder = None
        if init_trifinder:
            import matplotlib.tri as mtri
            trifinder = mtri.Triangulation(mesh.x, mesh.y, mesh.triangles).get_trifinder()
        f = H5Reader(fset_file)
        dim, nplan = int(f["dim"][0]), int(f["nplan"][0])
        if dim not in (2, 3):
            raise ValueError("Wrong dimension in fset_file")
        velz = f["velocity_z"] if dim == 3 else None
        opt = None
        if int(f["nopt"][0]) != 0:
            if optional_fields is None:
                if "opt_fields_names" in f:
                    optional_fields = [n.decode("utf-8") if isinstance(n, bytes) else str(n)
                                       for n in f["opt_fields_names"]]
                else:
                    optional_fields = ["opt_field_{}".format(i) for i in range(int(f["nopt"][0]))]
            opt = f["opt_fields"]
        else:
            optional_fields = None
        return EulerianFieldSet(mesh, dim, nplan if dim == 3 else 1, f["times"], f["elevation"],
                                f["velocity_x"], f["velocity_y"], velz, opt, optional_fields, trifinder,
                                friction_model=friction_model, friction_coefficient=friction_coefficient)

    def save_hdf5(self, fset_file="cylag_fset_file.h5", mesh_file="cylag_mesh_file.h5"):
        """Write the mesh file and the field-set file (eulerian_field_set.pyx:568-604);
        the optional field names are stored as fixed-length byte strings."""
        from .hdf5 import write_h5
        self.triangular_mesh.save_mesh_hdf5(mesh_file)
        d = dict(dim=np.array([self.dim], dtype=np.int64), nplan=np.array([self.nlayers], dtype=np.int64),
                 times=np.asarray(self.times), elevation=np.asarray(self.elevation),
                 velocity_x=np.asarray(self.velocity_x), velocity_y=np.asarray(self.velocity_y),
                 nopt=np.array([self.nopt], dtype=np.int64))
        if self.dim == 3:
            d["velocity_z"] = np.asarray(self.velocity_z)
        if self.nopt != 0:
            d["opt_fields"] = np.asarray(self.opt_fields)
            d["opt_fields_names"] = [str(n) for n in self.opt_fields_names]
        write_h5(fset_file, d)

    @staticmethod
    def _from_telemac(file_name, dim, names, bnd_file, optional_fields, init_trifinder, friction_model,
                      friction_coefficient):
        from .mesh import TriangularMesh, read_cli
        from .telemac import SelafinFile
        res = SelafinFile(file_name)
        if bnd_file is not None:
            bnd_pts, nopen = read_cli(bnd_file)
        else:
            bnd_pts, nopen = None, 0
        tri = res.tri
        mesh = TriangularMesh.from_matplotlib(tri, boundary_points=bnd_pts, nopen=nopen)
        trifinder = tri.get_trifinder() if init_trifinder else None
        times = res.times - res.times[0]
        niter = res.ntimestep
        get = lambda name: np.stack([res.get_data_value(name, k) for k in range(niter)])
        elev, velx, vely = get(names[0]), get(names[1]), get(names[2])
        velz = get(names[3]) if dim == 3 else None
        opt = None
        if optional_fields is not None:
            opt = np.stack([get(name) for name in optional_fields])
        return EulerianFieldSet(mesh, dim, res.nplan if dim == 3 else 1, times, elev, velx, vely, velz, opt,
                                optional_fields, trifinder, friction_model=friction_model,
                                friction_coefficient=friction_coefficient)

    def check_dim(self, field):
        ref_shape = (self.ntimes, self.npoints) if self.dim == 2 else \
            (self.ntimes, self.npoints * self.nlayers)
        shape = np.shape(field)
        if shape != ref_shape:
            raise ValueError("Field have wrong dimension, expected {}, found {}".format(
                (self.ntimes, self.npoints), shape))

    def check_regular_layers(self, rtol=1e-3):
        """Uniform sigma layers at the first record (eulerian_field_set.pyx:281-314)."""
        nl = self.nlayers
        if self.dim != 3 or nl < 2:
            return
        frac = (np.arange(nl) / (nl - 1.))[:, None]
        z = np.asarray(self.elevation)[0].reshape(nl, self.npoints)
        zb = z[0]
        hl = (z[-1] - zb) / (nl - 1.)
        expected = zb[None, :] + frac * (z[-1] - zb)[None, :]
        tol = rtol * np.abs(hl) + 1e-6 * (1. + np.abs(zb))
        err = np.abs(z - expected) - tol[None, :]
        if np.any(err > 0.):
            l, i = np.unravel_index(np.argmax(err), err.shape)
            raise ValueError(
                "Vertical layers are not regularly spaced (CyLag requires uniform sigma layers): "
                "at first time record, node {}, layer {}, elevation is {:.6e}, expected "
                "{:.6e}.".format(i, l, z[l, i], expected[l, i]))

    # ------------------------------------------------------------------
    def field_by_id(self, field_id):
        """(ntimes, npoints*nlayers) array of a field id (eulerian_field_set.pyx:673-706)."""
        if field_id == 0:
            return self.elevation
        if field_id == 1:
            return self.velocity_x
        if field_id == 2:
            return self.velocity_y
        if field_id == 3:
            return self.velocity_z
        if 3 < field_id <= 3 + self.nopt:
            return self.opt_fields[field_id - 4]
        raise ValueError("Wrong field id")

    def get_mtri_triangulation(self):
        import matplotlib.tri as mtri
        m = self.triangular_mesh
        return mtri.Triangulation(m.x, m.y, m.triangles)

    def get_field_list(self):
        field_list = self.field_names + list(self.opt_fields_names or [])
        return field_list, list(range(len(field_list)))

    def _interp_record(self, field, time):
        """Field values at `time` (record search and lerp of
        eulerian_field_set.pyx:845-864,940-992)."""
        if self.frozen_hydro or self.ntimes == 1:
            return np.asarray(field[0])
        t = self.times
        rec = 0
        for i in range(self.ntimes - 1):
            if t[i] <= time < t[i + 1]:
                rec = i
                break
        else:
            if time >= t[-1]:
                rec = self.ntimes - 2
        tf = (time - t[rec + 1]) / (t[rec] - t[rec + 1])
        return tf * field[rec] + (1. - tf) * field[rec + 1]

    def get_field(self, field_id, time, layer=-1):
        """Field at a time and layer; pre/post-processing only (:1770-1795)."""
        res = self._interp_record(self.field_by_id(field_id), time)
        if self.dim == 2 or layer == -1:
            return np.asarray(res)
        return np.asarray(res[layer * self.npoints:(layer + 1) * self.npoints])

    def _p1(self, k, x, y):
        m = self.triangular_mesh
        f = m.det[k] + x * m.gradx[k] + y * m.grady[k]
        return m.triangles[k], f

    def interpolate_field(self, field_id, point, time, k=None):
        """P1 interpolation at a point (2D; :1842-1885)."""
        if k is None:
            k = int(self.trifinder(point[0], point[1]))
        if k == -1:
            raise ValueError("Failed interpolation (5), point not in triangulation")
        nodes, f = self._p1(k, point[0], point[1])
        fld = self.get_field(field_id, time, -1 if self.dim == 2 else 0)
        return fld[nodes[0]] * f[0] + fld[nodes[1]] * f[1] + fld[nodes[2]] * f[2]

    def interpolate_velocity(self, point, time, k=None):
        if k is None:
            k = int(self.trifinder(point[0], point[1]))
        if k == -1:
            raise ValueError("Failed interpolation (4), point not in triangulation")
        return (self.interpolate_field(1, point, time, k), self.interpolate_field(2, point, time, k))
