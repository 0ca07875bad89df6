"""ctypes driver for the CPU restatement in oracle/cpu_ref.c.

TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module; the
product package cylag_amd never does.

The classes mirror the reference objects just enough to drive the serial step:
``RefMesh`` (cylag/geom/triangular_mesh.pyx:75-179), ``RefFields``
(cylag/core/eulerian_field_set.pyx:125-256), ``RefParticles``
(cylag/core/lagrangian_particle_set.pyx:158-261) and ``RefSolver``
(cylag/core/solver.pyx:74-236,382-585).  The global PCG32 stream of
cylag/lsm/random_utils.pyx:17-20 is the module-level ``RNG`` object, seeded at
import with (42, 0xda3e06d11bb2b203) like cylag/lsm/__init__.py:6-10.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()


class Pcg(C.Structure):
    _fields_ = [("state", C.c_uint64), ("inc", C.c_uint64)]


class MeshS(C.Structure):
    _fields_ = [("nv", C.c_int), ("nt", C.c_int), ("ne", C.c_int), ("nopen", C.c_int),
                ("x", _dp), ("y", _dp), ("tri", _ip), ("edges", _ip), ("edgetri", _ip),
                ("triedge", _ip), ("edge_labels", _ip), ("det", _dp), ("gradx", _dp),
                ("grady", _dp), ("centroids", _dp)]


class FsetS(C.Structure):
    _fields_ = [("dim", C.c_int), ("nlayers", C.c_int), ("npoints", C.c_int), ("ntimes", C.c_int),
                ("nopt", C.c_int), ("times", _dp), ("elev", _dp), ("u", _dp), ("v", _dp),
                ("w", _dp), ("opt", _dp), ("fid_zs", C.c_int), ("fid_ux", C.c_int),
                ("fid_uy", C.c_int), ("fid_uz", C.c_int), ("fid_turbeng", C.c_int),
                ("fid_dissip", C.c_int), ("fid_zb", C.c_int), ("friction_model", C.c_int),
                ("friction_coefficient", C.c_double),
                ("tmp_elev", _dp), ("tmp_u", _dp), ("tmp_v", _dp), ("tmp_w", _dp),
                ("tmp_z0", _dp), ("tmp_u0", _dp), ("tmp_v0", _dp), ("tmp_w0", _dp),
                ("tmp_opt", _dp), ("frozen", C.c_int), ("record", C.c_int),
                ("interp_done", C.c_int), ("prev_valid", C.c_int), ("c_valid", C.c_int),
                ("c_tri", C.c_int), ("c_x", C.c_double), ("c_y", C.c_double),
                ("c_f0", C.c_double), ("c_f1", C.c_double), ("c_f2", C.c_double)]


class PsetS(C.Structure):
    _fields_ = [("npart", C.c_int), ("dim", C.c_int), ("npart_active", C.c_int),
                ("maxtag", C.c_int), ("inactive", _ip), ("tri", _ip), ("lowerlayer", _ip),
                ("last_tri", _ip), ("tag", _ip), ("position", _dp), ("velocity", _dp),
                ("us", _dp), ("last_position", _dp), ("last_velocity", _dp),
                ("diameter", _dp), ("volume", _dp), ("depth", _dp), ("compute_depth", C.c_int),
                ("rho0", C.c_double), ("density", C.c_double), ("drag_coef", C.c_double),
                ("added_mass_coef", C.c_double), ("buoyancy_velocity_value", C.c_double),
                ("drag_coef_model", C.c_int), ("added_mass_force", C.c_int),
                ("buoyancy_velocity_model", C.c_int), ("a0c", C.c_double), ("a1c", C.c_double),
                ("a2c", C.c_double), ("a3c", C.c_double), ("max_stranded", C.c_int),
                ("stranded_count", C.c_int), ("stranded_positions", _dp),
                ("max_deposited", C.c_int), ("deposited_count", C.c_int), ("deposited_positions", _dp),
                ("deposited_iteration", C.POINTER(C.c_int))]


class ParS(C.Structure):
    _fields_ = [("model", C.c_int), ("time_scheme", C.c_int),
                ("particle_velocity_init", C.c_int), ("rng_method", C.c_int),
                ("lsm3_edge_opt", C.c_int), ("water_density", C.c_double),
                ("diffusion_model", C.c_int), ("horizontal_diffusivity", C.c_double),
                ("vertical_diffusivity", C.c_double), ("schmidt_number", C.c_double),
                ("diffusion_lsm2_option", C.c_int), ("diffusion_lsm3_tl_horizontal", C.c_double),
                ("diffusion_lsm3_tl_vertical", C.c_double), ("boundary_conditions", C.c_int),
                ("boundary_conditions_type", C.c_int), ("max_wall_rebounds", C.c_int),
                ("rebound_damping_coef", C.c_double), ("wall_roughness_angle", C.c_double),
                ("stranding_probability", C.c_double), ("bnd_statistics", C.c_int),
                ("stranding_output", C.c_int), ("bottom_deposition", C.c_int)]


class SolverS(C.Structure):
    _fields_ = [("mesh", C.POINTER(MeshS)), ("fset", C.POINTER(FsetS)),
                ("pset", C.POINTER(PsetS)), ("par", C.POINTER(ParS)), ("rng", Pcg),
                ("time", C.c_double), ("time_step", C.c_double), ("iteration", C.c_int),
                ("bnd_stats", _ip), ("init_tri", _ip), ("has_trifinder", C.c_int),
                ("err_code", C.c_int), ("err_particle", C.c_int), ("ncsec", C.c_int),
                ("csec_nseg", _ip), ("csec_seg", _dp), ("csec_stats", _ip)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_PP = C.POINTER(Pcg)
_sig("ref_pcg32_seed", None, [_PP, C.c_uint64, C.c_uint64])
_sig("ref_pcg32_advance", None, [_PP, C.c_uint64])
_sig("ref_pcg32_next", C.c_uint32, [_PP])
_sig("ref_random_uniform", C.c_double, [_PP])
_sig("ref_random_gaussian", C.c_double, [_PP])
_sig("ref_random_gaussian_pair", None, [_PP, _dp, _dp])
_sig("ref_random_gaussian_ziggurat", C.c_double, [_PP])
_sig("ref_ziggurat_tables", None, [C.POINTER(C.c_uint32), _dp, _dp])
_sig("ref_generate_random_pair", C.c_int, [_PP, C.c_int, _dp, _dp])
_sig("ref_generate_random_single", C.c_int, [_PP, C.c_int, _dp])
_sig("ref_fill_uint32", None, [_PP, C.c_int64, C.POINTER(C.c_uint32)])
_sig("ref_fill_normals", C.c_int, [_PP, C.c_int, C.c_int64, _dp])
_sig("ref_build_edges", C.c_int, [C.c_int, _ip, _ip])
_sig("ref_build_connectivity", C.c_int, [C.c_int, C.c_int, _ip, C.c_int, _ip, _ip, _ip])
_sig("ref_mesh_properties", C.c_int, [C.c_int, _ip, _dp, _dp, _dp, _dp, _dp, _dp])
_sig("ref_edge_labels", None, [C.c_int, _ip, _ip, _ip, _ip])
_MP = C.POINTER(MeshS)
_sig("ref_point_in_tri", C.c_int, [_MP, C.c_int, C.c_double, C.c_double])
_sig("ref_xy_localize", C.c_int, [_MP, C.c_int, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int, _ip])
_sig("ref_localize_kdtree", C.c_int, [_MP, C.c_double, C.c_double, C.c_int])
_sig("ref_localize_brute", C.c_int, [_MP, C.c_double, C.c_double])
_sig("ref_locate_candidates", None, [_MP, C.c_int, _dp, _dp, C.c_int, C.POINTER(C.c_longlong), _ip])
_sig("ref_intersec2d", C.c_int, [_dp, _dp, _dp, _dp])
_sig("ref_symmetric_point2d", None, [_dp, _dp, _dp, C.c_double, _dp])
_sig("ref_lower_layer_index", C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double])
_sig("ref_lower_index", C.c_int, [_dp, C.c_int, C.c_double, C.c_int])
_sig("ref_init_tmp", None, [C.POINTER(FsetS)])
_sig("ref_time_interpolation", C.c_int, [C.POINTER(FsetS), C.c_double])
_sig("ref_diffusivity_gradient", C.c_int, [C.POINTER(FsetS), _MP, C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp])
_sig("ref_lsm2_direct_coefficients", C.c_int, [C.c_double, C.c_double] + [_dp] * 7)
_sig("ref_lsm3_direct_coefficients", C.c_int, [C.c_double] * 4 + [C.c_int] + [_dp] * 3)
_SP = C.POINTER(SolverS)
_sig("ref_initialize_particle_state", C.c_int, [_SP, C.c_int])
_sig("ref_forward", C.c_int, [_SP])
_sig("ref_forward_n", C.c_int, [_SP, C.c_int])
_sig("ref_initialize_model_coefficients", None, [C.POINTER(PsetS)])
_sig("ref_buoyancy_velocity", C.c_double, [C.POINTER(PsetS), C.c_double, C.c_int, C.c_double])
_sig("ref_drag_coefficient", C.c_double, [C.POINTER(PsetS), C.c_double, C.c_double])
_sig("ref_interp_acceleration_2d", C.c_int, [C.POINTER(FsetS), _MP, C.c_int] + [C.c_double] * 5 + [_dp, _dp])
_sig("ref_interp_acceleration_3d", C.c_int, [C.POINTER(FsetS), _MP, C.c_int, C.c_int] + [C.c_double] * 7
     + [_dp, _dp, _dp])


def P(a, kind):
    if a is None:
        return C.cast(None, _dp if kind == "d" else _ip)
    return a.ctypes.data_as(_dp if kind == "d" else _ip)


# --------------------------------------------------------------------------------------
# RNG (random_utils.pyx)
# --------------------------------------------------------------------------------------
class RefRng:
    """A PCG32 stream with the reference's module API."""

    def __init__(self, seed=42, stream=0xda3e06d11bb2b203):
        self.g = Pcg()
        self.seed(seed, stream)

    def seed(self, seed, stream):
        lib.ref_pcg32_seed(C.byref(self.g), seed & (2**64 - 1), stream & (2**64 - 1))

    def advance(self, delta):
        lib.ref_pcg32_advance(C.byref(self.g), delta & (2**64 - 1))

    def next_u32(self):
        return lib.ref_pcg32_next(C.byref(self.g))

    def uniform(self):
        return lib.ref_random_uniform(C.byref(self.g))

    def state(self):
        return int(self.g.state), int(self.g.inc)

    def set_state(self, state, inc):
        self.g.state = state
        self.g.inc = inc

    def u32(self, n):
        out = np.empty(n, dtype=np.uint32)
        lib.ref_fill_uint32(C.byref(self.g), n, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out

    def normals(self, method, n):
        out = np.empty(n, dtype=np.float64)
        e = lib.ref_fill_normals(C.byref(self.g), method, n, P(out, "d"))
        if e:
            raise ValueError("Unknown RNG method")
        return out

    def pair(self, method):
        a, b = C.c_double(), C.c_double()
        if lib.ref_generate_random_pair(C.byref(self.g), method, C.byref(a), C.byref(b)):
            raise ValueError("Unknown RNG method")
        return a.value, b.value


RNG = RefRng()


def lsm3_direct_coefficients(a1, tl, dt, a2=0., edge_opt=1):
    """(transition (3, 3), unit-noise covariance (3, 3), force response (3,)) in
    the order (Us, Up, Xp) (update_state_lsm3.pyx:449-518)."""
    tr, cov, resp = np.zeros(9), np.zeros(9), np.zeros(3)
    e = lib.ref_lsm3_direct_coefficients(a1, tl, dt, a2, edge_opt, P(tr, "d"), P(cov, "d"), P(resp, "d"))
    if e:
        raise ValueError("LSM3 coefficients: error {}".format(e))
    return tr.reshape(3, 3), cov.reshape(3, 3), resp


def ziggurat_tables():
    kn = np.empty(128, np.uint32)
    wn = np.empty(128)
    fn = np.empty(128)
    lib.ref_ziggurat_tables(kn.ctypes.data_as(C.POINTER(C.c_uint32)), P(wn, "d"), P(fn, "d"))
    return kn, wn, fn


# --------------------------------------------------------------------------------------
# Mesh (triangular_mesh.pyx)
# --------------------------------------------------------------------------------------
class RefMesh:
    def __init__(self, x, y, triangles, edges=None, boundary_points=None, nopen=0):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.triangles = np.ascontiguousarray(triangles, dtype=np.int32)
        self.nv, self.nt = self.x.shape[0], self.triangles.shape[0]
        if edges is None:
            buf = np.empty((3 * self.nt, 2), dtype=np.int32)
            ne = lib.ref_build_edges(self.nt, P(self.triangles, "i"), P(buf, "i"))
            self.edges = np.ascontiguousarray(buf[:ne])
        else:
            self.edges = np.ascontiguousarray(edges, dtype=np.int32)
        self.ne = self.edges.shape[0]
        self.edgetri = np.empty((self.ne, 2), np.int32)
        self.triedge = np.empty((self.nt, 3), np.int32)
        e = lib.ref_build_connectivity(self.nv, self.nt, P(self.triangles, "i"), self.ne,
                                       P(self.edges, "i"), P(self.edgetri, "i"),
                                       P(self.triedge, "i"))
        if e:
            raise ValueError("MAX_VERTEX_ADJACENCY overflow")
        # _init_boundaries, triangular_mesh.pyx:490-558
        self.nopen = nopen
        vlab = -np.ones(self.nv, np.int32)
        if boundary_points is None:
            bnd = self.edges[self.edgetri[:, 1] == -1]
            vlab[bnd.ravel()] = 0
        else:
            bp = np.asarray(boundary_points, dtype=np.int32)
            for j, lab in bp:
                vlab[j] = lab
        self.vertex_labels = vlab
        self.edges_labels = np.empty(self.ne, np.int32)
        lib.ref_edge_labels(self.ne, P(self.edges, "i"), P(self.edgetri, "i"), P(vlab, "i"),
                            P(self.edges_labels, "i"))
        self.area = np.empty(self.nt)
        self.det = np.empty((self.nt, 3))
        self.gradx = np.empty((self.nt, 3))
        self.grady = np.empty((self.nt, 3))
        e = lib.ref_mesh_properties(self.nt, P(self.triangles, "i"), P(self.x, "d"), P(self.y, "d"),
                                    P(self.area, "d"), P(self.det, "d"), P(self.gradx, "d"),
                                    P(self.grady, "d"))
        if e:
            raise ValueError("Degenerate triangle detected at index {}".format(-1 - e))
        t = self.triangles
        # _build_kdtree, triangular_mesh.pyx:617-629
        self.centroids = np.ascontiguousarray(np.stack(
            [(self.x[t[:, 0]] + self.x[t[:, 1]] + self.x[t[:, 2]]) / 3.0,
             (self.y[t[:, 0]] + self.y[t[:, 1]] + self.y[t[:, 2]]) / 3.0], axis=1))
        self.s = MeshS(self.nv, self.nt, self.ne, nopen, P(self.x, "d"), P(self.y, "d"),
                       P(self.triangles, "i"), P(self.edges, "i"), P(self.edgetri, "i"),
                       P(self.triedge, "i"), P(self.edges_labels, "i"), P(self.det, "d"),
                       P(self.gradx, "d"), P(self.grady, "d"), P(self.centroids, "d"))

    def localize(self, k, p0x, p0y, x, y, method=3):
        b = C.c_int()
        r = lib.ref_xy_localize(C.byref(self.s), k, p0x, p0y, x, y, method, C.byref(b))
        return r, b.value

    def kdtree(self, x, y):
        return lib.ref_localize_kdtree(C.byref(self.s), x, y, 10)

    def kdtree_locator(self):
        """Callable (xs, ys) -> element ids by xy_localize_point_kdtree
        (xylocalizer.pyx:200-286), the initial search of a field set without a
        trifinder (solver.pyx:262-268), over scipy's cKDTree of the centroids as
        the reference builds it (triangular_mesh.pyx:617-629): the 10 nearest
        centroids are tested in cKDTree's order with point_in_tri; points none
        of them holds take the oracle's own expanding search."""
        from scipy.spatial import cKDTree
        tree = cKDTree(self.centroids)

        def locate(xs, ys):
            xs, ys = np.atleast_1d(np.asarray(xs, float)), np.atleast_1d(np.asarray(ys, float))
            out = np.full(xs.shape[0], -1, np.int32)
            if not xs.size:
                return out
            k = min(10, self.nt)
            _, cand = tree.query(np.stack([xs, ys], 1), k=k)
            cand = np.ascontiguousarray(cand.reshape(xs.shape[0], k), dtype=np.int64)
            xs, ys = np.ascontiguousarray(xs), np.ascontiguousarray(ys)
            lib.ref_locate_candidates(C.byref(self.s), xs.shape[0], P(xs, "d"), P(ys, "d"), k,
                                      cand.ctypes.data_as(C.POINTER(C.c_longlong)), P(out, "i"))
            return out

        return locate

    def brute(self, x, y):
        return lib.ref_localize_brute(C.byref(self.s), x, y)


# --------------------------------------------------------------------------------------
# Field set (eulerian_field_set.pyx)
# --------------------------------------------------------------------------------------
_TURB = ("TURBULENT ENERG.", "TURBULENT ENERGY", "ENERGIE TURBUL.")


class RefFields:
    def __init__(self, mesh, dim=2, nlayers=1, times=None, elevation=None, velocity_x=None,
                 velocity_y=None, velocity_z=None, opt_fields=None, opt_fields_names=None,
                 friction_model=1, friction_coefficient=40.):
        self.mesh = mesh
        self.dim, self.nlayers, self.npoints = dim, nlayers, mesh.nv
        self.times = np.ascontiguousarray(times, dtype=np.float64)
        self.ntimes = len(self.times)
        c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.elev, self.u, self.v, self.w = c(elevation), c(velocity_x), c(velocity_y), c(velocity_z)
        self.opt = c(opt_fields)
        names = list(opt_fields_names) if opt_fields_names is not None else []
        self.nopt = len(names) if opt_fields is not None else 0
        fid = lambda cands: next((4 + names.index(n) for n in cands if n in names), -1)
        size = self.npoints * nlayers
        self.tmp = {k: np.zeros(size) for k in ("elev", "u", "v", "w", "z0", "u0", "v0", "w0")}
        self.tmp_opt = np.zeros(max(1, self.nopt) * size)
        T = self.tmp
        self.s = FsetS(dim, nlayers, self.npoints, self.ntimes, self.nopt, P(self.times, "d"),
                       P(self.elev, "d"), P(self.u, "d"), P(self.v, "d"), P(self.w, "d"),
                       P(self.opt, "d"), 0, 1, 2, 3, fid(_TURB), fid(("DISSIPATION",)),
                       fid(("BOTTOM", "FOND")), friction_model, friction_coefficient,
                       P(T["elev"], "d"), P(T["u"], "d"), P(T["v"], "d"),
                       P(T["w"] if dim == 3 else None, "d"), P(T["z0"], "d"), P(T["u0"], "d"),
                       P(T["v0"], "d"), P(T["w0"] if dim == 3 else None, "d"),
                       P(self.tmp_opt, "d"), 1 if self.ntimes == 1 else 0, 0, 0, 0, 0, -1,
                       0., 0., 0., 0., 0.)
        lib.ref_init_tmp(C.byref(self.s))


# --------------------------------------------------------------------------------------
# Particle set (lagrangian_particle_set.pyx)
# --------------------------------------------------------------------------------------
class RefParticles:
    def __init__(self, position=None, velocity=None, fluid_velocity_seen=None, dim=2,
                 initial_pool_size=None, particle_density=1000., particle_diameter=0.01,
                 drag_coefficient_model=1, drag_coefficient=0.44, added_mass_force=False,
                 added_mass_coef=0.5, buoyancy_velocity_model=0, buoyancy_velocity=0.0,
                 compute_depth=False, diameter=None):
        if position is not None and initial_pool_size is None:
            n = np.shape(position)[0]
        else:
            n = initial_pool_size if initial_pool_size is not None else 10000
        self.npart, self.dim = n, dim
        na = 0 if position is None else np.shape(position)[0]
        self.inactive = np.ones(n, np.int32)
        self.inactive[:na] = 0
        self.position = np.zeros((n, dim))
        self.velocity = np.zeros((n, dim))
        self.us = np.zeros((n, dim))
        if position is not None:
            self.position[:na] = position
        if velocity is not None:
            self.velocity[:na] = velocity
        if fluid_velocity_seen is not None:
            self.us[:na] = fluid_velocity_seen
        self.particle_diameter = float(particle_diameter)
        self.diameter = np.full(n, float(particle_diameter))
        if diameter is not None:
            self.diameter[:] = diameter
        self.volume = np.pi * (self.diameter ** 3) / 6.
        self.tag = np.arange(n, dtype=np.int32)
        self.tri = -np.ones(n, np.int32)
        self.lowerlayer = np.zeros(n, np.int32)
        self.last_tri = -np.ones(n, np.int32)
        self.last_position = self.position.copy()
        self.last_velocity = self.velocity.copy()
        self.depth = np.zeros(n) if compute_depth else None
        self.stranded = np.zeros((n, 3))
        self.deposited = np.zeros((n, 3))
        self.deposited_iteration = np.zeros(n, np.int32)
        self.s = PsetS(n, dim, na, n - 1, P(self.inactive, "i"), P(self.tri, "i"),
                       P(self.lowerlayer, "i"), P(self.last_tri, "i"), P(self.tag, "i"),
                       P(self.position, "d"), P(self.velocity, "d"), P(self.us, "d"),
                       P(self.last_position, "d"), P(self.last_velocity, "d"),
                       P(self.diameter, "d"), P(self.volume, "d"), P(self.depth, "d"),
                       int(compute_depth), 1000., particle_density, drag_coefficient,
                       added_mass_coef, buoyancy_velocity, drag_coefficient_model,
                       int(added_mass_force), buoyancy_velocity_model, 0., 0., 0., 0., n, 0,
                       P(self.stranded, "d"), n, 0, P(self.deposited, "d"), P(self.deposited_iteration, "i"))

    @property
    def npart_active(self):
        return self.s.npart_active


# --------------------------------------------------------------------------------------
# Parameters + solver (parameters.pyx, solver.pyx)
# --------------------------------------------------------------------------------------
DEFAULTS = dict(initial_time=0., final_time=3600., time_step=1., frozen_eulerian_fields=False,
                model=1, time_scheme=1, particle_velocity_init=1, rng_method=2, lsm3_edge_opt=2,
                water_density=1000., diffusion_model=0, horizontal_diffusivity=0.,
                vertical_diffusivity=0., schmidt_number=0.72, diffusion_lsm2_option=1,
                diffusion_lsm3_tl_horizontal=1., diffusion_lsm3_tl_vertical=1.,
                boundary_conditions=True, boundary_conditions_type=1, max_wall_rebounds=4,
                rebound_damping_coef=0.5, wall_roughness_angle=15.0, stranding_probability=0.,
                bnd_statistics=False, stranding_output=False, bottom_deposition=False)


def resolve_params(user):
    """Default merge and time_scheme default (parameters.pyx:116-195)."""
    p = dict(DEFAULTS)
    p.update({k: v for k, v in user.items() if k in DEFAULTS})
    if p["frozen_eulerian_fields"]:
        p["initial_time"] = 0.
    if "time_scheme" not in user:
        m, dm = p["model"], p["diffusion_model"]
        if m == 1:
            p["time_scheme"] = 4 if dm == 0 else 1
        elif m in (2, 3):
            p["time_scheme"] = 5
    return p


class RefSolver:
    """Serial solver over the C restatement; rng is a RefRng (default: module RNG)."""

    def __init__(self, fields, particles, parameters=None, trifinder=None, rng=None,
                 control_sections=None):
        self.f, self.p, self.mesh = fields, particles, fields.mesh
        self.params = resolve_params(parameters or {})
        pr = self.params
        self.rng = rng if rng is not None else RNG
        self.trifinder = trifinder
        self.nt = int((pr["final_time"] - pr["initial_time"]) / pr["time_step"])
        self.par = ParS(pr["model"], pr["time_scheme"], pr["particle_velocity_init"],
                        pr["rng_method"], pr["lsm3_edge_opt"], pr["water_density"],
                        pr["diffusion_model"], pr["horizontal_diffusivity"],
                        pr["vertical_diffusivity"], pr["schmidt_number"],
                        pr["diffusion_lsm2_option"], pr["diffusion_lsm3_tl_horizontal"],
                        pr["diffusion_lsm3_tl_vertical"], int(pr["boundary_conditions"]),
                        pr["boundary_conditions_type"], pr["max_wall_rebounds"],
                        pr["rebound_damping_coef"], pr["wall_roughness_angle"],
                        pr["stranding_probability"], int(pr["bnd_statistics"]),
                        int(pr["stranding_output"]), int(pr["bottom_deposition"]))
        self.bnd_stats = np.zeros(max(1, self.mesh.nopen), np.int32)
        self.bnd_history = []
        self.init_tri = np.full(particles.npart, -1, np.int32)
        fs = fields.s
        frozen = pr["frozen_eulerian_fields"]
        fs.frozen = 1 if frozen else 0
        if not frozen:
            fs.record = lib.ref_lower_index(P(fields.times, "d"), fields.ntimes,
                                            pr["initial_time"], 0)
        fs.interp_done = 0
        fs.prev_valid = 0
        particles.s.rho0 = pr["water_density"]
        if pr["model"] in (2, 3):
            lib.ref_initialize_model_coefficients(C.byref(particles.s))
        self.s = SolverS(C.pointer(self.mesh.s), C.pointer(fs), C.pointer(particles.s),
                         C.pointer(self.par), Pcg(0, 0), pr["initial_time"], pr["time_step"], 0,
                         P(self.bnd_stats, "i"), P(self.init_tri, "i"), int(trifinder is not None),
                         0, 0)
        # control sections: polygons (npoint, 2); ControlSection counts
        # poly.shape[1] - 1 segments [poly[k], poly[k + 1]] (control_shapes.pyx:17-19)
        self.csec_history = []
        if control_sections is not None:
            polys = [np.asarray(q, dtype=np.float64) for q in control_sections]
            self._csec_nseg = np.array([q.shape[1] - 1 for q in polys], np.int32)
            segs = [[q[k, 0], q[k, 1], q[k + 1, 0], q[k + 1, 1]]
                    for q in polys for k in range(q.shape[1] - 1)]
            self._csec_seg = np.array(segs, np.float64).reshape(-1)
            self.csec_stats = np.zeros(max(1, len(polys)), np.int32)
            self.s.ncsec = len(polys)
            self.s.csec_nseg = P(self._csec_nseg, "i")
            self.s.csec_seg = P(self._csec_seg, "d")
            self.s.csec_stats = P(self.csec_stats, "i")
        if not frozen:
            e = lib.ref_time_interpolation(C.byref(fs), pr["initial_time"])
            if e:
                raise ValueError("time interpolation failed")
        self._sync_in()
        act = np.nonzero(particles.inactive == 0)[0]
        self._trifind(act)
        for i in act:
            e = lib.ref_initialize_particle_state(C.byref(self.s), int(i))
            if e:
                raise ValueError("initialize_particle_state failed ({})".format(e))
        self._sync_out()

    def _trifind(self, idx):
        if self.trifinder is not None and len(idx):
            pos = self.p.position
            self.init_tri[idx] = np.asarray(self.trifinder(pos[idx, 0], pos[idx, 1]),
                                            dtype=np.int32)

    def _sync_in(self):
        self.s.rng.state, self.s.rng.inc = self.rng.g.state, self.rng.g.inc

    def _sync_out(self):
        self.rng.g.state, self.rng.g.inc = self.s.rng.state, self.s.rng.inc

    @property
    def time(self):
        return self.s.time

    def forward(self, nsteps=1):
        self._sync_in()
        try:
            for _ in range(nsteps):
                if self.trifinder is not None:
                    new = np.nonzero((self.p.inactive == 0) & (self.p.tri == -2))[0]
                    self._trifind(new)
                e = lib.ref_forward(C.byref(self.s))
                if e:
                    raise ValueError("oracle forward failed: code {} particle {}".format(
                        self.s.err_code, self.s.err_particle))
                self.bnd_history.append(self.bnd_stats[:self.mesh.nopen].copy())
                if self.s.ncsec:
                    self.csec_history.append(self.csec_stats[:self.s.ncsec].copy())
        finally:
            self._sync_out()

    def solve(self):
        self.forward(self.nt)

    def add(self, positions):
        """LagrangianParticleSet.add first-fit (lagrangian_particle_set.pyx:620-710), pool
        large enough (no growth)."""
        p, j = self.p, 0
        positions = np.asarray(positions, dtype=np.float64)
        for i in range(p.npart):
            if j == len(positions):
                break
            if p.inactive[i] == 1:
                p.inactive[i] = 0
                p.position[i] = positions[j]
                p.diameter[i] = p.particle_diameter
                p.volume[i] = np.pi * (p.diameter[i] ** 3) / 6.
                p.tri[i] = -2
                p.s.npart_active += 1
                j += 1
        if j < len(positions):
            raise ValueError("pool too small for RefSolver.add")


def read_cli(file_name, vertex_id_column=11):
    """Telemac .cli boundary reader (cylag/geom/read_cli.pyx:24-98)."""
    pts, newb, count = [], False, 0
    with open(file_name) as fh:
        for line in fh:
            w = line.split()
            if not w:
                continue
            if int(w[0]) == 2 and int(w[1]) == 2 and int(w[2]) == 2:
                lab, newb = 0, False
            else:
                if not newb:
                    count += 1
                lab, newb = count, True
            pts.append([int(w[vertex_id_column]) - 1, lab])
    bp = np.asarray(pts).astype(np.int32)
    nb = bp.shape[0]
    if bp[0, 1] != 0 and bp[nb - 1, 1] != 0:
        last = bp[nb - 1, 1]
        bp[bp[:, 1] == last, 1] = bp[0, 1]
    return bp, count
