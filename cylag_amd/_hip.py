"""ctypes binding of libcylag_hip.so (declarations in include/cylag_hip.h).

The library is the only compute path of this package: importing the step
without it, or running it without a GPU, raises instead of falling back.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcylag_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_u64p = C.POINTER(C.c_uint64)


class CylParams(C.Structure):
    _fields_ = [("model", C.c_int), ("time_scheme", C.c_int), ("particle_velocity_init", C.c_int),
                ("rng_method", C.c_int), ("diffusion_model", C.c_int),
                ("diffusion_lsm2_option", C.c_int), ("horizontal_diffusivity", C.c_double),
                ("vertical_diffusivity", C.c_double), ("schmidt_number", C.c_double),
                ("boundary_conditions", C.c_int), ("boundary_conditions_type", C.c_int),
                ("max_wall_rebounds", C.c_int), ("bnd_statistics", C.c_int),
                ("rebound_damping_coef", C.c_double), ("wall_roughness_angle", C.c_double),
                ("stranding_probability", C.c_double), ("compute_depth", C.c_int),
                ("frozen", C.c_int), ("lsm3_edge_opt", C.c_int),
                ("diffusion_lsm3_tl_horizontal", C.c_double), ("diffusion_lsm3_tl_vertical", C.c_double),
                ("stranding_output", C.c_int), ("bottom_deposition", C.c_int)]


class CylPconst(C.Structure):
    _fields_ = [("rho0", C.c_double), ("density", C.c_double), ("drag_coef", C.c_double),
                ("added_mass_coef", C.c_double), ("buoyancy_velocity_value", C.c_double),
                ("drag_coef_model", C.c_int), ("added_mass_force", C.c_int),
                ("buoyancy_velocity_model", C.c_int), ("a0c", C.c_double), ("a1c", C.c_double),
                ("a2c", C.c_double), ("a3c", C.c_double)]


# name -> (restype, argtypes); every symbol include/cylag_hip.h declares
SIGNATURES = {
    "cyl_ctx_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "cyl_ctx_destroy": (C.c_int, [C.c_void_p]),
    "cyl_last_error": (C.c_int, [C.c_void_p, _ip, _ip, C.c_char_p, C.c_size_t]),
    "cyl_device_info": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t),
                                  C.POINTER(C.c_size_t)]),
    "cyl_mesh_upload": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _ip, _ip,
                                  _ip, _ip, _dp, _dp, _dp, C.c_int]),
    "cyl_fields_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp,
                                    _dp, _dp, C.c_int, _dp, _ip, C.c_int, C.c_double]),
    "cyl_fields_reset": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "cyl_fields_interpolate": (C.c_int, [C.c_void_p, C.c_double]),
    "cyl_fields_download_tmp": (C.c_int, [C.c_void_p, C.c_int, _dp]),
    "cyl_particles_upload": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip,
                                       _dp, _dp, _ip, C.POINTER(CylPconst)]),
    "cyl_particles_download": (C.c_int, [C.c_void_p, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _dp,
                                         _dp, _ip, _dp]),
    "cyl_particles_deleted": (C.c_int, [C.c_void_p, _ip, C.c_int, _ip]),
    "cyl_deleted_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "cyl_particles_add": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp, _dp, _dp, _ip]),
    "cyl_rng_set": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64]),
    "cyl_rng_get": (C.c_int, [C.c_void_p, _u64p, _u64p]),
    "cyl_pcg32_next": (C.c_uint32, [_u64p, C.c_uint64]),
    "cyl_pcg32_seed": (None, [_u64p, _u64p, C.c_uint64, C.c_uint64]),
    "cyl_pcg32_advance": (None, [_u64p, C.c_uint64, C.c_uint64]),
    "cyl_random_normal": (C.c_double, [_u64p, C.c_uint64, C.c_int]),
    "cyl_random_pair": (None, [_u64p, C.c_uint64, C.c_int, _dp, _dp]),
    "cyl_init_particles": (C.c_int, [C.c_void_p, C.POINTER(CylParams), C.c_double, C.c_double]),
    "cyl_step": (C.c_int, [C.c_void_p, C.POINTER(CylParams), _dp, C.c_double, C.c_int, _ip]),
    "cyl_set_tuning": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cyl_deposited_get": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
    "cyl_last_timing": (C.c_int, [C.c_void_p, _dp, _dp, _ip]),
    "cyl_last_prep_timing": (C.c_int, [C.c_void_p, _dp]),
    "cyl_last_path": (C.c_int, [C.c_void_p, _ip, _ip, _ip]),
    "cyl_control_sections_set": (C.c_int, [C.c_void_p, C.c_int, _ip, _dp]),
    "cyl_control_counts": (C.c_int, [C.c_void_p, _ip, C.c_int, _ip]),
    "cyl_stranded_get": (C.c_int, [C.c_void_p, _dp, C.c_int, _ip]),
    "cyl_scenarios_set": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_void_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), _ip]),
    "cyl_scenario_rng_get": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "cyl_batch_step": (C.c_int, [C.c_void_p, _dp, C.c_double, C.c_int, _ip]),
    "cyl_batch_collect": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _ip, C.c_int, _ip]),
    "cyl_batch_collect_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, _ip, _ip, _ip, _ip, C.c_int,
                                            _ip]),
    "cyl_batch_reshape": (C.c_int, [C.c_void_p, C.c_int, _ip, C.c_int, _ip, C.c_void_p, C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), _ip, _dp, _dp, _dp, _ip, _ip, _ip, _dp, _dp, _ip,
                                    C.c_double, C.c_double]),
}

_lib = None


def load():
    """Load libcylag_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libcylag_hip.so is not built; run __graft_entry__.build() "
                              "(make -C cylag_amd/csrc)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def ptr(a, kind="d"):
    if a is None:
        return C.cast(None, {"d": _dp, "i": _ip}[kind])
    return a.ctypes.data_as({"d": _dp, "i": _ip}[kind])


class HipError(RuntimeError):
    pass


class Context:
    """One device context: one HIP stream, mesh, fields and particle pool in HBM."""

    def __init__(self, device=0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.cyl_ctx_create(device, 0, C.byref(h))
        self.h = h
        if rc != 0:
            self.check(rc, "cyl_ctx_create")

    def check(self, rc, what):
        if rc == 0:
            return
        code, part = C.c_int(), C.c_int()
        buf = C.create_string_buffer(512)
        if self.h:
            self.lib.cyl_last_error(self.h, C.byref(code), C.byref(part), buf, 512)
        msg = buf.value.decode(errors="replace")
        if rc == -4:  # device-detected condition: the reference raises ValueError
            raise ValueError("{} (particle {})".format(msg, part.value))
        raise HipError("{} failed ({}): {}".format(what, rc, msg))

    def close(self):
        if self.h:
            self.lib.cyl_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        name = C.create_string_buffer(256)
        fb, tb = C.c_size_t(), C.c_size_t()
        self.check(self.lib.cyl_device_info(self.h, name, 256, C.byref(fb), C.byref(tb)),
                   "cyl_device_info")
        return name.value.decode(), fb.value, tb.value

    def last_timing(self):
        k, t, n = C.c_double(), C.c_double(), C.c_int()
        self.lib.cyl_last_timing(self.h, C.byref(k), C.byref(t), C.byref(n))
        return k.value, t.value, n.value

    def last_path(self):
        """(fused, generic) main-kernel launches of the last step call (eager steps)."""
        f, g, q = C.c_int(), C.c_int(), C.c_int()
        self.check(self.lib.cyl_last_path(self.h, C.byref(f), C.byref(g), C.byref(q)), "cyl_last_path")
        return f.value, g.value

    def last_graph_steps(self):
        """Steps of the last step call run by graph replays."""
        q = C.c_int()
        self.check(self.lib.cyl_last_path(self.h, None, None, C.byref(q)), "cyl_last_path")
        return q.value

    def last_prep_ms(self):
        q = C.c_double()
        self.lib.cyl_last_prep_timing(self.h, C.byref(q))
        return q.value


def contiguous(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)
