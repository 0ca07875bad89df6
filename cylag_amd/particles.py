"""Lagrangian particle pool (cylag/core/lagrangian_particle_set.pyx).

The host object keeps the reference's public arrays, shapes and pool
semantics: AoS (npart, dim) rows, an `inactive` flag per slot, first-fit
`add`, `delete_single` zeroing the row and renumbering its tag.  Once a
Solver runs, the rows live on the GPU; reading an array downloads the
current rows, and assigning an array (or calling `add`/`delete`) marks the
host copy as the newer one so it is uploaded before the next step.
"""
import numpy as np

from .constants import DEFAULT_INITIAL_POOL_SIZE, GRAV, NU0, EPSILON

_ARRAYS = ("position", "velocity", "fluid_velocity_seen", "last_position", "last_velocity",
           "inactive", "tri", "lowerlayer", "last_tri", "depth", "diameter", "volume", "tag")


def _prop(name):
    def get(self):
        self._sync()
        return self._h[name]

    def set(self, value):
        self._sync()
        old = self._h.get(name)
        dtype = old.dtype if isinstance(old, np.ndarray) else None
        self._h[name] = None if value is None else np.ascontiguousarray(value, dtype=dtype)
        self._host_dirty = True

    return property(get, set)


class LagrangianParticleSet:
    def __init__(self, position=None, velocity=None, fluid_velocity_seen=None, dim=2,
                 initial_pool_size=None, min_pool_size_update=1000, particle_density=1000.,
                 particle_diameter=0.01, particle_diameter_distribution=None,
                 drag_coefficient_model=1, drag_coefficient=0.44, added_mass_force=False,
                 added_mass_coef=0.5, buoyancy_velocity_model=0, buoyancy_velocity=0.0,
                 compute_depth=False, max_stranded=None, max_deposited=None):
        self._h = {}
        self._solver = None
        self._stale = False
        self._host_dirty = False
        if position is not None and initial_pool_size is None:
            self.npart = np.shape(position)[0]
        else:
            self.npart = initial_pool_size if initial_pool_size is not None else DEFAULT_INITIAL_POOL_SIZE
        n = self.npart
        self.dim = dim
        self.min_pool_size_update = min_pool_size_update
        self.npart_active = 0
        h = self._h
        h["inactive"] = np.ones(n, dtype=np.int32)
        self.rho0 = 1000.
        self.density = particle_density
        self.particle_diameter = particle_diameter
        self.particle_volume = np.pi * (particle_diameter ** 3) / 6.
        if particle_diameter_distribution is not None:
            self.diameter_distribution = particle_diameter_distribution[0]
            self.diameter_distribution_param = particle_diameter_distribution[1]
        else:
            self.diameter_distribution = 'monodisperse'
            self.diameter_distribution_param = 0.0
        h["diameter"] = np.full(n, self._draw_diameter()) if n else np.zeros(0)
        h["volume"] = np.pi * (h["diameter"] ** 3) / 6.
        h["position"] = np.zeros((n, dim))
        h["velocity"] = np.zeros((n, dim))
        h["fluid_velocity_seen"] = np.zeros((n, dim))
        self.compute_depth = compute_depth
        if compute_depth and dim != 3:
            raise ValueError("compute_depth can only be True for 3D pset")
        h["depth"] = np.zeros(n) if compute_depth else None
        h["tag"] = np.arange(n, dtype=np.int32)
        self.maxtag = int(h["tag"][n - 1])
        self.max_stranded = max_stranded if max_stranded is not None else n
        self.stranded_count = 0
        self.stranded_positions = np.zeros((self.max_stranded, 3))
        # deposits on the bed (bottom_deposition, an extension): position and iteration
        self.max_deposited = max_deposited if max_deposited is not None else n
        self.deposited_count = 0
        self.deposited_positions = np.zeros((self.max_deposited, 3))
        self.deposited_iterations = np.zeros(self.max_deposited, dtype=np.int64)
        self._init_state_vector(position, velocity, fluid_velocity_seen)
        h["tri"] = -np.ones(n, dtype=np.int32)
        h["lowerlayer"] = np.zeros(n, dtype=np.int32)
        h["last_tri"] = -np.ones(n, dtype=np.int32)
        h["last_position"] = h["position"].copy()
        h["last_velocity"] = h["velocity"].copy()
        self.drag_coef_model = drag_coefficient_model
        self.drag_coef = drag_coefficient
        self.added_mass_force = added_mass_force
        self.added_mass_coef = added_mass_coef
        self.buoyancy_velocity_model = buoyancy_velocity_model
        self.buoyancy_velocity_value = buoyancy_velocity
        self.additional_velocity = False
        self.additional_force = False
        self.a0c = self.a1c = self.a2c = self.a3c = 0.

    for _n in _ARRAYS:
        locals()[_n] = _prop(_n)
    del _n

    def _draw_diameter(self):
        if self.diameter_distribution == 'monodisperse':
            return self.particle_diameter
        raise ValueError("only monodisperse diameters are drawn here; set `diameter` directly")

    def _init_state_vector(self, position=None, velocity=None, fluid_velocity_seen=None):
        """lagrangian_particle_set.pyx:263-335"""
        h = self._h
        if position is not None:
            pshp = np.shape(position)
            if pshp[0] > self.npart:
                raise ValueError('size of position > initial_pool_size')
            if pshp[1] != self.dim:
                raise ValueError('dim of position != LagrangianParticleSet.dim')
            h["position"][:pshp[0]] = position
            h["inactive"][:pshp[0]] = 0
            self.npart_active = pshp[0]
        if velocity is not None:
            pshp = np.shape(velocity)
            if position is None:
                raise ValueError('position must be != None')
            if pshp[0] != self.npart_active:
                raise ValueError('size of velocity != size of position')
            if pshp[1] != self.dim:
                raise ValueError('dim of velocity != LagrangianParticleSet.dim')
            h["velocity"][:pshp[0]] = velocity
        if fluid_velocity_seen is not None:
            pshp = np.shape(fluid_velocity_seen)
            if position is None:
                raise ValueError('position must be != None')
            if velocity is None:
                raise ValueError('velocity must be != None')
            if pshp[0] != self.npart_active:
                raise ValueError('size of fluid_velocity_seen != size of velocity')
            if pshp[1] != self.dim:
                raise ValueError('dim of fluid_velocity_seen != LagrangianParticleSet.dim')
            h["fluid_velocity_seen"][:pshp[0]] = fluid_velocity_seen

    # -- device synchronisation ------------------------------------------------
    def _sync(self):
        if self._stale and self._solver is not None:
            self._stale = False
            self._solver._pull()

    def mark_modified(self):
        """Declare in-place edits of the host arrays (uploaded before the next step)."""
        self._sync()
        self._host_dirty = True

    # -- reference methods ---------------------------------------------------------
    def _initialize_model_coefficients(self):
        """A0..A3 constants (lagrangian_particle_set.pyx:349-375)."""
        rhop, Ca = self.density, self.added_mass_coef
        if not self.added_mass_force:
            self.a0c = (rhop - self.rho0) / rhop
            self.a1c = 0.5 * (self.rho0 / rhop)
            self.a2c = 0.
            self.a3c = 1. / rhop
        else:
            aux = 1. / (rhop + Ca * self.rho0)
            self.a0c = (rhop - self.rho0) * aux
            self.a1c = 0.5 * self.rho0 * aux
            self.a2c = self.rho0 * (1. + Ca) * aux
            self.a3c = aux

    def show_status(self, i):
        print("~~~~~~~~~~~~~~~~~~~~~")
        print("Particle n°", i)
        for k in ("position", "velocity", "fluid_velocity_seen"):
            print("{:20s}=".format(k), np.asarray(getattr(self, k)[i]))
        for k in ("tag", "inactive", "tri", "lowerlayer", "last_tri"):
            print("{:20s}=".format(k), getattr(self, k)[i])
        print("last_position       =", np.asarray(self.last_position[i]))

    def _sort_active_particles(self):
        """Re-sort the pool so active rows come first (lagrangian_particle_set.pyx:584-618).
        This changes the pool order, hence the random-stream order of later steps."""
        self._sync()
        h = self._h
        idx = np.argsort(h["inactive"])
        for k in _ARRAYS:
            if h.get(k) is not None:
                h[k] = np.ascontiguousarray(h[k][idx])
        self._host_dirty = True

    def get_state(self):
        na = self.npart_active
        self._sort_active_particles()
        h = self._h
        return h["position"][:na], h["velocity"][:na], h["fluid_velocity_seen"][:na]

    def get_particle_properties(self):
        na = self.npart_active
        self._sort_active_particles()
        return self._h["diameter"][:na], self._h["volume"][:na]

    def _increase_pool_size(self, size):
        """lagrangian_particle_set.pyx:522-582"""
        self._sync()
        h, d = self._h, self.dim
        add = max(self.min_pool_size_update, size)
        self.npart += add
        h["inactive"] = np.concatenate([h["inactive"], np.ones(add, np.int32)])
        h["diameter"] = np.concatenate([h["diameter"], np.zeros(add)])
        h["volume"] = np.concatenate([h["volume"], np.zeros(add)])
        for k in ("position", "velocity", "fluid_velocity_seen", "last_position", "last_velocity"):
            h[k] = np.concatenate([h[k], np.zeros((add, d))])
        if self.compute_depth:
            h["depth"] = np.concatenate([h["depth"], np.zeros(add)])
        h["tag"] = np.concatenate([h["tag"], np.arange(self.maxtag + 1, self.maxtag + 1 + add,
                                                       dtype=np.int32)])
        self.maxtag += add
        h["tri"] = np.concatenate([h["tri"], -np.ones(add, np.int32)])
        h["lowerlayer"] = np.concatenate([h["lowerlayer"], np.zeros(add, np.int32)])
        h["last_tri"] = np.concatenate([h["last_tri"], np.zeros(add, np.int32)])
        self._host_dirty = True

    def add(self, position):
        """First-fit activation of inactive slots (lagrangian_particle_set.pyx:620-710).
        Returns the slots used (rows get tri = -2, initialized by the next step)."""
        self._sync()
        position = np.asarray(position, dtype=np.float64)
        nadd = position.shape[0]
        if nadd == 0:
            return np.zeros(0, np.int32)
        if nadd > self.npart - self.npart_active:
            self._increase_pool_size(nadd - (self.npart - self.npart_active))
        h = self._h
        slots = np.nonzero(h["inactive"] == 1)[0][:nadd]
        h["inactive"][slots] = 0
        for s in slots:
            h["diameter"][s] = self._draw_diameter()
        h["volume"][slots] = np.pi * (h["diameter"][slots] ** 3) / 6.
        h["position"][slots] = position
        h["tri"][slots] = -2
        self.npart_active += nadd
        self._host_dirty = True
        return slots.astype(np.int32)

    def delete(self, particles_id):
        for i in particles_id:
            self.delete_single(int(i))

    def delete_single(self, i):
        """lagrangian_particle_set.pyx:754-788"""
        self._sync()
        h = self._h
        if h["inactive"][i] == 0:
            h["inactive"][i] = 1
            h["diameter"][i] = 0.
            h["volume"][i] = 0.
            for k in ("position", "velocity", "fluid_velocity_seen", "last_position", "last_velocity"):
                h[k][i, :] = 0.
            if self.compute_depth:
                h["depth"][i] = 0.
            h["tri"][i] = -1
            h["lowerlayer"][i] = 0
            h["last_tri"][i] = -1
            self.npart_active -= 1
            h["tag"][i] = self.maxtag + 1
            self.maxtag += 1
            self._host_dirty = True

    def drag_coefficient(self, U, dp):
        """lagrangian_particle_set.pyx:790-832"""
        Re = max(EPSILON, U * dp / NU0)
        if self.drag_coef_model == 0:
            return self.drag_coef
        if self.drag_coef_model == 1:
            return (24. / Re) * (1. + 0.15 * Re ** 0.687) if Re <= 1000. else 0.44
        if self.drag_coef_model == 2:
            phi1 = (24. / Re) ** 10. + (21. * Re ** -0.67) ** 10. + (4. * Re ** -0.33) ** 10. + 0.4 ** 10
            phi2 = 1. / ((0.148 * Re ** 0.11) ** -10. + 0.5 ** -10.)
            phi3 = (1.57 * (Re ** -1.625) * 1.e8) ** 10.
            phi4 = 1. / ((6e-17 * Re ** 2.63) ** -10. + 0.2 ** -10.)
            return (1. / ((phi1 + phi2) ** -1. + phi3 ** -1.) + phi4) ** (1 / 10.)
        return 0.44

    def buoyancy_velocity(self, dp, model_hr, rho0):
        """lagrangian_particle_set.pyx:834-890"""
        rhop, Cd, s = self.density, self.drag_coef, self.density / rho0
        if self.buoyancy_velocity_model == 0:
            return self.buoyancy_velocity_value
        if dp < 1.e-4:
            wb = abs(s - 1.) * (GRAV * dp ** 2) / (18. * NU0)
        elif dp > 1.e-3:
            wb = (np.sqrt(abs(s - 1.) * (4. * dp * GRAV) / (3. * Cd)) if model_hr == 0
                  else 1.1 * np.sqrt((s - 1.) * GRAV * dp))
        else:
            wb = (10. * NU0 / dp) * (np.sqrt(1. + abs(s - 1.) * GRAV * dp ** 3 / (100. * NU0 ** 2)) - 1.)
        if s > 1.:
            wb *= -1.
        return wb
