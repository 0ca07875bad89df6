"""Particle arrivals during a run: schedulers, initial positions in a polygon,
and the particle sources the Solver releases each step (the arrival path in
front of the device step: new rows are activated first-fit and initialized
on the device at the next step).

Restates cylag/extra/scheduler.pyx, cylag/extra/initial_conditions.pyx,
cylag/extra/particle_source.pyx and polygon_area (cylag/geom/utils.pyx:200-225).
"""
import numpy as np

from . import random_utils
from .constants import INTERN_REF  # label of an interior mesh vertex (-1)


# ---------------------------------------------------------------------------
# schedulers (scheduler.pyx:22-179)
# ---------------------------------------------------------------------------
class Scheduler:
    """Base scheduler: never triggers."""

    def now(self, simutime):
        return False


class TimeDeltaScheduler(Scheduler):
    """Every time_delta, from ini_time (scheduler.pyx:32-46)."""

    def __init__(self, time_delta, ini_time=None):
        self.time_delta = time_delta
        self.next_time = ini_time
        self.iteration = -1

    def now(self, simutime):
        if simutime.time <= self.next_time < simutime.time + simutime.time_step:
            self.next_time += self.time_delta
            self.iteration += 1
            return True
        return False


class IterationDeltaScheduler(Scheduler):
    """Every iteration_delta iterations (scheduler.pyx:48-59)."""

    def __init__(self, iteration_delta):
        self.iteration = -1
        self.iteration_delta = iteration_delta

    def now(self, simutime):
        if simutime.iteration % self.iteration_delta == 0:
            self.iteration += 1
            return True
        return False


class IterationScheduler(Scheduler):
    """At the listed iterations (scheduler.pyx:61-73)."""

    def __init__(self, iterations):
        self.iteration = -1
        self.iterations = iterations

    def now(self, simutime):
        if simutime.iteration in self.iterations:
            self.iteration += 1
            return True
        return False


class TimeScheduler(Scheduler):
    """At the listed times, in order (scheduler.pyx:75-90)."""

    def __init__(self, times):
        self.iteration = 0
        self.times = np.array(times, dtype='d')
        self.ntimes = len(times)

    def now(self, simutime):
        if self.iteration < self.ntimes and \
                simutime.time <= self.times[self.iteration] < simutime.time + simutime.time_step:
            self.iteration += 1
            return True
        return False


class CountScheduler(Scheduler):
    """count tasks equally spaced over [0, final_time] (scheduler.pyx:92-109)."""

    def __init__(self, count):
        self.iteration = 0
        self.count = count
        self.is_ready = True

    def now(self, simutime):
        times = np.linspace(0, simutime.final_time, self.count)
        if self.iteration < self.count and \
                simutime.time <= times[self.iteration] < simutime.time + simutime.time_step:
            self.iteration += 1
            return True
        return False


class AlwaysScheduler(Scheduler):
    """Every iteration (scheduler.pyx:111-118)."""

    def __init__(self):
        self.iteration = -1

    def now(self, simutime):
        self.iteration += 1
        return True


def schedules(ini_time=0., time_delta=None, iteration_delta=None, times=None, iterations=None, count=None,
              never=None, always=None):
    """One scheduler from exactly one of the trigger arguments (scheduler.pyx:120-179)."""
    args = [time_delta, iteration_delta, times, iterations, count, never, always]
    n = len(args) - args.count(None)
    if n == 0:
        raise ValueError("At least one of these argument required: "
                         "time_delta, iteration_delta, times, iterations, count, never, always")
    if n > 1:
        raise ValueError("Only one of these arguments possible: "
                         "time_delta, iteration_delta, times, iterations, count, never, always , got {}".format(n))
    if time_delta is not None:
        return TimeDeltaScheduler(time_delta, ini_time)
    if iteration_delta is not None:
        return IterationDeltaScheduler(iteration_delta)
    if times is not None:
        return TimeScheduler(times)
    if iterations is not None:
        return IterationScheduler(iterations)
    if count is not None:
        return CountScheduler(count)
    if never is not None:
        return Scheduler()
    return AlwaysScheduler()


# ---------------------------------------------------------------------------
# initial positions (initial_conditions.pyx)
# ---------------------------------------------------------------------------
def polygon_area(x, y):
    """Shoelace area of a simple polygon (geom/utils.pyx:200-225)."""
    x = np.asarray(x, dtype='d')
    y = np.asarray(y, dtype='d')
    n = x.shape[0]
    if n < 3:
        return 0.0
    area = 0.0
    for i in range(n - 1):
        area += x[i] * y[i + 1] - x[i + 1] * y[i]
    area += x[n - 1] * y[0] - x[0] * y[n - 1]
    return 0.5 * abs(area)


def circle(x0, y0, r, n):
    """n vertices of a circle (initial_conditions.pyx:218-231)."""
    poly = np.zeros((n, 2), dtype='d')
    dtheta = 2. * np.pi / n
    for i in range(n):
        theta = i * dtheta
        poly[i, 0] = x0 + r * np.cos(theta)
        poly[i, 1] = y0 + r * np.sin(theta)
    return poly


def _xy_grid(poly, poly_area, xy_method, grid_res, xy_npart, xy_density):
    xlim = (np.min(poly[:, 0]), np.max(poly[:, 0]))
    ylim = (np.min(poly[:, 1]), np.max(poly[:, 1]))
    if xy_method == 0:
        ni, nj = grid_res[0] + 1, grid_res[1] + 1
    elif xy_method == 1:
        dxy = 1. / np.sqrt(xy_npart / poly_area)
        ni = int((xlim[1] - xlim[0]) / dxy) + 1
        nj = int((ylim[1] - ylim[0]) / dxy) + 1
    elif xy_method == 2:
        dxy = 1. / np.sqrt(xy_density)
        ni = int((xlim[1] - xlim[0]) / dxy) + 1
        nj = int((ylim[1] - ylim[0]) / dxy) + 1
    else:
        raise ValueError("Invalid initialization method")
    return np.linspace(xlim[0], xlim[1], ni), np.linspace(ylim[0], ylim[1], nj)


def _inside(poly, positions):
    """Cell centres kept: those a matplotlib Polygon patch of the vertices
    contains (as the reference filters them)."""
    from matplotlib import patches
    polygon = patches.Polygon(poly)
    keep = np.array([polygon.contains_point((p[0], p[1])) for p in positions], dtype=bool)
    return positions[keep] if keep.size else positions


def init_positions_2d(poly, poly_area=None, xy_method=0, grid_res=np.array([5, 5]), xy_npart=25, xy_density=25.):
    """Centres of a grid's cells inside a polygon (initial_conditions.pyx:25-106)."""
    poly = np.asarray(poly, dtype='d')
    if poly.shape[0] < 3:
        raise ValueError("Invalid polygon")
    if poly_area is None:
        poly_area = polygon_area(poly[:, 0], poly[:, 1])
    xl, yl = _xy_grid(poly, poly_area, xy_method, grid_res, xy_npart, xy_density)
    ni, nj = xl.size, yl.size
    positions = np.zeros(((ni - 1) * (nj - 1), 2), dtype='d')
    for i in range(ni - 1):
        for j in range(nj - 1):
            k = i + j * (ni - 1)
            positions[k, 0] = 0.5 * (xl[i] + xl[i + 1])
            positions[k, 1] = 0.5 * (yl[j] + yl[j + 1])
    return _inside(poly, positions)


def init_positions_3d(poly, poly_area=None, xy_method=0, grid_res=np.array([5, 5, 1]), xy_npart=25,
                      xy_density=25., z_method=0, z_npart=1, z_density=1., zmin=0., zmax=1.):
    """Centres of a 3D grid's cells above a polygon (initial_conditions.pyx:108-216)."""
    poly = np.asarray(poly, dtype='d')
    if poly.shape[0] < 3:
        raise ValueError("Invalid polygon")
    if poly_area is None:
        poly_area = polygon_area(poly[:, 0], poly[:, 1])
    xl, yl = _xy_grid(poly, poly_area, xy_method, grid_res, xy_npart, xy_density)
    if z_method == 0:
        nl = grid_res[2] + 1
    elif z_method == 1:
        nl = z_npart + 1
    elif z_method == 2:
        nl = int(z_density * (zmax - zmin)) + 1
    else:
        raise ValueError("Invalid initialization method")
    zl = np.linspace(zmin, zmax, nl)
    ni, nj = xl.size, yl.size
    positions = np.zeros(((ni - 1) * (nj - 1) * (nl - 1), 3), dtype='d')
    for i in range(ni - 1):
        for j in range(nj - 1):
            for l in range(nl - 1):
                k = i + j * (ni - 1) + l * (ni - 1) * (nj - 1)
                positions[k, 0] = 0.5 * (xl[i] + xl[i + 1])
                positions[k, 1] = 0.5 * (yl[j] + yl[j + 1])
                positions[k, 2] = 0.5 * (zl[l] + zl[l + 1])
    return _inside(poly, positions)


# ---------------------------------------------------------------------------
# sources (particle_source.pyx)
# ---------------------------------------------------------------------------
class ParticleSource:
    """Releases the grid positions of a polygon when its scheduler triggers
    (particle_source.pyx:31-97)."""

    def __init__(self, scheduler, poly, grid_res=np.array([5, 5, 1], dtype='int32'), xy_method=0, xy_npart=25,
                 xy_density=25., z_method=0, z_npart=1, z_density=1., zmin=0., zmax=1.):
        self.scheduler = scheduler
        self.poly = np.asarray(poly, dtype='d')
        self.poly_area = polygon_area(self.poly[:, 0], self.poly[:, 1])
        self.grid_res = grid_res
        self.xy_method = xy_method
        self.xy_npart = xy_npart
        self.xy_density = xy_density
        self.z_method = z_method
        self.z_npart = z_npart
        self.z_density = z_density
        self.zmin = zmin
        self.zmax = zmax

    def release_condition(self, simutime, fset):
        return self.scheduler.now(simutime)

    def init_positions(self, fset, dim=-1):
        if dim == -1:
            dim = fset.dim
        if dim == 2:
            return init_positions_2d(self.poly, self.poly_area, self.xy_method, self.grid_res, self.xy_npart,
                                     self.xy_density)
        return init_positions_3d(self.poly, self.poly_area, self.xy_method, self.grid_res, self.xy_npart,
                                 self.xy_density, self.z_method, self.z_npart, self.z_density, self.zmin, self.zmax)


_OPS = {"eq": np.equal, "neq": np.not_equal, "geq": np.greater_equal, "leq": np.less_equal,
        "gt": np.greater, "lt": np.less}


class ParticleSourceFromField:
    """Releases particles around the interior nodes where a field (or its
    variation between the current and next record) meets a criterion, as many
    as density x dual-cell area / virtual_mass, uniformly in a disc of the
    dual cell's area, with the module's uniform draws (particle_source.pyx:99-333)."""

    def __init__(self, scheduler, field_id_criteria, field_id_density, criteria, threshold, virtual_mass=1.,
                 single_release=False):
        self.scheduler = scheduler
        self.field_id_criteria = field_id_criteria
        self.field_id_density = field_id_density
        self.criteria = criteria
        self.threshold = threshold
        self.virtual_mass = virtual_mass
        self.loc = np.empty(0, dtype='int64')
        self.single_release = single_release
        self.release_rec = -1

    def release_condition(self, simutime, fset):
        rec = fset.record
        if not self.scheduler.now(simutime):
            return 0
        kind, _, op = self.criteria.partition('_')
        if kind not in ("var", "val") or op not in _OPS:
            raise ValueError("Unkwown criteria")
        field = fset.field_by_id(self.field_id_criteria)
        f0 = np.asarray(field[rec])
        value = (np.asarray(field[rec + 1]) - f0) if kind == "var" else f0
        self.loc = np.where(_OPS[op](value, self.threshold))[0]
        if self.loc.shape[0] > 0:
            if self.single_release:
                if rec != self.release_rec:
                    self.release_rec = rec
                    return 1
            else:
                self.release_rec = rec
                return 1
        return 0

    def init_positions(self, fset, dim=-1):
        if fset.dim != 2:
            raise ValueError("Dimension must be 2")
        if dim == 3:
            raise ValueError("ParticleSourceFromField only generates 2D positions, "
                             "it cannot be used with a 3D LagrangianParticleSet")
        m = fset.triangular_mesh
        dens_field = np.asarray(fset.field_by_id(self.field_id_density)[fset.record])
        out = [np.zeros((0, 2), dtype='d')]
        for k in self.loc:
            if m.vertex_labels[k] != INTERN_REF:
                continue
            tris = m.vertex_adjacency[k]
            dual_cell_area = 0.
            for itri in tris:
                if itri != -1:
                    dual_cell_area += abs(m.signed_area[itri]) / 3.
            density = dens_field[k] / self.virtual_mass
            r0 = np.sqrt(dual_cell_area / np.pi)
            if density != 0.:
                npart_new = max(1, int(density * dual_cell_area))
                pos = np.zeros((npart_new, 2), dtype='d')
                x0, y0 = m.x[k], m.y[k]
                for j in range(npart_new):
                    r = r0 * np.sqrt(random_utils.random_uniform())
                    a = 2. * np.pi * random_utils.random_uniform()
                    pos[j, 0] = x0 + r * np.cos(a)
                    pos[j, 1] = y0 + r * np.sin(a)
                out.append(pos)
        return np.concatenate(out, axis=0)
