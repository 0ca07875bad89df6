"""Synthetic workloads shared by the parity tests, smoke() and bench.py.

Each case builds the same mesh, fields and initial particles for the product
(cylag_amd, HIP step) and for the oracle (oracle/refpy.py, CPU restatement),
starts both from the same global PCG32 state, and returns pool-order rows.
"""
import os
from types import SimpleNamespace

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SEED = (42, 0xda3e06d11bb2b203)


def structured_square(n, length=100.):
    """n x n nodes on [0, length]^2, each cell split by one diagonal."""
    xs = np.linspace(0., length, n)
    X, Y = np.meshgrid(xs, xs)
    x, y = X.ravel().copy(), Y.ravel().copy()
    idx = np.arange(n * n).reshape(n, n)
    a, b = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel()
    c, d = idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)]).astype(np.int32)
    return x, y, tri


def jittered_delaunay(nx, ny, lx, ly, seed=7, jitter=0.35):
    """Delaunay triangulation of a jittered grid (matplotlib/qhull)."""
    from matplotlib.tri import Triangulation
    rng = np.random.default_rng(seed)
    xs, ys = np.linspace(0., lx, nx), np.linspace(0., ly, ny)
    X, Y = np.meshgrid(xs, ys)
    dx, dy = lx / (nx - 1), ly / (ny - 1)
    inner = np.ones_like(X, bool)
    inner[[0, -1], :] = False
    inner[:, [0, -1]] = False
    X = X + inner * rng.uniform(-jitter, jitter, X.shape) * dx
    Y = Y + inner * rng.uniform(-jitter, jitter, Y.shape) * dy
    tri = Triangulation(X.ravel(), Y.ravel())
    return tri.x.copy(), tri.y.copy(), tri.triangles.astype(np.int32), tri


class Case:
    """mesh + fields + initial rows; `fields` is a dict of EulerianFieldSet kwargs."""

    def __init__(self, x, y, tri, fields, position, velocity=None, pool=None, boundary_points=None,
                 nopen=0, use_trifinder=True, pset_kw=None, mpl_tri=None, pdim=None, diameter=None):
        self.x, self.y, self.tri = x, y, tri
        self.fields = fields
        self.pdim = pdim if pdim is not None else fields.get("dim", 2)
        self.position = np.asarray(position, dtype=np.float64)
        self.velocity = velocity
        self.pool = pool
        self.boundary_points = boundary_points
        self.nopen = nopen
        self.use_trifinder = use_trifinder
        self.pset_kw = pset_kw or {}
        self.diameter = diameter        # per-row diameters (None: the particle set's own)
        self._mpl = mpl_tri

    def trifinder(self):
        if not self.use_trifinder:
            return None
        if self._mpl is None:
            from matplotlib.tri import Triangulation
            self._mpl = Triangulation(self.x, self.y, self.tri)
        return self._mpl.get_trifinder()

    # -- product -------------------------------------------------------------------
    def build_gpu(self, cy, params, **solver_kw):
        mesh = cy.TriangularMesh(self.x, self.y, self.tri, boundary_points=self.boundary_points,
                                 nopen=self.nopen)
        fs = cy.EulerianFieldSet(mesh, trifinder=self.trifinder(), **self.fields)
        ps = cy.LagrangianParticleSet(self.position, self.velocity, dim=self.pdim,
                                      initial_pool_size=self.pool, **self.pset_kw)
        if self.diameter is not None:
            d = np.zeros(ps.npart)
            d[:len(self.diameter)] = self.diameter
            ps.diameter = d
            ps.volume = np.pi * d ** 3 / 6.
        cy.pcg32_seed(*SEED)
        return cy.Solver(fs, ps, params, **solver_kw)

    def run_gpu(self, cy, params, nsteps=None, **solver_kw):
        s = self.build_gpu(cy, params, **solver_kw)
        if nsteps is None:
            s.solve()
        else:
            for _ in range(nsteps):
                s.forward(s.simutime.time_step)
        ps = s.particle_set
        return SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(),
                               fluid_velocity_seen=ps.fluid_velocity_seen.copy(),
                               tri=ps.tri.copy(), inactive=ps.inactive.copy(),
                               lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy(), solver=s)

    # -- oracle -----------------------------------------------------------------------
    def run_oracle(self, R, params, nsteps=None, **solver_kw):
        mesh = R.RefMesh(self.x, self.y, self.tri, boundary_points=self.boundary_points,
                         nopen=self.nopen)
        f = dict(self.fields)
        fs = R.RefFields(mesh, **{k: v for k, v in f.items() if k != "trifinder"})
        kw = {k: v for k, v in self.pset_kw.items()}
        if self.diameter is not None:
            n = self.pool if self.pool is not None else len(self.position)
            d = np.zeros(n)
            d[:len(self.diameter)] = self.diameter
            kw["diameter"] = d
        ps = R.RefParticles(self.position, self.velocity, dim=self.pdim,
                            initial_pool_size=self.pool, **kw)
        rng = R.RefRng(*SEED)
        s = R.RefSolver(fs, ps, params, trifinder=self.trifinder(), rng=rng, **solver_kw)
        if nsteps is None:
            s.solve()
        else:
            s.forward(nsteps)
        return SimpleNamespace(position=ps.position.copy(), velocity=ps.velocity.copy(),
                               fluid_velocity_seen=ps.us.copy(),
                               tri=ps.tri.copy(), inactive=ps.inactive.copy(),
                               lowerlayer=ps.lowerlayer.copy(), tag=ps.tag.copy(), rng=rng,
                               last_position=ps.last_position.copy(),
                               last_velocity=ps.last_velocity.copy(), solver=s)


def square_case(n=101, npart=10000, u=0.1, v=0.0, depth=1.0, centre=50., length=100.):
    """The plain 2D isotropic-diffusion workload: uniform flow on a structured
    square, all particles released at the centre (a mesh vertex when n is odd)."""
    x, y, tri = structured_square(n, length)
    nn = x.size
    fields = dict(dim=2, nlayers=1, times=np.zeros(1), elevation=np.full((1, nn), depth),
                  velocity_x=np.full((1, nn), u), velocity_y=np.full((1, nn), v))
    pos = np.full((npart, 2), centre)
    return Case(x, y, tri, fields, pos)


def reference_square_case(position, velocity=None, pool=None):
    """The reference tests' square mesh with a solid-body rotation field
    (tests/test_20_lsm1_analytical_advection_2d.py:15-66)."""
    x = np.loadtxt(os.path.join(GOLDEN, "square_meshx.txt"))
    y = np.loadtxt(os.path.join(GOLDEN, "square_meshy.txt"))
    tri = np.loadtxt(os.path.join(GOLDEN, "square_meshtri.txt"), dtype=np.int32)
    from oracle.refpy import read_cli
    bp, _ = read_cli(os.path.join(GOLDEN, "square_geo.cli"))
    om = 2. * np.pi / 3600.
    r, phi = np.sqrt(x ** 2 + y ** 2), np.arctan2(x, y)
    vx, vy = om * r * np.cos(phi), -om * r * np.sin(phi)
    n = x.size
    fields = dict(dim=2, nlayers=1, times=np.zeros(1), elevation=np.ones((1, n)),
                  velocity_x=vx[None], velocity_y=vy[None])
    return Case(x, y, tri, fields, position, velocity, pool, boundary_points=bp)


def _square_mesh():
    x = np.loadtxt(os.path.join(GOLDEN, "square_meshx.txt"))
    y = np.loadtxt(os.path.join(GOLDEN, "square_meshy.txt"))
    tri = np.loadtxt(os.path.join(GOLDEN, "square_meshtri.txt"), dtype=np.int32)
    from oracle.refpy import read_cli
    bp, _ = read_cli(os.path.join(GOLDEN, "square_geo.cli"))
    return x, y, tri, bp


P4_3D = np.array([[5000., 0., 0.5], [5000., 500., 0.9999], [1000., 0., 0.5], [1000., 500., 0.9999]])


def reference_square_3d_case(position=P4_3D, pool=4, nlayers=11):
    """The reference's 3D test field set: 11 uniform sigma planes on [0, 1], the
    solid-body rotation scaled by z, w = 0
    (tests/test_21_lsm1_analytical_advection_3d.py:15-84)."""
    x, y, tri, bp = _square_mesh()
    om = 2. * np.pi / 3600.
    r, phi = np.sqrt(x ** 2 + y ** 2), np.arctan2(x, y)
    n = x.size
    dz = 1. / max(nlayers - 1, 1)
    z = np.repeat(np.arange(nlayers) * dz, n)
    rr, pp = np.tile(r, nlayers), np.tile(phi, nlayers)
    fields = dict(dim=3, nlayers=nlayers, times=np.zeros(1), elevation=z[None],
                  velocity_x=(om * rr * np.cos(pp) * z)[None], velocity_y=(-om * rr * np.sin(pp) * z)[None],
                  velocity_z=np.zeros((1, n * nlayers)))
    return Case(x, y, tri, fields, position, None, pool, boundary_points=bp)


def reference_square_p3d_case(position=P4_3D, pool=4):
    """Pseudo-3D: the 2D rotation field with a zero BOTTOM field, unit depth and
    log-law friction (friction_model 1, coefficient 40), 3D particles
    (tests/test_22_lsm1_analytical_advection_p3d.py:15-72)."""
    x, y, tri, bp = _square_mesh()
    om = 2. * np.pi / 3600.
    r, phi = np.sqrt(x ** 2 + y ** 2), np.arctan2(x, y)
    n = x.size
    fields = dict(dim=2, nlayers=1, times=np.zeros(1), elevation=np.ones((1, n)),
                  velocity_x=(om * r * np.cos(phi))[None], velocity_y=(-om * r * np.sin(phi))[None],
                  opt_fields=np.zeros((1, 1, n)), opt_fields_names=['BOTTOM'],
                  friction_model=1, friction_coefficient=40.)
    return Case(x, y, tri, fields, position, None, pool, boundary_points=bp, pdim=3)


def vortex_case(nx=41, ny=11, lx=2000., ly=500., npart=2000, seed=3, ntimes=2, keps=True,
                wavelength=400., dead_zone=False):
    """Small version of the unstructured 2D workload: jittered Delaunay mesh,
    two frames of a cellular vortex, k-epsilon diffusivity varying in space.
    dead_zone: no turbulence (K = 0) for x < lx/4, where particles draw no
    normal, so the number of draws per particle is not uniform."""
    x, y, tri, mpl = jittered_delaunay(nx, ny, lx, ly, seed=7)
    k = 2. * np.pi / wavelength
    n = x.size
    frames_u, frames_v = [], []
    for t in range(ntimes):
        ph = 0.3 * t
        frames_u.append(np.sin(k * x + ph) * np.cos(k * y))
        frames_v.append(-np.cos(k * x + ph) * np.sin(k * y))
    times = np.linspace(0., 10000., ntimes)
    depth = 2. + 18. * (0.5 + 0.5 * np.sin(x / lx * 3.1) * np.cos(y / ly * 2.3))
    K = 0.1 + 4.9 * (0.5 + 0.5 * np.cos(x / lx * 5.) * np.sin(y / ly * 4.))
    eps = np.full(n, 1e-3)
    tke = np.sqrt(K * eps / (0.09 / 0.72))
    if dead_zone:
        tke[x < 0.25 * lx] = 0.
    fields = dict(dim=2, nlayers=1, times=times, elevation=np.tile(depth, (ntimes, 1)),
                  velocity_x=np.array(frames_u), velocity_y=np.array(frames_v))
    if keps:
        opt = np.stack([np.tile(tke, (ntimes, 1)), np.tile(eps, (ntimes, 1))])
        fields.update(opt_fields=opt, opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(0.05 * lx, 0.95 * lx, npart),
                    rng.uniform(0.05 * ly, 0.95 * ly, npart)], 1)
    return Case(x, y, tri, fields, pos, mpl_tri=mpl)


def vortex_3d_case(nx=21, ny=7, lx=2000., ly=500., npart=1500, seed=4, nlayers=5, ntimes=2,
                   wavelength=400.):
    """3D sigma-layer version of the vortex workload: nlayers planes between a
    smooth bottom -h(x, y) and the surface 0, horizontal velocity growing with
    height, w = 0, k-epsilon fields giving a parabolic K(z) (nodal k, eps per
    plane); particles spread through the water column."""
    x, y, tri, mpl = jittered_delaunay(nx, ny, lx, ly, seed=7)
    n = x.size
    k = 2. * np.pi / wavelength
    h = 5. + 10. * (0.5 + 0.5 * np.sin(x / lx * 3.1) * np.cos(y / ly * 2.3))
    sig = np.linspace(0., 1., nlayers)           # 0 bottom .. 1 surface
    z = np.concatenate([-h + s_ * h for s_ in sig])
    prof = np.concatenate([np.full(n, 0.3 + 0.7 * s_) for s_ in sig])
    xx, yy = np.tile(x, nlayers), np.tile(y, nlayers)
    us, vs = [], []
    for t in range(ntimes):
        ph = 0.3 * t
        us.append(np.sin(k * xx + ph) * np.cos(k * yy) * prof)
        vs.append(-np.cos(k * xx + ph) * np.sin(k * yy) * prof)
    K = np.concatenate([1e-2 + 0.2 * s_ * (1.05 - s_) * np.ones(n) for s_ in sig])
    eps = np.full(n * nlayers, 1e-3)
    tke = np.sqrt(K * eps / (0.09 / 0.72))
    fields = dict(dim=3, nlayers=nlayers, times=np.linspace(0., 10000., ntimes),
                  elevation=np.tile(z, (ntimes, 1)), velocity_x=np.array(us), velocity_y=np.array(vs),
                  velocity_z=np.zeros((ntimes, n * nlayers)),
                  opt_fields=np.stack([np.tile(tke, (ntimes, 1)), np.tile(eps, (ntimes, 1))]),
                  opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    rng = np.random.default_rng(seed)
    px = rng.uniform(0.05 * lx, 0.95 * lx, npart)
    py = rng.uniform(0.05 * ly, 0.95 * ly, npart)
    # depth at the particle from the nodal depth (nearest node: only to place it in the column)
    from scipy.spatial import cKDTree
    hn = h[cKDTree(np.stack([x, y], 1)).query(np.stack([px, py], 1))[1]]
    pz = -hn * rng.uniform(0.15, 0.85, npart)
    return Case(x, y, tri, fields, np.stack([px, py, pz], 1), mpl_tri=mpl)


def open_channel_case(n=21, npart=400, u=0.8, seed=5):
    """Structured square with open boundaries on the right (label 1) and left
    (label 2) sides, walls elsewhere: particles advected to the right leave
    the domain (deletions, tag renumbering, boundary statistics)."""
    x, y, tri = structured_square(n)
    nn = x.size
    right = np.nonzero(np.isclose(x, 100.) & (y > 0.) & (y < 100.))[0]
    left = np.nonzero(np.isclose(x, 0.) & (y > 0.) & (y < 100.))[0]
    bp = np.concatenate([np.stack([right, np.full(right.size, 1)], 1),
                         np.stack([left, np.full(left.size, 2)], 1)]).astype(np.int32)
    fields = dict(dim=2, nlayers=1, times=np.zeros(1), elevation=np.ones((1, nn)),
                  velocity_x=np.full((1, nn), u), velocity_y=np.zeros((1, nn)))
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(5., 95., npart), rng.uniform(5., 95., npart)], 1)
    return Case(x, y, tri, fields, pos, pool=npart + 16, boundary_points=bp, nopen=2)


def still_column_case(npart=600, h=10., nlayers=5, seed=31, ws=1e-3):
    """Still water over a flat bed at depth h (sigma planes from -h to 0, no
    flow, no turbulence fields), particles spread through the column, sinking
    at the constant settling velocity ws (buoyancy_velocity_model 0)."""
    x, y, tri, mpl = jittered_delaunay(11, 11, 100., 100., seed=3)
    n = x.size
    sig = np.linspace(0., 1., nlayers)
    z = np.concatenate([np.full(n, -h + s_ * h) for s_ in sig])
    zero = np.zeros((1, n * nlayers))
    fields = dict(dim=3, nlayers=nlayers, times=np.zeros(1), elevation=z[None, :], velocity_x=zero,
                  velocity_y=zero, velocity_z=zero)
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(10., 90., npart), rng.uniform(10., 90., npart),
                    -h * rng.uniform(0.02, 0.98, npart)], 1)
    case = Case(x, y, tri, fields, pos, mpl_tri=mpl)
    case.pset_kw = dict(buoyancy_velocity_model=0, buoyancy_velocity=-ws)
    return case


def settling_diameters(rng, n, rho_p=2650.):
    """Stokes diameters of settling velocities w_s ~ LN(median 1 mm/s, sigma 0.5)
    for sediment of density rho_p (a way to draw a diameter spread; the step
    itself uses the reference's constants)."""
    ws = np.exp(np.log(1e-3) + 0.5 * rng.standard_normal(n))
    return np.sqrt(18. * 1e-6 * ws / ((rho_p / 1000. - 1.) * 9.81))


def settling_3d_case(npart=1500, seed=21, spread="regimes", **kw):
    """vortex_3d_case with sediment (density 2650, analytic settling velocity)
    of per-row diameters: "regimes" spreads them log-uniformly over
    [5e-6, 3e-3] m, across the Stokes (< 1e-4), Ruby-Zanke (1e-4 - 1e-3) and
    constant-drag (> 1e-3) branches of the settling law; "coarse" over
    [1.5e-3, 5e-3] m (relaxation times of a tenth of a second, for the explicit
    LSM-2 schemes); "lognormal" draws settling_diameters (most rows in the
    Stokes branch)."""
    case = vortex_3d_case(npart=npart, **kw)
    rng = np.random.default_rng(seed)
    if spread == "regimes":
        case.diameter = np.exp(rng.uniform(np.log(5e-6), np.log(3e-3), npart))
    elif spread == "coarse":
        case.diameter = np.exp(rng.uniform(np.log(1.5e-3), np.log(5e-3), npart))
    else:
        case.diameter = settling_diameters(rng, npart)
    case.pset_kw = dict(particle_density=2650., buoyancy_velocity_model=1)
    return case


def c2_workload(npts_x=2000, npts_y=500, npart=10_000_000, seed=7):
    lx, ly = 20000., 5000.
    x, y, tri, mpl = jittered_delaunay(npts_x, npts_y, lx, ly, seed=seed)
    n = x.size
    k = 2. * np.pi / 2000.
    times = np.array([0., 10000.])
    u = np.stack([np.sin(k * x + ph) * np.cos(k * y) for ph in (0., 0.5)])
    v = np.stack([-np.cos(k * x + ph) * np.sin(k * y) for ph in (0., 0.5)])
    rng = np.random.default_rng(seed)
    depth = 2. + 18. * (0.5 + 0.5 * np.sin(2 * np.pi * x / 7000. + 1.) * np.cos(2 * np.pi * y / 3000.))
    K = 0.1 + 4.9 * (0.5 + 0.5 * np.cos(2 * np.pi * x / 5000.) * np.sin(2 * np.pi * y / 4000. + 0.3))
    eps = np.full(n, 1e-3)
    tke = np.sqrt(K * eps / (0.09 / 0.72))
    fields = dict(dim=2, nlayers=1, times=times, elevation=np.stack([depth, depth]),
                  velocity_x=u, velocity_y=v,
                  opt_fields=np.stack([np.stack([tke, tke]), np.stack([eps, eps])]),
                  opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    pos = np.stack([rng.uniform(0.01 * lx, 0.99 * lx, npart), rng.uniform(0.01 * ly, 0.99 * ly, npart)], 1)
    params = dict(final_time=10000., time_step=5., model=1, diffusion_model=2, schmidt_number=0.72)
    return x, y, tri, fields, pos, params


def c3_workload(npts_x=1000, npts_y=250, nlayers=10, seed=7):
    """The 3D sigma-layer channel: the 20 x 5 km domain with 250k jittered
    points (~500k triangles), 10 uniform sigma planes between the bottom -h
    (h = 2-20 m, smooth) and the surface 0; log-law u(z) along x, v = w = 0;
    k and eps giving a parabolic K(z) = kappa u* z' (1 - z'/h); frozen (one frame)."""
    lx, ly = 20000., 5000.
    x, y, tri, mpl = jittered_delaunay(npts_x, npts_y, lx, ly, seed=seed)
    n = x.size
    h = 2. + 18. * (0.5 + 0.5 * np.sin(2 * np.pi * x / 7000. + 1.) * np.cos(2 * np.pi * y / 3000.))
    kappa, ustar, z0, cmu = 0.41, 0.05, 1e-3, 0.09
    sig = np.linspace(0., 1., nlayers)
    zp = np.concatenate([np.maximum(s_ * h, z0 * 10.) for s_ in sig])   # height above the bottom
    hh = np.tile(h, nlayers)
    z = np.concatenate([-h + s_ * h for s_ in sig])
    u = (ustar / kappa) * np.log(zp / z0)
    frac = np.maximum(1. - zp / hh, 0.05)
    tke = ustar ** 2 / np.sqrt(cmu) * frac
    eps = ustar ** 3 / (kappa * zp) * frac
    fields = dict(dim=3, nlayers=nlayers, times=np.zeros(1), elevation=z[None], velocity_x=u[None],
                  velocity_y=np.zeros((1, n * nlayers)), velocity_z=np.zeros((1, n * nlayers)),
                  opt_fields=np.stack([tke[None], eps[None]]),
                  opt_fields_names=['TURBULENT ENERG.', 'DISSIPATION'])
    return x, y, tri, h, fields


def c3_release(x, y, h, npart, seed=17, lx=20000., ly=5000.):
    """Rows of the 3D sediment workload on the c3_workload channel: uniform over
    the interior, at a uniform fraction 0.1-0.9 of the local depth below the
    surface (the nearest node's depth places them in the column), with
    settling_diameters."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    px, py = rng.uniform(0.02 * lx, 0.98 * lx, npart), rng.uniform(0.02 * ly, 0.98 * ly, npart)
    hn = h[cKDTree(np.stack([x, y], 1)).query(np.stack([px, py], 1))[1]]
    pos = np.stack([px, py, -hn * rng.uniform(0.1, 0.9, npart)], 1)
    return pos, settling_diameters(rng, npart)
