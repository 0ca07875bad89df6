// One particle's time step on the device: random draws at the particle's
// serial offset, LSM-1 / test-kernel / LSM-2 updates, boundary conditions and
// arrival initialization.  The control flow and expression order restate
// cylag/lsm/update_state_lsm1.pyx, update_state_test.pyx,
// update_state_lsm2.pyx, init_lsm2_fluctuations.pyx,
// cylag/boundary_conditions/compute_bc.pyx and cylag/core/solver.pyx:238-549.
#pragma once
#include "cyl_geom.h"
#include "cyl_interp.h"

namespace cyl {

// Thread-local copy of one pool row (LagrangianParticleSet arrays of row i).
struct Part {
  double x[3], up[3], us[3];
  double lx[3], lup[3];  // last_position / last_velocity
  int tri, ll, last_tri;
  int inactive, deleted, nrebound;
  double diameter, volume;
};

__device__ inline void delete_part(Part& p, int dim) {
  // delete_single, lagrangian_particle_set.pyx:754-788 (tag renumbering is done on the host)
  if (p.inactive) return;
  p.inactive = 1;
  p.deleted = 1;
  p.diameter = 0.;
  p.volume = 0.;
  for (int c = 0; c < 3; ++c) {
    p.x[c] = 0.; p.up[c] = 0.; p.us[c] = 0.; p.lx[c] = 0.; p.lup[c] = 0.;
  }
  p.tri = -1;
  p.ll = 0;
  p.last_tri = -1;
}

// ---------------------------------------------------------------------------
// Random draws.  A "unit" is one normal for rng_method 1 and 2 and one call of
// the polar sampler for rng_method 3.  The particle's first unit comes from the
// pool-order scan; its stream position is found through the slow-event table
// of the step (Ziggurat, polar) and the particle then samples sequentially.
// ---------------------------------------------------------------------------
struct Draws {
  const StreamDev* st;
  const JumpTable* jt;
  const JumpTable8* jt8;  // byte-window table (null: bit table only)
  const ZigTables* zt;
  int method;
  bool dry;
  int64_t unit0;   // first unit of this particle
  int64_t nunits;  // units consumed
  int64_t pos0;    // stream position of the first draw
  int64_t ncalls;  // generator calls since pos0
  bool live;
  int err;
  Pcg g;
  // stream position and generator state of each slot's first draw, prepared
  // by a separate pass (null: found here)
  const int64_t* pre_pos;
  const uint64_t* pre_state;
  int slot;

  __device__ uint32_t next() {
    ++ncalls;
    return pcg_next(g);
  }

  // #events with key < u: bucket start, then a short forward scan (binary
  // search past the end of the bucket table)
  __device__ int64_t events_before(int64_t u) const {
    int64_t b = u >> st->ev_bshift, lo;
    if (b < st->n_bucket) {
      lo = st->ev_bucket[b];
      while (lo < st->n_ev && st->ev_key[lo] < u) ++lo;
      return lo;
    }
    lo = 0;
    int64_t hi = st->n_ev;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (st->ev_key[mid] < u) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  }

  __device__ int64_t zig_position(int64_t u) const {
    // number of on-chain slow events with normal index < u, then extra positions
    int64_t lo = events_before(u);
    return u + (lo > 0 ? st->ev_extra[lo - 1] : 0);
  }

  __device__ int64_t polar_position(int64_t q) const {
    // first attempt of call q = q + #{rejected attempts j : r_j - j < q}
    return 2 * (q + events_before(q));
  }

  __device__ void start() {
    live = true;
    if (pre_pos) {
      pos0 = pre_pos[slot];
      if (pos0 < 0) err = E_RNG_WINDOW;
      g.inc = st->inc;
      g.state = pre_state[slot];
      ncalls = 0;
      return;
    }
    if (method == 1) pos0 = unit0;
    else if (method == 2) pos0 = zig_position(unit0);
    else pos0 = polar_position(unit0);
    if (method != 1 && pos0 >= st->window) err = E_RNG_WINDOW;
    g.inc = st->inc;
    g.state = jt8 ? pcg_jump8(st->base, (uint64_t)pos0, *jt8, *jt) : pcg_jump(st->base, (uint64_t)pos0, *jt);
    ncalls = 0;
  }

  __device__ double zig() {
    // random_gaussian_ziggurat, random_utils.pyx:154-194
    double v;
    int32_t hz;
    if (zig_head(v, hz)) return v;
    return zig_tail(hz);
  }

  // The Ziggurat loop up to a base-strip draw: true with the normal in v, or
  // false with the base-strip draw hz, whose tail loop zig_tail finishes
  __device__ bool zig_head(double& v, int32_t& hz_tail) {
    for (;;) {
      int32_t hz = (int32_t)next();
      uint32_t iz = (uint32_t)hz & 127u;
      uint32_t ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;
      if (ahz < zt->kn[iz]) {
        v = hz * zt->wn[iz];
        return true;
      }
      if (iz == 0) {
        hz_tail = hz;
        return false;
      }
      double x = hz * zt->wn[iz];
      if (lt_exp(zt->fn[iz] + u32_to_uniform(next()) * (zt->fn[iz - 1] - zt->fn[iz]), -0.5 * x * x)) {
        v = x;
        return true;
      }
    }
  }

  __device__ double zig_tail(int32_t hz) {
    for (;;) {
      double x = -cr_log_call(u32_to_uniform(next())) * 0.2904764;  // the base strip's tail: rare
      double y = -cr_log_call(u32_to_uniform(next()));
      if (y + y >= x * x) return hz > 0 ? ZIG_R + x : -ZIG_R - x;
    }
  }

  __device__ void polar(double& a, double& b) {
    // random_gaussian_pair, random_utils.pyx:90-108
    double x1 = 0., x2 = 0., w = 2.0;
    while (w >= 1.0 || w == 0.0) {
      x1 = 2.0 * u32_to_uniform(next()) - 1.0;
      x2 = 2.0 * u32_to_uniform(next()) - 1.0;
      w = x1 * x1 + x2 * x2;
    }
    w = sqrt((-2.0 * cr_log(w)) / w);
    a = x1 * w;
    b = x2 * w;
  }

  // generate_random_pair, random_utils.pyx:202-233
  __device__ void pair(double& a, double& b) {
    if (dry) {
      nunits += method == 3 ? 1 : 2;
      a = b = 0.;
      return;
    }
    if (!live) start();
    if (method == 1) {
      a = (u32_to_uniform(next()) - 0.5) * TWO_SQRT3;
      b = (u32_to_uniform(next()) - 0.5) * TWO_SQRT3;
      nunits += 2;
    } else if (method == 2) {
      a = zig();
      b = zig();
      nunits += 2;
    } else {
      polar(a, b);
      nunits += 1;
    }
  }

  // generate_random_single, random_utils.pyx:235-262
  __device__ double single() {
    if (dry) {
      nunits += 1;
      return 0.;
    }
    if (!live) start();
    nunits += 1;
    if (method == 1) return (u32_to_uniform(next()) - 0.5) * TWO_SQRT3;
    if (method == 2) return zig();
    double a, b;
    polar(a, b);
    return a;
  }
};

// Boundary-condition draws (stranding, rough-wall normal perturbation) come
// from a per-particle sub-stream: their serial position depends on the
// particle's own diffusion draws, which a parallel step cannot know in advance.
struct BcDraws {
  Pcg g;
  int method;
  const ZigTables* zt;
  __device__ double uniform() { return u32_to_uniform(pcg_next(g)); }
  __device__ double normal() {
    Draws d;
    d.method = method;
    d.zt = zt;
    d.pre_pos = nullptr;
    d.dry = false;
    d.live = true;
    d.g = g;
    d.nunits = 0;
    d.ncalls = 0;
    double r = d.single();
    g = d.g;
    return r;
  }
};

// Stranded-particle records (compute_bc.pyx:107-122): position at stranding,
// keyed by (step, pool index) so that the host can restore the serial order.
// Deposits on the bed (bottom_deposition) share the buffer, their pool index
// marked with REC_DEPOSIT: a row is recorded at most once (both delete it).
constexpr int REC_DEPOSIT = 1 << 30;
struct StrandRec {
  double* xyz;  // cap * 3
  int2* key;    // cap: (step, pool index)
  int* count;   // [0] records so far, [1] records at the start of the step
  int cap;
  int step, pool;
};

struct Ctx {
  const MeshDev& m;
  const FieldsDev& f;
  const ParamsDev& par;
  const ParticleConst& pc;
  double dt;
  int pdim;  // particle-set dimension
  int err;
  StrandRec* sr = nullptr;  // stranding_output: where stranded positions go
};

#define CYL_CHK_TRI(k)              \
  if ((k) < 0) {                    \
    c.err = E_NOT_IN_TRI;           \
    return;                         \
  }
#define CYL_RET_ERR() \
  if (c.err) return;

// mean fluid velocity at a point, with pseudo-3D corrections (update_state_lsm1.pyx:149-164)
__device__ inline void mean_velocity(Ctx& c, int k, int ll, double x, double y, double z, double& ux,
                                     double& uy, double& uz) {
  CYL_CHK_TRI(k);
  if (c.f.dim == 2) {
    vel2d(c.f, c.m, k, x, y, ux, uy);
    uz = 0.;
    if (c.pdim == 3) {
      double w, hor;
      pseudo3d(c.f, c.m, k, x, y, z, c.dt, ux, uy, w, hor);
      uz += w;
      ux *= hor;
      uy *= hor;
    }
  } else {
    vel3d(c.f, c.m, k, ll, x, y, z, ux, uy, uz, c.err);
  }
}

// ---------------------------------------------------------------------------
// LSM-1
// ---------------------------------------------------------------------------
// RK2/RK3/RK4 (update_state_lsm1.pyx:179-656); returns bnd_j of an early exit
__device__ inline int rk_lsm1(Ctx& c, Part& p, int order) {
  const int d = c.pdim;
  double dt = c.dt, ux = 0., uy = 0., uz = 0.;
  double kx[4] = {0., 0., 0., 0.}, ky[4] = {0., 0., 0., 0.}, kz[4] = {0., 0., 0., 0.};
  mean_velocity(c, p.tri, d == 3 ? p.ll : 0, p.x[0], p.x[1], d == 3 ? p.x[2] : 0., ux, uy, uz);
  if (c.err) return -1;
  p.up[0] = ux; p.up[1] = uy; p.us[0] = ux; p.us[1] = uy;
  if (d == 3) { p.up[2] = uz; p.us[2] = uz; }
  for (int st = 0; st < order - 1; ++st) {
    kx[st] = ux * dt; ky[st] = uy * dt; kz[st] = d == 3 ? uz * dt : 0.;
    double xt, yt, zt = 0.;
    if (order == 2) {
      xt = p.x[0] + 0.5 * ux * dt;
      yt = p.x[1] + 0.5 * uy * dt;
      if (d == 3) zt = p.x[2] + 0.5 * uz * dt;
    } else if (order == 3) {
      double cc = st == 0 ? (1. / 3.) : (2. / 3.);
      xt = p.x[0] + cc * kx[st];
      yt = p.x[1] + cc * ky[st];
      zt = d == 3 ? p.x[2] + cc * kz[st] : 0.;
    } else {
      double cc = st < 2 ? 0.5 : 1.;
      if (st < 2) {
        xt = p.x[0] + cc * kx[st];
        yt = p.x[1] + cc * ky[st];
        zt = d == 3 ? p.x[2] + cc * kz[st] : 0.;
      } else {
        xt = p.x[0] + kx[st];
        yt = p.x[1] + ky[st];
        zt = d == 3 ? p.x[2] + kz[st] : 0.;
      }
    }
    p.last_tri = p.tri;
    int bnd = -1, tl = 0;
    int tt = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], xt, yt, 3, bnd, c.err);
    if (d == 3) tl = z_localize(c.f, c.m, tt, xt, yt, zt);
    if (bnd != -1 || (d == 3 && (tl == -2 || tl == -3))) {
      p.tri = tt;
      if (d == 3) p.ll = tl;
      p.x[0] += p.up[0] * dt;
      p.x[1] += p.up[1] * dt;
      if (d == 3) p.x[2] += p.up[2] * dt;
      return bnd;
    }
    mean_velocity(c, tt, tl, xt, yt, zt, ux, uy, uz);
    if (c.err) return -1;
  }
  if (order == 2) {
    p.x[0] += ux * dt;
    p.x[1] += uy * dt;
    if (d == 3) p.x[2] += uz * dt;
  } else if (order == 3) {
    kx[2] = ux * dt; ky[2] = uy * dt; kz[2] = d == 3 ? uz * dt : 0.;
    p.x[0] += (0.25 * kx[0] + 0.75 * kx[2]);
    p.x[1] += (0.25 * ky[0] + 0.75 * ky[2]);
    if (d == 3) p.x[2] += (0.25 * kz[0] + 0.75 * kz[2]);
  } else {
    kx[3] = ux * dt; ky[3] = uy * dt; kz[3] = d == 3 ? uz * dt : 0.;
    p.x[0] += (kx[0] + 2. * kx[1] + 2. * kx[2] + kx[3]) / 6.;
    p.x[1] += (ky[0] + 2. * ky[1] + 2. * ky[2] + ky[3]) / 6.;
    if (d == 3) p.x[2] += (kz[0] + 2. * kz[1] + 2. * kz[2] + kz[3]) / 6.;
  }
  return -1;
}

// diffusion_lsm1 (update_state_lsm1.pyx:658-726) == diffusion_lsm2 (update_state_lsm2.pyx:1023-1086)
__device__ inline void position_diffusion(Ctx& c, Part& p, Draws& dr, int tri0, int ll0) {
  const int d = c.pdim;
  int dm = c.par.diffusion_model;
  double dt = c.dt, sqrtdt = sqrt(dt);
  double xix = 0., xiy = 0., xiz = 0., diffx = 0., diffy = 0., diffz = 0.;
  double driftx = 0., drifty = 0., driftz = 0.;
  if (dm == 2 || dm == 3) {
    CYL_CHK_TRI(tri0);
    double Kp = 0., dkdx = 0., dkdy = 0., dkdz = 0.;
    diffusivity_gradient(c.f, c.m, tri0, ll0, p.lx[0], p.lx[1], d == 3 ? p.lx[2] : 0., c.par.schmidt_number,
                         Kp, dkdx, dkdy, dkdz, c.err);
    CYL_RET_ERR();
    diffx = sqrt(2. * Kp) * sqrtdt;
    diffy = diffx;
    driftx = dkdx * dt;
    drifty = dkdy * dt;
    if (d == 3) {
      if (dm == 2) diffz = sqrt(2. * c.par.vertical_diffusivity) * sqrtdt;
      else {
        diffz = diffx;
        driftz = dkdz * dt;
      }
    }
  } else {
    diffx = sqrt(2. * c.par.horizontal_diffusivity) * sqrtdt;
    diffy = diffx;
    if (d == 3) diffz = sqrt(2. * c.par.vertical_diffusivity) * sqrtdt;
  }
  if (diffx != 0.) dr.pair(xix, xiy);
  if (d == 3 && diffz != 0.) xiz = dr.single();
  p.x[0] += driftx + diffx * xix;
  p.x[1] += drifty + diffy * xiy;
  if (d == 3) p.x[2] += driftz + diffz * xiz;
}

// buoyancy_velocity, lagrangian_particle_set.pyx:834-890 (model_hr = 0)
__device__ inline double buoyancy_velocity(const ParticleConst& pc, double dp) {
  if (pc.buoyancy_velocity_model == 0) return pc.buoyancy_velocity_value;
  double s = pc.density / pc.rho0, wb;
  if (dp < 1.e-4) wb = fabs(s - 1.) * (GRAV * (dp * dp)) / (18. * NU0);
  else if (dp > 1.e-3) wb = sqrt(fabs(s - 1.) * (4. * dp * GRAV) / (3. * pc.drag_coef));
  else wb = (10. * NU0 / dp) * (sqrt(1. + fabs(s - 1.) * GRAV * cr_pow(dp, 3.0) / (100. * (NU0 * NU0))) - 1.);
  if (s > 1.) wb *= -1.;
  return wb;
}

// update_state_lsm1, update_state_lsm1.pyx:30-133
__device__ inline int update_lsm1(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int bnd_j = -1, tri0 = p.tri, ll0 = d == 3 ? p.ll : 0;
  for (int q = 0; q < d; ++q) { p.lx[q] = p.x[q]; p.lup[q] = p.up[q]; }
  int ts = c.par.time_scheme;
  if (ts == 1) {
    double ux = 0., uy = 0., uz = 0.;
    mean_velocity(c, p.tri, d == 3 ? p.ll : 0, p.x[0], p.x[1], d == 3 ? p.x[2] : 0., ux, uy, uz);
    if (c.err) return -1;
    p.x[0] += ux * c.dt;
    p.x[1] += uy * c.dt;
    p.up[0] = ux; p.up[1] = uy; p.us[0] = ux; p.us[1] = uy;
    if (d == 3) { p.x[2] += uz * c.dt; p.up[2] = uz; p.us[2] = uz; }
  } else if (ts >= 2 && ts <= 4) {
    bnd_j = rk_lsm1(c, p, ts);
    if (c.err) return -1;
  } else {
    c.err = E_SCHEME;
    return -1;
  }
  if (c.par.diffusion_model != 0) {
    position_diffusion(c, p, dr, tri0, ll0);
    if (c.err) return -1;
  }
  if (d == 3) {
    double wb = buoyancy_velocity(c.pc, p.diameter);
    p.x[2] += wb * c.dt;
    p.up[2] += wb;
  }
  if (bnd_j == -1) {
    p.last_tri = p.tri;
    p.tri = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], p.x[0], p.x[1], 3, bnd_j, c.err);
    if (d == 3) p.ll = z_localize(c.f, c.m, p.tri, p.x[0], p.x[1], p.x[2]);
  }
  return bnd_j;
}

// update_state_test, update_state_test.pyx:28-87
__device__ inline int update_test(Ctx& c, Part& p) {
  const int d = c.pdim;
  int bnd_j = -1;
  for (int q = 0; q < d; ++q) { p.lx[q] = p.x[q]; p.lup[q] = p.up[q]; }
  for (int q = 0; q < d; ++q) p.x[q] += p.up[q] * c.dt;
  p.last_tri = p.tri;
  p.tri = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], p.x[0], p.x[1], 2, bnd_j, c.err);
  if (d == 3) p.ll = z_localize(c.f, c.m, p.tri, p.x[0], p.x[1], p.x[2]);
  return bnd_j;
}

// ---------------------------------------------------------------------------
// LSM-2
// ---------------------------------------------------------------------------
// drag_coefficient, lagrangian_particle_set.pyx:790-832
__device__ inline double drag_coefficient(const ParticleConst& pc, double U, double dp) {
  double Re = fmax(EPS, U * dp / NU0);
  if (pc.drag_coef_model == 0) return pc.drag_coef;
  if (pc.drag_coef_model == 1) return Re <= 1000. ? (24. / Re) * (1. + 0.15 * cr_pow(Re, 0.687)) : 0.44;
  if (pc.drag_coef_model == 2) {
    double phi1 = cr_pow(24. / Re, 10.) + cr_pow(21. * cr_pow(Re, -0.67), 10.) + cr_pow(4. * cr_pow(Re, -0.33), 10.) + cr_pow(0.4, 10.);
    double phi2 = 1. / (cr_pow(0.148 * cr_pow(Re, 0.11), -10.) + cr_pow(0.5, -10.));
    double phi3 = cr_pow(1.57 * cr_pow(Re, -1.625) * 1.e8, 10.);
    double phi4 = 1. / (cr_pow(6e-17 * cr_pow(Re, 2.63), -10.) + cr_pow(0.2, -10.));
    return cr_pow(1. / (cr_pow(phi1 + phi2, -1.) + cr_pow(phi3, -1.)) + phi4, 1 / 10.);
  }
  return 0.44;
}

// a1 of a particle of diameter dp with slip velocity unorm (update_state_lsm2.pyx:1346-1350)
__device__ inline double relax_rate_dp(const ParticleConst& pc, double dp, double unorm) {
  unorm = fmax(EPS_LSM2, unorm);
  double Cd = drag_coefficient(pc, unorm, dp);
  return pc.a1c * (1.5 / dp) * Cd * unorm;
}

__device__ inline double relax_rate(const Ctx& c, const Part& p, double unorm) {
  return relax_rate_dp(c.pc, p.diameter, unorm);
}

__device__ inline double rel_norm(int d, double ux, double uy, double uz, double a, double b, double e) {
  double dx = ux - a, dy = uy - b;
  if (d == 2) return sqrt(dx * dx + dy * dy);
  double dz = uz - e;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ inline void accel_at(Ctx& c, int k, int ll, double x, double y, double z, double ux, double uy,
                                double uz, double& ax, double& ay, double& az) {
  ax = ay = az = 0.;
  if (!c.pc.added_mass_force) return;
  if (c.f.dim == 2) accel2d(c.f, c.m, k, x, y, c.dt, ux, uy, ax, ay);
  else accel3d(c.f, c.m, k, ll, x, y, z, c.dt, ux, uy, uz, ax, ay, az, c.err);
}

// euler_lsm2 (update_state_lsm2.pyx:141-217) and rk2/rk3/rk4_lsm2 (:219-1021)
__device__ inline int rk_lsm2(Ctx& c, Part& p, int order) {
  const int d = c.pdim;
  double dt = c.dt, a2c = c.pc.a2c, a0c = c.pc.a0c;
  double ux = 0., uy = 0., uz = 0., ax, ay, az, a1 = 0.;
  double kx[4], ky[4], kz[4], fx[4], fy[4], fz[4];
  int tri = p.tri, ll = d == 3 ? p.ll : 0;
  double xt = p.x[0], yt = p.x[1], zt = d == 3 ? p.x[2] : 0.;
  for (int st = 0; st < order; ++st) {
    mean_velocity(c, tri, ll, xt, yt, zt, ux, uy, uz);
    if (c.err) return -1;
    accel_at(c, tri, ll, xt, yt, zt, ux, uy, uz, ax, ay, az);
    if (c.err) return -1;
    if (st == 0) {
      a1 = relax_rate(c, p, rel_norm(d, ux, uy, uz, p.up[0], p.up[1], p.up[2]));
      kx[0] = p.up[0] * dt;
      ky[0] = p.up[1] * dt;
      fx[0] = (ux - p.up[0]) * dt * a1 + ax * a2c * dt;
      fy[0] = (uy - p.up[1]) * dt * a1 + ay * a2c * dt;
      if (d == 3) {
        kz[0] = p.up[2] * dt;
        fz[0] = (uz - p.up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
      }
    } else {
      // velocity fraction applied to the previous stage's increment
      double cv = order == 3 ? (st == 1 ? (1. / 3.) : (2. / 3.)) : (order == 4 && st == 3 ? 1. : 0.5);
      double ex = p.up[0] + cv * fx[st - 1], ey = p.up[1] + cv * fy[st - 1], ez = 0.;
      double rx = ux - p.up[0] - cv * fx[st - 1], ry = uy - p.up[1] - cv * fy[st - 1];
      double un;
      if (d == 2) un = sqrt(rx * rx + ry * ry);
      else {
        double rz = uz - p.up[2] - cv * fz[st - 1];
        un = sqrt(rx * rx + ry * ry + rz * rz);
        ez = p.up[2] + cv * fz[st - 1];
      }
      a1 = relax_rate(c, p, un);
      kx[st] = ex * dt;
      ky[st] = ey * dt;
      fx[st] = (ux - ex) * dt * a1 + ax * a2c * dt;
      fy[st] = (uy - ey) * dt * a1 + ay * a2c * dt;
      if (d == 3) {
        kz[st] = ez * dt;
        fz[st] = (uz - ez) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
      }
    }
    if (st == order - 1) break;
    double cp = order == 3 ? (st == 0 ? (1. / 3.) : (2. / 3.)) : (order == 4 && st == 2 ? 1. : 0.5);
    if (order == 4 && st == 2) {
      xt = p.x[0] + kx[st];
      yt = p.x[1] + ky[st];
      zt = d == 3 ? p.x[2] + kz[st] : 0.;
    } else {
      xt = p.x[0] + cp * kx[st];
      yt = p.x[1] + cp * ky[st];
      zt = d == 3 ? p.x[2] + cp * kz[st] : 0.;
    }
    p.last_tri = p.tri;
    int bnd = -1;
    tri = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], xt, yt, 3, bnd, c.err);
    if (d == 3) ll = z_localize(c.f, c.m, tri, xt, yt, zt);
    if (bnd != -1 || (d == 3 && (ll == -2 || ll == -3))) {
      p.tri = tri;
      if (d == 3) p.ll = ll;
      p.us[0] = ux;
      p.us[1] = uy;
      p.up[0] += (ux - p.up[0]) * dt * a1 + ax * a2c * dt;
      p.up[1] += (uy - p.up[1]) * dt * a1 + ay * a2c * dt;
      if (d == 3) {
        p.us[2] = uz;
        p.up[2] += (uz - p.up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt;
      }
      p.x[0] += p.up[0] * dt;
      p.x[1] += p.up[1] * dt;
      if (d == 3) p.x[2] += p.up[2] * dt;
      return bnd;
    }
  }
  p.us[0] = ux;
  p.us[1] = uy;
  if (d == 3) p.us[2] = uz;
  if (order == 1) {
    p.up[0] += fx[0];
    p.up[1] += fy[0];
    p.x[0] += p.up[0] * dt;
    p.x[1] += p.up[1] * dt;
    if (d == 3) {
      p.up[2] += fz[0];
      p.x[2] += p.up[2] * dt;
    }
  } else if (order == 2) {
    p.up[0] += fx[1];
    p.up[1] += fy[1];
    p.x[0] += kx[1];
    p.x[1] += ky[1];
    if (d == 3) { p.up[2] += fz[1]; p.x[2] += kz[1]; }
  } else if (order == 3) {
    p.up[0] += (0.25 * fx[0] + 0.75 * fx[2]);
    p.up[1] += (0.25 * fy[0] + 0.75 * fy[2]);
    p.x[0] += (0.25 * kx[0] + 0.75 * kx[2]);
    p.x[1] += (0.25 * ky[0] + 0.75 * ky[2]);
    if (d == 3) {
      p.up[2] += (0.25 * fz[0] + 0.75 * fz[2]);
      p.x[2] += (0.25 * kz[0] + 0.75 * kz[2]);
    }
  } else {
    p.up[0] += (1. / 6.) * (fx[0] + 2. * fx[1] + 2. * fx[2] + fx[3]);
    p.up[1] += (1. / 6.) * (fy[0] + 2. * fy[1] + 2. * fy[2] + fy[3]);
    p.x[0] += (1. / 6.) * (kx[0] + 2. * kx[1] + 2. * kx[2] + kx[3]);
    p.x[1] += (1. / 6.) * (ky[0] + 2. * ky[1] + 2. * ky[2] + ky[3]);
    if (d == 3) {
      p.up[2] += (1. / 6.) * (fz[0] + 2. * fz[1] + 2. * fz[2] + fz[3]);
      p.x[2] += (1. / 6.) * (kz[0] + 2. * kz[1] + 2. * kz[2] + kz[3]);
    }
  }
  return -1;
}

// turbulent energy and dissipation at the particle (update_state_lsm2.pyx:1164-1176)
__device__ inline void keps_at(Ctx& c, const Part& p, double x, double y, double z, double& k, double& e) {
  if (c.f.dim == 2) {
    k = field2d(c.f, c.m, c.f.fid_turbeng, p.tri, -1, x, y);
    e = field2d(c.f, c.m, c.f.fid_dissip, p.tri, -1, x, y);
  } else {
    k = field3d(c.f, c.m, c.f.fid_turbeng, p.tri, p.ll, x, y, z, c.err);
    e = field3d(c.f, c.m, c.f.fid_dissip, p.tri, p.ll, x, y, z, c.err);
  }
}

// lsm2_direct_coefficients, update_state_lsm2.pyx:1240-1327: the drift part
// (beta, D1, D2 and the force coefficient), and um_beta = -expm1(-a1 dt), the
// weight of the fluid velocity in di_lsm2 (update_state_lsm2.pyx:1434), which
// is also D1's numerator when a1 dt >= 0.1; false when a1 or dt is invalid
__device__ inline bool lsm2_drift_coefficients(double a1, double dt, double& beta, double& D1, double& D2,
                                               double& fxc, double& um_beta) {
  double r = a1 * dt;
  if (!isfinite(a1) || !isfinite(dt) || a1 < 0. || dt < 0. || !isfinite(r)) return false;
  beta = cr_exp(-r);
  um_beta = -cr_expm1(-r);
  if (r < 0.1) {
    double phi1 = 1., phi2 = 0.5, term = 1.;
    for (int n = 1; n < 13; ++n) {
      term *= -r / (n + 1.);
      phi1 += term;
      phi2 += term / (n + 2.);
    }
    D1 = dt * phi1;
    D2 = dt * r * phi2;
    fxc = dt * dt * phi2;
  } else {
    D1 = um_beta / a1;
    D2 = dt - D1;
    fxc = D2 / a1;
  }
  return true;
}

// lsm2_direct_coefficients with the covariance terms of the noise
__device__ inline bool lsm2_coefficients(double a1, double dt, double& beta, double& D1, double& D2, double& fxc,
                                         double& um_beta, double& quu, double& qxu, double& qxx) {
  if (!lsm2_drift_coefficients(a1, dt, beta, D1, D2, fxc, um_beta)) return false;
  double r = a1 * dt;
  quu = (r == 0.) ? dt : dt * (-cr_expm1(-2. * r) / (2. * r));
  qxu = 0.5 * D1 * D1;
  if (r < 1.) {
    double term = 1. / 6., power2 = 4., q = 1. / 3.;
    for (int n = 1; n < 24; ++n) {
      term *= -r / (n + 3.);
      power2 *= 2.;
      q += (power2 - 2.) * term;
    }
    q *= dt * dt * dt;
    qxx = q;
  } else {
    qxx = ((dt - 2. * D1 + quu) / a1) / a1;
  }
  return true;
}

// di_lsm2, update_state_lsm2.pyx:1329-1543
__device__ inline void di_lsm2(Ctx& c, Part& p, Draws& dr, int dm, double nux, double nuz, int nuopt) {
  const int d = c.pdim;
  double dt = c.dt, x0[3] = {p.x[0], p.x[1], p.x[2]}, u0[3] = {p.up[0], p.up[1], p.up[2]};
  double ux = 0., uy = 0., uz = 0., ax, ay, az;
  int ll = d == 3 ? p.ll : 0;
  double z0 = d == 3 ? x0[2] : 0.;
  mean_velocity(c, p.tri, ll, x0[0], x0[1], z0, ux, uy, uz);
  CYL_RET_ERR();
  accel_at(c, p.tri, ll, x0[0], x0[1], z0, ux, uy, uz, ax, ay, az);
  CYL_RET_ERR();
  double a1 = relax_rate(c, p, rel_norm(d, ux, uy, uz, u0[0], u0[1], u0[2]));
  double beta, D1, D2, fxc, um_beta, quu, qxu, qxx;
  if (!lsm2_coefficients(a1, dt, beta, D1, D2, fxc, um_beta, quu, qxu, qxx)) {
    c.err = E_LSM2_COEF;
    return;
  }
  double b2[3] = {0., 0., 0.};
  if (dm == 2 || dm == 3) {
    double tk, ep;
    keps_at(c, p, x0[0], x0[1], z0, tk, ep);
    CYL_RET_ERR();
    ep = fmax(ep, EPS);
    double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
    b2[0] = a1 * a1 * C0KOLM * ep * (tml * tml);
    b2[1] = b2[0];
    if (dm == 2) b2[2] = nuopt == 2 ? nuz * nuz : 2. * nuz * a1 * a1;
    else b2[2] = b2[0];
  } else if (dm != 0) {
    if (nuopt == 2) { b2[0] = nux * nux; b2[1] = b2[0]; b2[2] = nuz * nuz; }
    else if (nuopt == 3) { b2[0] = 2. * nux * a1 * a1; b2[1] = b2[0]; b2[2] = 2. * nuz * a1 * a1; }
  }
  for (int k = 0; k < d; ++k) {
    double uk = k == 0 ? ux : (k == 1 ? uy : uz);
    double Fp = k == 0 ? c.pc.a2c * ax : (k == 1 ? c.pc.a2c * ay : c.pc.a2c * az - GRAV * c.pc.a0c);
    double upm = u0[k] * beta + uk * um_beta + Fp * D1;
    double xpm = x0[k] + u0[k] * D1 + uk * D2 + Fp * fxc;
    if (b2[k] == 0. || dt == 0.) {
      p.us[k] = uk;
      p.up[k] = upm;
      p.x[k] = xpm;
      continue;
    }
    double vgp = b2[k] * quu, cov = b2[k] * qxu, vgx = b2[k] * qxx;
    double Lpp = sqrt(fmax(0., vgp));
    double Lxp = Lpp > 0. ? cov / Lpp : 0.;
    double Lxx = sqrt(fmax(0., vgx - Lxp * Lxp));
    double xix, xiy;
    dr.pair(xix, xiy);
    p.us[k] = uk;
    p.up[k] = upm + Lpp * xix;
    p.x[k] = xpm + (Lxp * xix + Lxx * xiy);
  }
}

// euler_maruyama_lsm2_u, update_state_lsm2.pyx:1088-1230
__device__ inline void em_lsm2_u(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int dm = c.par.diffusion_model, nuopt = c.par.diffusion_lsm2_option;
  double dt = c.dt, sqrtdt = sqrt(dt), ux = 0., uy = 0., uz = 0., ax, ay, az;
  double nux = c.par.horizontal_diffusivity, nuz = c.par.vertical_diffusivity;
  int ll = d == 3 ? p.ll : 0;
  double z = d == 3 ? p.x[2] : 0.;
  mean_velocity(c, p.tri, ll, p.x[0], p.x[1], z, ux, uy, uz);
  CYL_RET_ERR();
  accel_at(c, p.tri, ll, p.x[0], p.x[1], z, ux, uy, uz, ax, ay, az);
  CYL_RET_ERR();
  double a1 = relax_rate(c, p, rel_norm(d, ux, uy, uz, p.up[0], p.up[1], p.up[2]));
  double xix = 0., xiy = 0., xiz = 0., diffx = 0., diffy = 0., diffz = 0.;
  if (dm != 0) {
    dr.pair(xix, xiy);
    if (d == 3) xiz = dr.single();
  }
  if (dm == 2 || dm == 3) {
    double tk, ep;
    keps_at(c, p, p.x[0], p.x[1], z, tk, ep);
    CYL_RET_ERR();
    ep = ep > EPS ? ep : EPS;
    double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
    diffx = a1 * tml * sqrt(C0KOLM * ep) * sqrtdt;
    diffy = diffx;
    if (d == 3) diffz = dm == 2 ? (nuopt == 2 ? nuz * sqrtdt : a1 * sqrt(2. * nuz) * sqrtdt) : diffx;
  } else if (nuopt == 2) {
    diffx = nux * sqrtdt; diffy = diffx;
    if (d == 3) diffz = nuz * sqrtdt;
  } else if (nuopt == 3) {
    diffx = a1 * sqrt(2. * nux) * sqrtdt; diffy = diffx;
    if (d == 3) diffz = a1 * sqrt(2. * nuz) * sqrtdt;
  }
  double a2c = c.pc.a2c, a0c = c.pc.a0c;
  p.us[0] = ux;
  p.us[1] = uy;
  p.up[0] += (ux - p.up[0]) * dt * a1 + ax * a2c * dt + diffx * xix;
  p.up[1] += (uy - p.up[1]) * dt * a1 + ay * a2c * dt + diffy * xiy;
  p.x[0] += p.up[0] * dt;
  p.x[1] += p.up[1] * dt;
  if (d == 3) {
    p.us[2] = uz;
    p.up[2] += (uz - p.up[2]) * dt * a1 + az * a2c * dt - GRAV * a0c * dt + diffz * xiz;
    p.x[2] += p.up[2] * dt;
  }
}

// update_state_lsm2, update_state_lsm2.pyx:30-139
__device__ inline int update_lsm2(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int bnd_j = -1, tri0 = p.tri, ll0 = d == 3 ? p.ll : 0;
  int ts = c.par.time_scheme, dm = c.par.diffusion_model, nuopt = c.par.diffusion_lsm2_option;
  for (int q = 0; q < d; ++q) { p.lx[q] = p.x[q]; p.lup[q] = p.up[q]; }
  if (nuopt == 1 || (nuopt >= 2 && dm == 0)) {
    if (ts >= 1 && ts <= 4) bnd_j = rk_lsm2(c, p, ts);
    else if (ts == 5) di_lsm2(c, p, dr, 0, 0., 0., 2);
    else c.err = E_SCHEME;
    if (c.err) return -1;
    if (dm != 0) position_diffusion(c, p, dr, tri0, ll0);
  } else if (nuopt >= 2 && dm != 0) {
    if (ts == 1) em_lsm2_u(c, p, dr);
    else if (ts == 5) di_lsm2(c, p, dr, dm, c.par.horizontal_diffusivity, c.par.vertical_diffusivity, nuopt);
    else c.err = E_SCHEME;
  } else c.err = E_SCHEME;
  if (c.err) return -1;
  p.last_tri = p.tri;
  p.tri = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], p.x[0], p.x[1], 3, bnd_j, c.err);
  if (d == 3) p.ll = z_localize(c.f, c.m, p.tri, p.x[0], p.x[1], p.x[2]);
  return bnd_j;
}

// update_lsm2_fluctuations, init_lsm2_fluctuations.pyx:28-129
__device__ inline void lsm2_fluctuations(Ctx& c, Part& p, Draws& dr, double& f1, double& f2, double& f3) {
  f1 = f2 = f3 = 0.;
  int dm = c.par.diffusion_model, nuopt = c.par.diffusion_lsm2_option;
  if (dm == 0) return;
  double unorm = EPS_LSM2;
  double Cd = drag_coefficient(c.pc, unorm, p.diameter);
  double a1 = c.pc.a1c * (1.5 / p.diameter) * Cd * unorm;
  double nux = c.par.horizontal_diffusivity, nuz = c.par.vertical_diffusivity;
  if (a1 == 0.) {
    if (nuopt == 2 && (((dm != 2 && dm != 3) && nux != 0.) || (c.pdim == 3 && dm != 3 && nuz != 0.)))
      c.err = E_LSM2_COEF;
    return;
  }
  double b0, b1, bz;
  if (dm == 2 || dm == 3) {
    double tk, ep;
    keps_at(c, p, p.x[0], p.x[1], c.pdim == 3 ? p.x[2] : 0., tk, ep);
    CYL_RET_ERR();
    ep = fmax(ep, EPS);
    double tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
    b0 = a1 * C0KOLM * ep * (tml * tml);
    b1 = b0;
    if (dm == 2) bz = nuopt == 2 ? (nuz * nuz) / a1 : 2. * nuz * a1;
    else bz = b0;
  } else if (nuopt == 2) {
    b0 = (nux * nux) / a1; b1 = b0; bz = (nuz * nuz) / a1;
  } else {
    b0 = 2. * nux * a1; b1 = b0; bz = 2. * nuz * a1;
  }
  if (b0 > 0.) {
    double xix, xiy;
    dr.pair(xix, xiy);
    f1 = sqrt(0.5 * b0) * xix;
    f2 = sqrt(0.5 * b1) * xiy;
  }
  if (c.pdim == 3 && bz > 0.) f3 = sqrt(0.5 * bz) * dr.single();
}

// ---------------------------------------------------------------------------
// LSM-3 (update_state_lsm3.pyx): the seen-fluid velocity Us is an
// Ornstein-Uhlenbeck process relaxing to the mean flow over TL; the particle
// velocity relaxes to Us at the drag rate a1 (plus added mass A2 dUs).  Over a
// frozen step the (Us, Ur = Up - A2 Us, Xp) system of each component is linear
// with lower-triangular transition exp(dt A), A = [[-b,0,0],[a(1-A2),-a,0],[A2,1,0]].
// ---------------------------------------------------------------------------
constexpr double EPS_LSM3 = 1.e-8;        // constants.pyx:8
constexpr double LSM3_CLIP_STEP = 1.e-4;  // constants.pyx:11
constexpr double DBL_EPS = 2.220446049250313e-16;

struct Lsm3Coef {
  double phi[9], q[9], force[3], flow[3];
};

// input checks of the coefficient routines (update_state_lsm3.pyx:74-81)
__device__ inline bool lsm3_rates_ok(double a, double a2, double tl, double dt) {
  if (!isfinite(a) || !isfinite(a2) || !isfinite(tl) || !isfinite(dt) || a < 0. || tl <= 0. || dt < 0.)
    return false;
  double b = 1. / tl;
  return isfinite(b) && isfinite(a * dt) && isfinite(b * dt) && isfinite(a * (1. - a2) * dt);
}

__device__ inline void lsm3_identity(Lsm3Coef& C) {
  for (int k = 0; k < 9; ++k) {
    C.phi[k] = 0.;
    C.q[k] = 0.;
  }
  C.phi[0] = C.phi[4] = C.phi[8] = 1.;
  for (int k = 0; k < 3; ++k) C.force[k] = C.flow[k] = 0.;
}

__device__ inline bool lsm3_finite(const Lsm3Coef& C) {
  bool ok = isfinite(C.force[1]) && isfinite(C.force[2]);
  for (int k = 0; k < 9; ++k) ok = ok && isfinite(C.phi[k]) && isfinite(C.q[k]);
  for (int k = 0; k < 3; ++k) ok = ok && isfinite(C.flow[k]);
  return ok;
}

// _lsm3_direct_coefficients, update_state_lsm3.pyx:32-217: 24-term Taylor
// series of the normalized system at max(a h, b h) <= 1/2, then step doublings
__device__ __noinline__ bool lsm3_series_coefficients(double a, double a2, double tl, double dt, Lsm3Coef& C) {
  if (!lsm3_rates_ok(a, a2, tl, dt)) return false;
  double b = 1. / tl;
  lsm3_identity(C);
  if (dt == 0.) return true;
  double h = dt;
  int nd = 0;
  while (h > 0. && fmax(a * h, b * h) > 0.5) {
    h *= 0.5;
    ++nd;
  }
  double ah = a * h, bh = b * h, kh = a * (1. - a2) * h;
  // second moments of (s, p, x, si): ss sp sx pp px xx si pi xi ii
  double t[10] = {1., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  double qb[10] = {1., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
  double s = 1., pv = 0., xv = 0., si = 0.;
  double sum_p = 0., sum_x = 0., sum_si = 0., int_p = 0., int_x = 0., int_si = 0.;
  double bterm = 1., phi1 = 1., phi2 = 0.5;
  for (int n = 1; n < 25; ++n) {
    double dv = n + 1., u[10];
    u[0] = -2. * bh * t[0] / dv;
    u[1] = (t[0] - (ah + bh) * t[1]) / dv;
    u[2] = (t[1] - bh * t[2]) / dv;
    u[3] = (2. * t[1] - 2. * ah * t[3]) / dv;
    u[4] = (t[2] + t[3] - ah * t[4]) / dv;
    u[5] = 2. * t[4] / dv;
    u[6] = (t[0] - bh * t[6]) / dv;
    u[7] = (t[6] + t[1] - ah * t[7]) / dv;
    u[8] = (t[7] + t[2]) / dv;
    u[9] = 2. * t[6] / dv;
    for (int k = 0; k < 10; ++k) {
      t[k] = u[k];
      qb[k] += t[k];
    }
    double sn = -bh * s / n, pn = (s - ah * pv) / n, xn = pv / n, sin_ = s / n;
    s = sn;
    pv = pn;
    xv = xn;
    si = sin_;
    sum_p += pv;
    sum_x += xv;
    sum_si += si;
    int_p += pv / (n + 1.);
    int_x += xv / (n + 1.);
    int_si += si / (n + 1.);
    bterm *= -ah / n;
    phi1 += bterm / (n + 1.);
    phi2 += bterm / ((n + 1.) * (n + 2.));
  }
  double* phi = C.phi;
  double* q = C.q;
  q[0] = h * qb[0];
  q[1] = q[3] = h * kh * qb[1];
  q[2] = q[6] = h * h * (kh * qb[2] + a2 * qb[6]);
  q[4] = h * kh * kh * qb[3];
  q[5] = q[7] = h * h * kh * (kh * qb[4] + a2 * qb[7]);
  q[8] = h * h * h * (kh * kh * qb[5] + 2. * kh * a2 * qb[8] + a2 * a2 * qb[9]);
  phi[0] = cr_exp(-bh);
  phi[3] = kh * sum_p;
  phi[4] = cr_exp(-ah);
  phi[6] = h * (kh * sum_x + a2 * sum_si);
  phi[7] = h * phi1;
  C.force[1] = phi[7];
  C.force[2] = h * h * phi2;
  C.flow[0] = -cr_expm1(-bh);
  C.flow[1] = kh * bh * int_p;
  C.flow[2] = h * bh * (kh * int_x + a2 * int_si);
  for (int n = 0; n < nd; ++n) {
    // Q(2h) = Q(h) + Phi(h) Q(h) Phi(h)^T, Phi lower triangular
    double w[9], qn[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double v = 0.;
        for (int k = 0; k <= r; ++k) v += phi[3 * r + k] * q[3 * k + c];
        w[3 * r + c] = v;
      }
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c <= r; ++c) {
        double v = q[3 * r + c];
        for (int k = 0; k <= c; ++k) v += w[3 * r + k] * phi[3 * c + k];
        qn[3 * r + c] = qn[3 * c + r] = v;
      }
    for (int k = 0; k < 9; ++k) q[k] = qn[k];
    C.flow[2] = 2. * C.flow[2] + phi[6] * C.flow[0] + phi[7] * C.flow[1];
    C.flow[1] = (1. + phi[4]) * C.flow[1] + phi[3] * C.flow[0];
    C.flow[0] = -cr_expm1(-2. * b * h);
    phi[6] = phi[6] * (1. + phi[0]) + phi[7] * phi[3];
    phi[3] *= phi[0] + phi[4];
    C.force[2] = 2. * C.force[2] + phi[7] * phi[7];
    phi[7] *= 1. + phi[4];
    C.force[1] = phi[7];
    h *= 2.;
    phi[0] = cr_exp(-b * h);
    phi[4] = cr_exp(-a * h);
  }
  return lsm3_finite(C);
}

__device__ inline double lsm3_phi1(double z) {  // (exp(z) - 1)/z, update_state_lsm3.pyx:219-223
  return z == 0. ? 1. : cr_expm1(z) / z;
}

__device__ inline double lsm3_exp_divdiff(double x, double y) {  // update_state_lsm3.pyx:225-227
  return cr_exp(-fmin(x, y)) * lsm3_phi1(-fabs(x - y));
}

__device__ inline double lsm3_phi2(double z) {  // (exp(z) - 1 - z)/z^2, update_state_lsm3.pyx:229-234
  if (fabs(z) < 0.05)
    return 0.5 + z * (1. / 6. + z * (1. / 24. + z * (1. / 120. + z * (1. / 720. +
           z * (1. / 5040. + z * (1. / 40320. + z / 362880.))))));
  return (cr_expm1(z) - z) / (z * z);
}

// _lsm3_closed_covariance, update_state_lsm3.pyx:236-266
__device__ inline void lsm3_closed_covariance(double a, double b, double a2, double h, double* q) {
  double d = a - b, c = a * (1. - a2) / d, bb = (1. - a2) / d, aa = (a2 + c) / b;
  double sb = h * lsm3_phi1(-b * h), sa = h * lsm3_phi1(-a * h);
  double s2b = h * lsm3_phi1(-2. * b * h), s2a = h * lsm3_phi1(-2. * a * h);
  double sab = h * lsm3_phi1(-(a + b) * h);
  q[0] = s2b;
  q[1] = q[3] = c * (s2b - sab);
  q[2] = q[6] = aa * (sb - s2b) - bb * (sb - sab);
  q[4] = c * c * (s2b - 2. * sab + s2a);
  q[5] = q[7] = c * (aa * (sb - s2b - sa + sab) - bb * (sb - sab - sa + s2a));
  q[8] = (aa * aa * (h - 2. * sb + s2b) - 2. * aa * bb * (h - sb - sa + sab) + bb * bb * (h - 2. * sa + s2a));
}

// _lsm3_clipped_covariance, update_state_lsm3.pyx:268-310
__device__ __noinline__ void lsm3_clipped_covariance(double a, double b, double a2, double h, double* q) {
  double d = a - b, m = 0.5 * (a + b), x = m * h, delta;
  if (x > 1.) delta = fmin(0.5, LSM3_CLIP_STEP * sqrt(x)) / h;
  else delta = fmin(0.5, LSM3_CLIP_STEP / sqrt(x)) * m;
  if (fabs(d) >= delta) {
    lsm3_closed_covariance(a, b, a2, h, q);
  } else {
    double qp[9], qm[9];
    lsm3_closed_covariance(m + 0.5 * delta, m - 0.5 * delta, a2, h, qp);
    lsm3_closed_covariance(m - 0.5 * delta, m + 0.5 * delta, a2, h, qm);
    double w = 0.5 * (1. + d / delta);
    for (int k = 0; k < 9; ++k) q[k] = qm[k] + w * (qp[k] - qm[k]);
  }
  for (int k = 0; k < 3; ++k) q[4 * k] = fmax(q[4 * k], 0.);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < r; ++c) {
      double bound = sqrt(q[4 * r] * q[4 * c]);
      q[3 * r + c] = q[3 * c + r] = fmax(-bound, fmin(bound, q[3 * r + c]));
    }
}

// _lsm3_clipped_coefficients, update_state_lsm3.pyx:312-374
__device__ __noinline__ bool lsm3_clipped_coefficients(double a, double a2, double tl, double dt, Lsm3Coef& C) {
  if (!lsm3_rates_ok(a, a2, tl, dt)) return false;
  double b = 1. / tl;
  lsm3_identity(C);
  if (dt == 0.) return true;
  double ah = a * dt, bh = b * dt, k = a * (1. - a2);
  double alpha = cr_exp(-bh), beta = cr_exp(-ah), divdiff = lsm3_exp_divdiff(ah, bh);
  double ib = dt * lsm3_phi1(-bh), ia = dt * lsm3_phi1(-ah);
  C.phi[0] = alpha;
  C.phi[3] = k * dt * divdiff;
  C.phi[4] = beta;
  C.phi[6] = ib - (1. - a2) * dt * divdiff;
  C.phi[7] = ia;
  C.force[1] = ia;
  C.force[2] = dt * dt * lsm3_phi2(-ah);
  C.flow[0] = -cr_expm1(-bh);
  C.flow[1] = k * dt * (lsm3_phi1(-ah) - divdiff);
  C.flow[2] = (a2 * (dt - ib) + (1. - a2) * (dt - ia - ib + dt * divdiff));
  lsm3_clipped_covariance(a, b, a2, dt, C.q);
  return lsm3_finite(C);
}

__device__ inline bool lsm3_coefficients(double a, double a2, double tl, double dt, int edge_opt, Lsm3Coef& C) {
  return edge_opt == 2 ? lsm3_clipped_coefficients(a, a2, tl, dt, C) : lsm3_series_coefficients(a, a2, tl, dt, C);
}

// _lsm3_cholesky, update_state_lsm3.pyx:520-553 (false: not factorable)
__device__ inline bool lsm3_cholesky(const double* q, double* L) {
  for (int k = 0; k < 9; ++k) L[k] = 0.;
  for (int r = 0; r < 3; ++r) {
    double pivot = q[3 * r + r];
    for (int c = 0; c < r; ++c) {
      double res = q[3 * r + c];
      for (int k = 0; k < c; ++k) res -= L[3 * r + k] * L[3 * c + k];
      if (L[3 * c + c] > 0.) {
        L[3 * r + c] = res / L[3 * c + c];
      } else {
        double tol = (128. * DBL_EPS * sqrt(fabs(q[3 * r + r])) * sqrt(fabs(q[3 * c + c])));
        if (fabs(res) > tol) return false;
      }
      pivot -= L[3 * r + c] * L[3 * r + c];
    }
    if (pivot < 0.) {
      if (-pivot > 128. * DBL_EPS * fabs(q[3 * r + r])) return false;
      pivot = 0.;
    }
    L[3 * r + r] = sqrt(pivot);
  }
  return true;
}

// _lsm3_cholesky_clipped, update_state_lsm3.pyx:555-585
__device__ inline void lsm3_cholesky_clipped(const double* q, double* L) {
  for (int k = 0; k < 9; ++k) L[k] = 0.;
  for (int r = 0; r < 3; ++r) {
    double rem = fmax(q[3 * r + r], 0.);
    for (int c = 0; c < r; ++c) {
      if (L[3 * c + c] > 0.) {
        double v = q[3 * r + c];
        for (int k = 0; k < c; ++k) v -= L[3 * r + k] * L[3 * c + k];
        v /= L[3 * c + c];
        if (v * v > rem) v = copysign(sqrt(rem), v);
        L[3 * r + c] = v;
        rem -= v * v;
      }
    }
    L[3 * r + r] = sqrt(fmax(rem, 0.));
  }
}

// mean velocity (pseudo-3D corrected) and mean fluid acceleration, which LSM-3
// always evaluates (update_state_lsm3.pyx:702-727, 876-909)
__device__ inline void lsm3_mean_flow(Ctx& c, const Part& p, const double* x, double* u, double* acc) {
  const int d = c.pdim;
  int ll = d == 3 ? p.ll : 0;
  double z = d == 3 ? x[2] : 0.;
  mean_velocity(c, p.tri, ll, x[0], x[1], z, u[0], u[1], u[2]);
  CYL_RET_ERR();
  acc[0] = acc[1] = acc[2] = 0.;
  if (c.f.dim == 2) accel2d(c.f, c.m, p.tri, x[0], x[1], c.dt, u[0], u[1], acc[0], acc[1]);
  else accel3d(c.f, c.m, p.tri, ll, x[0], x[1], x[2], c.dt, u[0], u[1], u[2], acc[0], acc[1], acc[2], c.err);
}

// a1 = a1c (1.5/dp) unorm Cd with the slip floored (update_state_lsm3.pyx:729-735)
__device__ inline double lsm3_relax_rate(const Ctx& c, const Part& p, double unorm) {
  unorm = fmax(EPS_LSM2, unorm);
  double Cd = drag_coefficient(c.pc, unorm, p.diameter);
  return c.pc.a1c * (1.5 / p.diameter) * unorm * Cd;
}

// euler_maruyama_lsm3, update_state_lsm3.pyx:664-820
__device__ inline void em_lsm3(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int dm = c.par.diffusion_model;
  double dt = c.dt, sqrtdt = sqrt(dt), u[3], acc[3];
  double tlh = c.par.lsm3_tl_horizontal, tlv = c.par.lsm3_tl_vertical;
  lsm3_mean_flow(c, p, p.x, u, acc);
  CYL_RET_ERR();
  double du0 = p.us[0] - p.up[0], du1 = p.us[1] - p.up[1], unorm;
  if (d == 2) {
    unorm = sqrt(du0 * du0 + du1 * du1);
  } else {
    double du2 = p.us[2] - p.up[2];
    unorm = sqrt(du0 * du0 + du1 * du1 + du2 * du2);
  }
  double a1 = lsm3_relax_rate(c, p, unorm);
  double xix, xiy, xiz = 0., tk = 0., ep = 0., tml, unstml, diffx, diffz;
  dr.pair(xix, xiy);
  if (d == 3) xiz = dr.single();
  if (dm == 2 || dm == 3) {
    keps_at(c, p, p.x[0], p.x[1], d == 3 ? p.x[2] : 0., tk, ep);
    CYL_RET_ERR();
    ep = fmax(ep, EPS);
    tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
    unstml = 1. / fmax(EPS_LSM3, tml);
    diffx = sqrt(C0KOLM * ep) * sqrtdt;
  } else {
    unstml = 1. / tlh;
    diffx = sqrt(2. * c.par.horizontal_diffusivity) * sqrtdt / tlh;
  }
  double a2c = c.pc.a2c;
  double dusx = (acc[0] + (u[0] - p.us[0]) * unstml) * dt + diffx * xix;
  double dusy = (acc[1] + (u[1] - p.us[1]) * unstml) * dt + diffx * xiy;
  p.up[0] += (p.us[0] - p.up[0]) * dt * a1;
  p.up[1] += (p.us[1] - p.up[1]) * dt * a1;
  p.us[0] += dusx;
  p.us[1] += dusy;
  if (c.pc.added_mass_force) {
    p.up[0] += a2c * dusx;
    p.up[1] += a2c * dusy;
  }
  p.x[0] += p.up[0] * dt;
  p.x[1] += p.up[1] * dt;
  if (d == 3) {
    if (dm == 1 || dm == 2) {
      unstml = 1. / tlv;
      diffz = sqrt(2. * c.par.vertical_diffusivity) * sqrtdt / tlv;
    } else {
      tml = (tk / ep) / (0.5 + (3. / 4.) * C0KOLM);
      unstml = 1. / fmax(EPS_LSM3, tml);
      diffz = sqrt(C0KOLM * ep) * sqrtdt;
    }
    double dusz = (acc[2] + (u[2] - p.us[2]) * unstml) * dt + diffz * xiz;
    p.up[2] += (p.us[2] - p.up[2]) * dt * a1 - GRAV * c.pc.a0c * dt;
    p.us[2] += dusz;
    if (c.pc.added_mass_force) p.up[2] += a2c * dusz;
    p.x[2] += p.up[2] * dt;
  }
}

// di_lsm3, update_state_lsm3.pyx:822-995
__device__ __noinline__ void di_lsm3(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int dm = c.par.diffusion_model, eo = c.par.lsm3_edge_opt;
  double dt = c.dt;
  if (dt == 0.) return;
  double x0[3] = {0., 0., 0.}, u0[3] = {0., 0., 0.}, s0[3] = {0., 0., 0.}, u[3], acc[3], L[9];
  for (int k = 0; k < d; ++k) {
    x0[k] = p.x[k];
    u0[k] = p.up[k];
    s0[k] = p.us[k];
  }
  double a2 = c.pc.added_mass_force ? c.pc.a2c : 0.;
  double tlh = c.par.lsm3_tl_horizontal, tlv = c.par.lsm3_tl_vertical, tk = 0., ep = 0., sigh, sigv = 0.;
  lsm3_mean_flow(c, p, x0, u, acc);
  CYL_RET_ERR();
  if (dm == 2 || dm == 3) {
    keps_at(c, p, x0[0], x0[1], x0[2], tk, ep);
    CYL_RET_ERR();
  }
  double unorm = (s0[0] - u0[0]) * (s0[0] - u0[0]) + (s0[1] - u0[1]) * (s0[1] - u0[1]);
  if (d == 3) unorm += (s0[2] - u0[2]) * (s0[2] - u0[2]);
  double a1 = lsm3_relax_rate(c, p, sqrt(unorm));
  if (dm == 2 || dm == 3) {
    ep = fmax(ep, EPS);
    tlh = fmax(EPS_LSM3, (tk / ep) / (0.5 + (3. / 4.) * C0KOLM));
    sigh = sqrt(C0KOLM * ep);
  } else {
    sigh = sqrt(2. * c.par.horizontal_diffusivity) / tlh;
  }
  if (d == 3) {
    if (dm == 3) {
      tlv = tlh;
      sigv = sigh;
    } else {
      sigv = sqrt(2. * c.par.vertical_diffusivity) / tlv;
    }
  }
  Lsm3Coef C;
  if (!lsm3_coefficients(a1, a2, tlh, dt, eo, C)) {
    c.err = E_LSM3_COEF;
    return;
  }
  bool factored = false;
  double sig = sigh;
  for (int k = 0; k < d; ++k) {
    double tlk = tlh;
    if (k == 2) {
      tlk = tlv;
      sig = sigv;
      if (tlv != tlh) {
        if (!lsm3_coefficients(a1, a2, tlv, dt, eo, C)) {
          c.err = E_LSM3_COEF;
          return;
        }
        factored = false;
      }
    }
    double uk = u[k], Fp = k == 2 ? -GRAV * c.pc.a0c : 0.;
    double ur0 = u0[k] - a2 * s0[k];
    double us_m = (s0[k] * C.phi[0] + uk * C.flow[0] + acc[k] * tlk * C.flow[0]);
    double ur_m = (ur0 * C.phi[4] + s0[k] * C.phi[3] + uk * C.flow[1] + acc[k] * tlk * C.flow[1] + Fp * C.force[1]);
    double xp_m = (x0[k] + ur0 * C.force[1] + s0[k] * C.phi[6] + uk * C.flow[2] + acc[k] * tlk * C.flow[2] +
                   Fp * C.force[2]);
    if (sig > 0. && C.q[0] > 0.) {
      if (!factored) {
        if (eo == 2) {
          lsm3_cholesky_clipped(C.q, L);
        } else if (!lsm3_cholesky(C.q, L)) {
          c.err = E_LSM3_CHOL;
          return;
        }
        factored = true;
      }
      double xix, xiy = 0., xiz = 0.;
      if (L[4] > 0.) dr.pair(xix, xiy);
      else xix = dr.single();
      if (L[8] > 0.) xiz = dr.single();
      us_m += sig * L[0] * xix;
      ur_m += sig * (L[3] * xix + L[4] * xiy);
      xp_m += sig * (L[6] * xix + L[7] * xiy + L[8] * xiz);
    }
    p.us[k] = us_m;
    p.up[k] = ur_m + a2 * us_m;
    p.x[k] = xp_m;
  }
}

// update_state_lsm3, update_state_lsm3.pyx:587-662
__device__ inline int update_lsm3(Ctx& c, Part& p, Draws& dr) {
  const int d = c.pdim;
  int bnd_j = -1;
  for (int q = 0; q < d; ++q) {
    p.lx[q] = p.x[q];
    p.lup[q] = p.up[q];
  }
  if (c.par.time_scheme == 1) em_lsm3(c, p, dr);
  else if (c.par.time_scheme == 5) di_lsm3(c, p, dr);
  else c.err = E_SCHEME;
  if (c.err) return -1;
  p.last_tri = p.tri;
  p.tri = xy_localize(c.m, p.last_tri, p.lx[0], p.lx[1], p.x[0], p.x[1], 3, bnd_j, c.err);
  if (d == 3) p.ll = z_localize(c.f, c.m, p.tri, p.x[0], p.x[1], p.x[2]);
  return bnd_j;
}

// update_lsm3_fluctuations, init_lsm3_fluctuations.pyx:28-90
__device__ inline void lsm3_fluctuations(Ctx& c, Part& p, Draws& dr, double& f1, double& f2, double& f3) {
  f1 = f2 = f3 = 0.;
  int dm = c.par.diffusion_model;
  double sigh;
  if (dm == 2 || dm == 3) {
    double tk, ep;
    keps_at(c, p, p.x[0], p.x[1], c.pdim == 3 ? p.x[2] : 0., tk, ep);
    CYL_RET_ERR();
    ep = fmax(ep, EPS);
    double tml = fmax(EPS_LSM3, (tk / ep) / (0.5 + (3. / 4.) * C0KOLM));
    sigh = sqrt(0.5 * C0KOLM * ep * tml);
  } else {
    sigh = sqrt(c.par.horizontal_diffusivity / c.par.lsm3_tl_horizontal);
  }
  if (sigh > 0.) {
    double xix, xiy;
    dr.pair(xix, xiy);
    f1 = sigh * xix;
    f2 = sigh * xiy;
  }
  if (c.pdim == 3) {
    double sigv = dm == 3 ? sigh : sqrt(c.par.vertical_diffusivity / c.par.lsm3_tl_vertical);
    if (sigv > 0.) f3 = sigv * dr.single();
  }
}

// ---------------------------------------------------------------------------
// Arrival: initialize_particle_state, solver.pyx:238-380.  `tri_found` is the
// element found on the host (trifinder) or -3 to search on the device.
// ---------------------------------------------------------------------------
__device__ inline void init_particle(Ctx& c, Part& p, Draws& dr, int tri_found) {
  const int d = c.pdim;
  int bj;
  p.tri = tri_found != -3 ? tri_found : xy_localize(c.m, -1, 0., 0., p.x[0], p.x[1], 0, bj, c.err);
  if (p.tri == -1) delete_part(p, d);
  if (d == 3 && !p.inactive) {
    double zs = field2d(c.f, c.m, c.f.fid_zs, p.tri, c.f.nlayers - 1, p.x[0], p.x[1]);
    double zb = c.f.dim == 2 ? field2d(c.f, c.m, c.f.fid_zb, p.tri, 0, p.x[0], p.x[1])
                             : field2d(c.f, c.m, c.f.fid_zs, p.tri, 0, p.x[0], p.x[1]);
    p.ll = lower_layer_index(c.f.nlayers, zs, zb, p.x[2]);
    if (p.x[2] < zb) p.x[2] = zb + EPS;
    if (p.x[2] > zs) p.x[2] = zs - EPS;
  }
  if (p.inactive) return;
  if (c.par.particle_velocity_init == 1) {
    double ux, uy, uz = 0.;
    if (c.f.dim == 2) {
      vel2d(c.f, c.m, p.tri, p.x[0], p.x[1], ux, uy);
    } else {
      vel3d(c.f, c.m, p.tri, p.ll, p.x[0], p.x[1], p.x[2], ux, uy, uz, c.err);
      CYL_RET_ERR();
    }
    p.up[0] = ux; p.up[1] = uy; p.us[0] = ux; p.us[1] = uy;
    if (d == 3 && c.f.dim == 3) { p.up[2] = uz; p.us[2] = uz; }
    if (c.par.model == 2 && c.par.diffusion_lsm2_option >= 2 && c.par.diffusion_model != 0) {
      double f1, f2, f3;
      lsm2_fluctuations(c, p, dr, f1, f2, f3);
      CYL_RET_ERR();
      p.up[0] += f1;
      p.up[1] += f2;
      if (d == 3) p.up[2] += f3;
    }
    for (int q = 0; q < d; ++q) p.us[q] = p.up[q];
    if (c.par.model == 3) {
      double f1, f2, f3;
      lsm3_fluctuations(c, p, dr, f1, f2, f3);
      CYL_RET_ERR();
      p.us[0] += f1;
      p.us[1] += f2;
      if (d == 3) p.us[2] += f3;
    }
  }
}

// ---------------------------------------------------------------------------
// Boundary conditions, compute_bc.pyx:35-395
// ---------------------------------------------------------------------------
__device__ inline void reflect_velocity(double* v, const double nn[2], double damp, int d, bool damped) {
  double up0 = v[0];
  v[0] = up0 * nn[0] + v[1] * nn[1];
  v[1] = -up0 * nn[1] + v[1] * nn[0];
  v[0] = -1. * v[0];
  up0 = v[0];
  v[0] = up0 * nn[0] - v[1] * nn[1];
  v[1] = up0 * nn[1] + v[1] * nn[0];
  if (damped) {
    v[0] *= 1. - damp;
    v[1] *= 1. - damp;
    if (d == 3) v[2] *= 1. - damp;
  }
}

__device__ inline void rough_normal(const Ctx& c, BcDraws& bd, double nn[2]) {
  double rough = c.par.wall_roughness_angle * DEG_TO_RAD, max_theta = 85.0 * DEG_TO_RAD;
  if (c.par.boundary_conditions_type == 2 && rough > 0.) {
    double th = bd.normal() * rough;
    th = fmax(-max_theta, fmin(max_theta, th));
    double ct = cos(th), st = sin(th);
    double a = nn[0] * ct - nn[1] * st, b = nn[0] * st + nn[1] * ct;
    nn[0] = a;
    nn[1] = b;
  }
}

__device__ inline void compute_bc(Ctx& c, Part& p, int bnd_j, BcDraws& bd, int* bnd_stats, int nopen) {
  const int d = c.pdim;
  const MeshDev& m = c.m;
  double damp = fmin(c.par.rebound_damping_coef, 1.0 - EPS_REBOUND);
  double strand = c.par.stranding_probability;
  int strand_cond = 0;
  if (bnd_j != -1) {
    int lab = m.edge_labels[bnd_j];
    if (lab == 0 && strand > 0.) strand_cond = bd.uniform() < strand;
    if (strand_cond && c.sr && c.par.stranding_output) {
      int q = atomicAdd(c.sr->count, 1);
      if (q < c.sr->cap) {
        c.sr->xyz[3 * q] = p.x[0];
        c.sr->xyz[3 * q + 1] = p.x[1];
        c.sr->xyz[3 * q + 2] = d == 2 ? 0.0 : p.x[2];
        c.sr->key[q] = make_int2(c.sr->step, c.sr->pool);
      }
    }
    if (lab == 0 && !strand_cond) {
      int bjj = bnd_j, rc = 0;
      while (p.tri == PART_LOC_O && rc < c.par.max_wall_rebounds) {
        int i0 = m.edges[2 * bjj], i1 = m.edges[2 * bjj + 1];
        double v0x = m.x[i0], v0y = m.y[i0], v1x = m.x[i1], v1y = m.y[i1];
        // compute_symetric_point2d, intersections.pyx:27-62
        double n0 = v0y - v1y, n1 = v1x - v0x;
        double norm2 = n0 * n0 + n1 * n1;
        double dist = (p.x[0] - v0x) * n0 + (p.x[1] - v0y) * n1;
        double coef = 1. - damp / 2.;
        double px = p.x[0] - 2. * ((dist * coef) / norm2) * n0;
        double py = p.x[1] - 2. * ((dist * coef) / norm2) * n1;
        p.x[0] = px;
        p.x[1] = py;
        p.tri = xy_localize(m, p.last_tri, p.lx[0], p.lx[1], px, py, 3, bjj, c.err);
        ++p.nrebound;
        if (p.tri == PART_LOC_O && bjj != -1 && m.edge_labels[bjj] >= 1) {
          if (c.par.bnd_statistics && bnd_stats && m.edge_labels[bjj] - 1 < nopen)
            atomicAdd(&bnd_stats[m.edge_labels[bjj] - 1], 1);
          delete_part(p, d);
          break;
        }
        int model = c.par.model;
        if (model == 0 || model >= 2) {
          // normal2u, utils.pyx:99-119 (x**2 as x*x)
          double ex = v1x - v0x, ey = v1y - v0y;
          double nrm = sqrt(ex * ex + ey * ey);
          double nn[2] = {(v0y - v1y) / nrm, (v1x - v0x) / nrm};
          double nv[2] = {nn[0], nn[1]};
          rough_normal(c, bd, nv);
          reflect_velocity(p.up, nv, damp, d, true);
          if (model == 3) {
            double ns[2] = {nn[0], nn[1]};
            rough_normal(c, bd, ns);
            reflect_velocity(p.us, ns, damp, d, false);
          }
        }
        ++rc;
        if (p.tri == PART_LOC_O && bjj == -1) break;
      }
      if (d == 3 && p.tri != PART_LOC_O && !p.inactive) p.ll = z_localize(c.f, m, p.tri, p.x[0], p.x[1], p.x[2]);
    }
    if (lab >= 1) {
      delete_part(p, d);
      if (c.par.bnd_statistics && bnd_stats && lab - 1 < nopen) atomicAdd(&bnd_stats[lab - 1], 1);
    }
  }
  if (d == 3) {
    for (int pass = 0; pass < 2; ++pass) {
      bool below = pass == 0;
      int want = below ? PART_LOC_B : PART_LOC_A;
      if (!(p.ll == want && p.tri != PART_LOC_O) || p.inactive) continue;
      double zbp, zsp;
      if (c.f.dim == 2) {
        zbp = field2d(c.f, m, c.f.fid_zb, p.tri, 0, p.x[0], p.x[1]);
        zsp = field2d(c.f, m, c.f.fid_zs, p.tri, 0, p.x[0], p.x[1]);
      } else {
        zbp = field2d(c.f, m, c.f.fid_zs, p.tri, 0, p.x[0], p.x[1]);
        zsp = field2d(c.f, m, c.f.fid_zs, p.tri, c.f.nlayers - 1, p.x[0], p.x[1]);
      }
      double hp = fmax(0., zsp - zbp);
      if (below && c.par.bottom_deposition) {
        // deposition on the bed (an extension; the reference reflects): the
        // particle is recorded at (x, y, bed elevation) and leaves the column
        if (c.sr) {
          int q = atomicAdd(c.sr->count, 1);
          if (q < c.sr->cap) {
            c.sr->xyz[3 * q] = p.x[0];
            c.sr->xyz[3 * q + 1] = p.x[1];
            c.sr->xyz[3 * q + 2] = zbp;
            c.sr->key[q] = make_int2(c.sr->step, c.sr->pool | REC_DEPOSIT);
          }
        }
        delete_part(p, d);
        continue;
      }
      if (below) {
        double dz = fabs(zbp - p.x[2]);
        p.x[2] = fmin(zbp + (1. - damp) * dz, zbp + 0.5 * hp);
        p.ll = lower_layer_index(c.f.nlayers, zsp, zbp, p.x[2]);
        if (c.par.model >= 2) p.up[2] = fabs(p.up[2]) * (1. - damp);
        if (c.par.model == 3) p.us[2] = fabs(p.us[2]) * (1. - damp);
      } else {
        double dz = fabs(p.x[2] - zsp);
        p.x[2] = fmax(zsp - (1. - damp) * dz, zsp - 0.5 * hp);
        p.ll = lower_layer_index(c.f.nlayers, zsp, zbp, p.x[2]);
        if (c.par.model >= 2) p.up[2] = -fabs(p.up[2]) * (1. - damp);
        if (c.par.model == 3) p.us[2] = -fabs(p.us[2]) * (1. - damp);
      }
    }
  }
  if (!p.inactive) {
    if (p.tri == PART_LOC_O) delete_part(p, d);
    if (p.ll == PART_LOC_B) delete_part(p, d);
    if (p.ll == PART_LOC_A) delete_part(p, d);
  }
}

}  // namespace cyl
