// Fused 3D steps (sigma-layer prisms) on packed records: LSM-1 with the Euler
// scheme and LSM-2 with the direct integrator without velocity noise, each
// followed by position diffusion with the 3D k-epsilon K and its gradient,
// the xy walk, the z-localization and the 3D boundary conditions (wall
// rebounds, bottom and surface reflections, deletions).
//
// Packed layout: the element records and walk tables of the 2D fused kernel
// (ElemF, int4 node/edge ids, edge records) and one NodeL record (64 B, one
// cache line) per node and sigma plane, plane-major in the Hilbert node order:
// the plane's elevation z, the fluid velocity u, v, w and the nodal k-epsilon
// diffusivity K.  A particle in layer l reads the six records of its element's
// nodes on planes l and l + 1.
//
// The arithmetic is the same, operation for operation, as the generic step:
// update_state_lsm1.pyx (euler_lsm1 :135-177, diffusion_lsm1 :658-726,
// buoyancy :100-108), update_state_lsm2.pyx (di_lsm2 :1329-1543),
// eulerian_field_set.pyx (_interpolate_velocity_3d :1289-1394,
// _interpolate_diffusivity_gradient_c :1644-1763, z_localize :1117-1182),
// zlocalizer.pyx:51-71 and compute_bc.pyx:35-395.
#pragma once
#include "cyl_fast.h"

namespace cyl {

struct __align__(64) NodeL {
  double z, u, v, w, K, pad[3];
};

// tmp -> packed plane records: record l * n + i is mesh node old[i] on plane l
__global__ void __launch_bounds__(BLK) k_pack_nodes3(const double* z, const double* u, const double* v,
                                                     const double* w, const double* kt, const double* ep,
                                                     double fac, int n, int nl, const int* old, NodeL* out) {
  int64_t q = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (q >= (int64_t)n * nl) return;
  int l = (int)(q / n), i = (int)(q - (int64_t)l * n);
  int64_t o = (int64_t)l * n + old[i];
  NodeL r;
  r.z = z[o];
  r.u = u[o];
  r.v = v[o];
  r.w = w ? w[o] : 0.;
  r.K = kt ? fac * kt[o] * kt[o] / fmax(ep[o], EPS) : 0.;
  r.pad[0] = r.pad[1] = r.pad[2] = 0.;
  out[q] = r;
}

struct Fast3Args {
  FastArgs f;              // 2D records, x/y/tri/up/us (components 0, 1), scenarios, streams
  const NodeL* nodes3;     // [nl * nv]
  int nv, nl;
  const double* zA;
  double* zB;
  const int* llA;
  int* llB;
  const double* upA2;
  double* upB2;
  const double* usA2;
  double* usB2;
};

// P1 sums over the element's three records
#define CYL_P1(r0, r1, r2, fld, a0, a1, a2) ((r0).fld * (a0) + (r1).fld * (a1) + (r2).fld * (a2))

// z_localize on packed records (eulerian_field_set.pyx:1117-1182 with zlocalizer.pyx:51-71)
__device__ __forceinline__ int z_localize3(const Fast3Args& a, const ElemF& ef, int4 en, int k, double x, double y,
                                           double z) {
  if (k == -1) return PART_LOC_O;
  double f0 = ef.det[0] + x * ef.gx[0] + y * ef.gy[0];
  double f1 = ef.det[1] + x * ef.gx[1] + y * ef.gy[1];
  double f2 = ef.det[2] + x * ef.gx[2] + y * ef.gy[2];
  const NodeL* top = a.nodes3 + (size_t)(a.nl - 1) * a.nv;
  double zbp = a.nodes3[en.x].z * f0 + a.nodes3[en.y].z * f1 + a.nodes3[en.z].z * f2;
  double zsp = top[en.x].z * f0 + top[en.y].z * f1 + top[en.z].z * f2;
  int ll = lower_layer_index(a.nl, zsp, zbp, z);
  if (z < zbp) ll = PART_LOC_B;
  else if (z > zsp) ll = PART_LOC_A;
  return ll;
}

// bottom and surface elevations at (x, y) in element k
__device__ __forceinline__ void column3(const Fast3Args& a, int k, double x, double y, double& zbp, double& zsp) {
  const ElemF& ef = a.f.fm.ef[k];
  int4 en = a.f.fm.en[k];
  double f0 = ef.det[0] + x * ef.gx[0] + y * ef.gy[0];
  double f1 = ef.det[1] + x * ef.gx[1] + y * ef.gy[1];
  double f2 = ef.det[2] + x * ef.gx[2] + y * ef.gy[2];
  const NodeL* top = a.nodes3 + (size_t)(a.nl - 1) * a.nv;
  zbp = a.nodes3[en.x].z * f0 + a.nodes3[en.y].z * f1 + a.nodes3[en.z].z * f2;
  zsp = top[en.x].z * f0 + top[en.y].z * f1 + top[en.z].z * f2;
}

// PREGEN: the slot's three normals are the generator's record of its active
// rank (uniform draw assignment); otherwise they are sampled here from the
// prepared start (k_draw_start).
template <int MODEL, int RNG, bool PREGEN>
__device__ __forceinline__ unsigned long long fast3d_particle(const Fast3Args& a3, int j, const ZigTables& zt) {
  const FastArgs& a = a3.f;
  if (a.inactive[j] != 0) return 0;
  const int s = a.scen ? a.scen[j] : 0;
  // the record's load is issued here, before the branches, so that its latency
  // overlaps the element and node loads
  const int64_t xidx = PREGEN ? xi_index(a.chunk_base, a.gen_shift, s, a.rank[j]) : 0;
  XiRec3 xr{0., 0., 0., 0.};
  if (PREGEN) xr = reinterpret_cast<const XiRec3*>(a.xi)[(uint64_t)xidx < (uint64_t)a.xi_cap ? xidx : 0];
  int tri0 = a.triA[j], ll0 = a3.llA[j];
  double x0 = a.xA[j], y0 = a.yA[j], z0 = a3.zA[j];
  if (a.scen_end && batch_step_of(a) >= a.scen_end[s]) {
    // the scenario has run all its steps: its rows carry over unchanged
    a.xB[j] = x0;
    a.yB[j] = y0;
    a3.zB[j] = z0;
    a.triB[j] = tri0;
    a3.llB[j] = ll0;
    a.upB0[j] = a.upA0[j];
    a.upB1[j] = a.upA1[j];
    a3.upB2[j] = a3.upA2[j];
    a.usB0[j] = a.usA0[j];
    a.usB1[j] = a.usA1[j];
    a3.usB2[j] = a3.usA2[j];
    return 0;
  }
  const StreamDev& st = a.stream[s];
  const FastScen& q = a.fsc[s];
  int perr = 0;
  if (tri0 < 0) perr = E_NOT_IN_TRI;
  else if (ll0 == PART_LOC_B) perr = E_BELOW_BOTTOM;
  else if (ll0 == PART_LOC_A) perr = E_ABOVE_SURFACE;
  if (perr) {
    atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)perr);
    a.xB[j] = x0;
    a.yB[j] = y0;
    a3.zB[j] = z0;
    a.triB[j] = tri0;
    a3.llB[j] = ll0;
    return 0;
  }
  const ElemF& ef = a.fm.ef[tri0];
  int4 en = a.fm.en[tri0];
  double f0 = ef.det[0] + x0 * ef.gx[0] + y0 * ef.gy[0];
  double f1 = ef.det[1] + x0 * ef.gx[1] + y0 * ef.gy[1];
  double f2 = ef.det[2] + x0 * ef.gx[2] + y0 * ef.gy[2];
  const NodeL* Ll = a3.nodes3 + (size_t)ll0 * a3.nv;
  const NodeL* Lu = Ll + a3.nv;
  // references: each field is loaded where it is used, so the six records
  // never occupy registers whole
  const NodeL &l0 = Ll[en.x], &l1 = Ll[en.y], &l2 = Ll[en.z];
  const NodeL &u0 = Lu[en.x], &u1 = Lu[en.y], &u2 = Lu[en.z];
  // _interpolate_velocity_3d at the start of the step
  double zlp = CYL_P1(l0, l1, l2, z, f0, f1, f2), zup = CYL_P1(u0, u1, u2, z, f0, f1, f2);
  double dzv = zlp - zup;
  if (fabs(dzv) < EPS) perr = E_DEGENERATE_LAYER;
  double zf0 = (z0 - zup) / dzv, zf1 = (zlp - z0) / dzv;
  double ux = zf0 * CYL_P1(l0, l1, l2, u, f0, f1, f2) + zf1 * CYL_P1(u0, u1, u2, u, f0, f1, f2);
  double uy = zf0 * CYL_P1(l0, l1, l2, v, f0, f1, f2) + zf1 * CYL_P1(u0, u1, u2, v, f0, f1, f2);
  double uz = zf0 * CYL_P1(l0, l1, l2, w, f0, f1, f2) + zf1 * CYL_P1(u0, u1, u2, w, f0, f1, f2);
  double x, y, z, upx = ux, upy = uy, upz = uz;
  if (MODEL == 1) {
    // euler_lsm1
    x = x0 + ux * a.dt;
    y = y0 + uy * a.dt;
    z = z0 + uz * a.dt;
  } else {
    // di_lsm2: exact Ornstein-Uhlenbeck drift of the particle velocity, gravity on z
    double v0x = a.upA0[j], v0y = a.upA1[j], v0z = a3.upA2[j];
    double dx = ux - v0x, dy = uy - v0y, dz = uz - v0z;
    double a1 = relax_rate_dp(a.pc, a.diameter[j], sqrt(dx * dx + dy * dy + dz * dz));
    double beta, D1, D2, fxc, um_beta;
    if (!lsm2_drift_coefficients(a1, a.dt, beta, D1, D2, fxc, um_beta)) {
      perr = E_LSM2_COEF;
      beta = D1 = D2 = fxc = um_beta = 0.;
    }
    double Fx = a.pc.a2c * 0., Fy = a.pc.a2c * 0.;  // no added mass: zero fluid acceleration
    double Fz = a.pc.a2c * 0. - GRAV * a.pc.a0c;
    upx = v0x * beta + ux * um_beta + Fx * D1;
    upy = v0y * beta + uy * um_beta + Fy * D1;
    upz = v0z * beta + uz * um_beta + Fz * D1;
    x = x0 + v0x * D1 + ux * D2 + Fx * fxc;
    y = y0 + v0y * D1 + uy * D2 + Fy * fxc;
    z = z0 + v0z * D1 + uz * D2 + Fz * fxc;
  }
  Draws dr;
  dr.st = &st;
  dr.jt = &a.jt[s];
  dr.jt8 = nullptr;
  dr.pre_pos = PREGEN ? nullptr : a.pos0;
  dr.pre_state = a.state0;
  dr.slot = j;
  dr.zt = &zt;
  dr.method = RNG;
  dr.dry = false;
  dr.nunits = 0;
  dr.ncalls = 0;
  dr.pos0 = 0;
  dr.live = false;
  dr.err = 0;
  dr.unit0 = 0;
  // position diffusion with K and grad K at the start-of-step element and layer
  if (q.diffusion_model != 0 && !perr) {
    double diffx, diffz, driftx = 0., drifty = 0., driftz = 0.;
    if (q.diffusion_model >= 2) {
      double Kl = CYL_P1(l0, l1, l2, K, f0, f1, f2);
      double gxl = CYL_P1(l0, l1, l2, K, ef.gx[0], ef.gx[1], ef.gx[2]);
      double gyl = CYL_P1(l0, l1, l2, K, ef.gy[0], ef.gy[1], ef.gy[2]);
      double zlx = CYL_P1(l0, l1, l2, z, ef.gx[0], ef.gx[1], ef.gx[2]);
      double zly = CYL_P1(l0, l1, l2, z, ef.gy[0], ef.gy[1], ef.gy[2]);
      double Ku = CYL_P1(u0, u1, u2, K, f0, f1, f2);
      double gxu = CYL_P1(u0, u1, u2, K, ef.gx[0], ef.gx[1], ef.gx[2]);
      double gyu = CYL_P1(u0, u1, u2, K, ef.gy[0], ef.gy[1], ef.gy[2]);
      double zux = CYL_P1(u0, u1, u2, z, ef.gx[0], ef.gx[1], ef.gx[2]);
      double zuy = CYL_P1(u0, u1, u2, z, ef.gy[0], ef.gy[1], ef.gy[2]);
      double dzk = zup - zlp;
      if (fabs(dzk) < EPS) perr = E_DEGENERATE_LAYER;
      double sk = (z0 - zlp) / dzk;
      double K = (1. - sk) * Kl + sk * Ku;
      double dkdz = (Ku - Kl) / dzk;
      double dkdx = (1. - sk) * gxl + sk * gxu - dkdz * ((1. - sk) * zlx + sk * zux);
      double dkdy = (1. - sk) * gyl + sk * gyu - dkdz * ((1. - sk) * zly + sk * zuy);
      diffx = sqrt(2. * fmax(0., K)) * a.sqrtdt;
      driftx = dkdx * a.dt;
      drifty = dkdy * a.dt;
      if (q.diffusion_model == 2) {
        diffz = q.diff_const_z;
      } else {
        diffz = diffx;
        driftz = dkdz * a.dt;
      }
    } else {
      diffx = q.diff_const;
      diffz = q.diff_const_z;
    }
    double xix = 0., xiy = 0., xiz = 0.;
    if (PREGEN) {
      // the generator's values stand for the first draws of the rank; a row
      // that skips the pair (zero amplitude) draws fewer units than predicted
      // and the step is rerun in general mode
      const bool have = st.units_per_part > 0 && xidx < a.xi_cap;
      if (diffx != 0.) {
        xix = have ? xr.a : 0.;
        xiy = have ? xr.b : 0.;
        dr.nunits += RNG == 3 ? 1 : 2;
      }
      if (diffz != 0.) {
        xiz = have ? xr.c : 0.;
        dr.nunits += 1;
      }
    } else {
      if (diffx != 0.) dr.pair(xix, xiy);
      if (diffz != 0.) xiz = dr.single();
    }
    x += driftx + diffx * xix;
    y += drifty + diffx * xiy;
    z += driftz + diffz * xiz;
  }
  if (MODEL == 1) {
    double wb = buoyancy_velocity(a.pc, a.diameter[j]);
    z += wb * a.dt;
    upz += wb;
  }
  // localization from (tri0, x0) along the path, then the layer
  int bnd_j = -1, tri = tri0;
  if (!in_elem(ef, x, y)) {
    tri = walk_fast(a.fm, tri0, ef, x0, y0, x, y, bnd_j);
    if (tri == -1 && bnd_j == -1) tri = locate_grid(a.m, x, y);
  }
  int ll = tri >= 0 ? z_localize3(a3, a.fm.ef[tri], a.fm.en[tri], tri, x, y, z) : PART_LOC_O;
  bool deleted = false;
  const double damp = q.damp;
  if (q.boundary_conditions) {
    if (bnd_j != -1) {
      int lab = a.fm.edge_labels[bnd_j];
      int strand_cond = 0;
      if (lab == 0 && q.strand > 0.) {
        BcDraws bd;
        bd.method = RNG;
        bd.zt = &zt;
        bd.g.inc = 0x6c62272e07bb0143ULL | 1u;
        int lpool = a.pool[j] - (a.scen_base ? a.scen_base[s] : 0);  // row of the scenario's own pool
        bd.g.state = st.base ^ ((uint64_t)(unsigned)lpool * 0x9e3779b97f4a7c15ULL);
        strand_cond = bd.uniform() < q.strand;
      }
      if (lab == 0 && !strand_cond) {
        int bjj = bnd_j, rc = 0;
        while (tri == PART_LOC_O && rc < q.max_wall_rebounds) {
          int4 ed = a.fm.ed[bjj];
          double2 v0 = a.fm.xy[ed.x], v1 = a.fm.xy[ed.y];
          // compute_symetric_point2d, intersections.pyx:27-62
          double nx = v0.y - v1.y, ny = v1.x - v0.x;
          double norm2 = nx * nx + ny * ny;
          double dist = (x - v0.x) * nx + (y - v0.y) * ny;
          double coef = 1. - damp / 2.;
          double px = x - 2. * ((dist * coef) / norm2) * nx;
          double py = y - 2. * ((dist * coef) / norm2) * ny;
          x = px;
          y = py;
          tri = localize_fast(a.fm, a.m, tri0, x0, y0, px, py, bjj);
          if (tri == PART_LOC_O && bjj != -1 && a.fm.edge_labels[bjj] >= 1) {
            int l2 = a.fm.edge_labels[bjj];
            if (q.bnd_statistics && a.bnd_stats && l2 - 1 < a.nopen) atomicAdd(&a.bnd_stats[s * a.nopen_stride + l2 - 1], 1);
            deleted = true;
            break;
          }
          if (MODEL == 2) {
            // reflection of the particle velocity on the wall (normal2u, utils.pyx:99-119), damped
            double ex = v1.x - v0.x, ey = v1.y - v0.y;
            double nrm = sqrt(ex * ex + ey * ey);
            double nn0 = (v0.y - v1.y) / nrm, nn1 = (v1.x - v0.x) / nrm;
            double w0 = upx;
            upx = w0 * nn0 + upy * nn1;
            upy = -w0 * nn1 + upy * nn0;
            upx = -1. * upx;
            w0 = upx;
            upx = w0 * nn0 - upy * nn1;
            upy = w0 * nn1 + upy * nn0;
            upx *= 1. - damp;
            upy *= 1. - damp;
            upz *= 1. - damp;
          }
          ++rc;
          if (tri == PART_LOC_O && bjj == -1) break;
        }
        if (tri != PART_LOC_O && !deleted) ll = z_localize3(a3, a.fm.ef[tri], a.fm.en[tri], tri, x, y, z);
      }
      if (lab >= 1 && !deleted) {
        deleted = true;
        if (q.bnd_statistics && a.bnd_stats && lab - 1 < a.nopen) atomicAdd(&a.bnd_stats[s * a.nopen_stride + lab - 1], 1);
      }
    }
    // bottom, then surface: damped reflection into the column
    for (int pass = 0; pass < 2 && !deleted; ++pass) {
      bool below = pass == 0;
      if (ll != (below ? PART_LOC_B : PART_LOC_A) || tri == PART_LOC_O) continue;
      double zbp, zsp;
      column3(a3, tri, x, y, zbp, zsp);
      double hp = fmax(0., zsp - zbp);
      if (below && q.deposition) {
        // deposition on the bed (an extension; the reference reflects): the
        // particle is recorded at (x, y, bed elevation) and leaves the column
        if (a.rec_xyz) {
          int r = atomicAdd(a.rec_count, 1);
          if (r < a.rec_cap) {
            a.rec_xyz[3 * r] = x;
            a.rec_xyz[3 * r + 1] = y;
            a.rec_xyz[3 * r + 2] = zbp;
            a.rec_key[r] = make_int2(step_of(a), a.pool[j] | REC_DEPOSIT);
          }
        }
        deleted = true;
        continue;
      }
      if (below) {
        double dz = fabs(zbp - z);
        z = fmin(zbp + (1. - damp) * dz, zbp + 0.5 * hp);
        ll = lower_layer_index(a3.nl, zsp, zbp, z);
        if (MODEL == 2) upz = fabs(upz) * (1. - damp);
      } else {
        double dz = fabs(z - zsp);
        z = fmax(zsp - (1. - damp) * dz, zsp - 0.5 * hp);
        ll = lower_layer_index(a3.nl, zsp, zbp, z);
        if (MODEL == 2) upz = -fabs(upz) * (1. - damp);
      }
    }
    if (!deleted && (tri == PART_LOC_O || ll == PART_LOC_B || ll == PART_LOC_A)) deleted = true;
  } else if (tri == PART_LOC_O || ll == PART_LOC_B || ll == PART_LOC_A) {
    deleted = true;
  }
  if (perr) atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)perr);
  if (dr.err) atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)dr.err);
  if (!a.general && dr.nunits != st.units_per_part) atomicOr(&a.flag[0], 1);
  if (deleted) {
    a.inactive[j] = 2;
    a.del_step[j] = step_of(a);
    atomicAdd(a.n_deleted, 1);
    x = y = z = 0.;
    tri = -1;
    ll = 0;
    ux = uy = uz = 0.;
    upx = upy = upz = 0.;
  }
  a.xB[j] = x;
  a.yB[j] = y;
  a3.zB[j] = z;
  a.triB[j] = tri;
  a3.llB[j] = ll;
  a.upB0[j] = upx;
  a.upB1[j] = upy;
  a3.upB2[j] = upz;
  a.usB0[j] = ux;
  a.usB1[j] = uy;
  a3.usB2[j] = uz;
  return dr.live ? (unsigned long long)(dr.pos0 + dr.ncalls) : 0ull;
}

#undef CYL_P1

template <int MODEL, int RNG, bool PREGEN>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(PREGEN && MODEL == 1 ? 4 : 3))) k_fast_3d(Fast3Args a3) {
  const FastArgs& a = a3.f;
  if (a.guarded && a.flag[0] == 0) return;
  __shared__ ZigTables zt;
  if (RNG == 2 && !PREGEN) {
    for (int q = threadIdx.x; q < 128; q += BLK) {
      zt.kn[q] = a.zt->kn[q];
      zt.wn[q] = a.zt->wn[q];
      zt.fn[q] = a.zt->fn[q];
    }
  }
  __syncthreads();
  const int64_t chunk = xcd_chunk(a.n);
  if (chunk < 0) return;
  const int64_t jl = chunk * BLK + threadIdx.x;
  if (jl < a.n) (void)fast3d_particle<MODEL, RNG, PREGEN>(a3, a.j0 + (int)jl, zt);
}

}  // namespace cyl
