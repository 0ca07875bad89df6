// Field interpolation at a particle: P1 barycentric factors in a triangle,
// prism (sigma-layer) interpolation in 3D, k-epsilon diffusivity and its
// gradient, vertical localization, pseudo-3D log-law corrections and the
// fluid acceleration.  Each routine restates the expression order of
// cylag/core/eulerian_field_set.pyx so results are bit-identical.
#pragma once
#include "cyl_crmath.h"
#include "cyl_types.h"

namespace cyl {

struct Bary {
  int i0, i1, i2;
  double f0, f1, f2;
};

__device__ __forceinline__ Bary bary(const MeshDev& m, int k, double x, double y) {
  // eulerian_field_set.pyx:1044-1046 ; factors f_j = det + x*gradx + y*grady
  Bary b;
  const double* d = m.det + 3 * k;
  const double* gx = m.gx + 3 * k;
  const double* gy = m.gy + 3 * k;
  b.f0 = d[0] + x * gx[0] + y * gy[0];
  b.f1 = d[1] + x * gx[1] + y * gy[1];
  b.f2 = d[2] + x * gx[2] + y * gy[2];
  b.i0 = m.tri[3 * k];
  b.i1 = m.tri[3 * k + 1];
  b.i2 = m.tri[3 * k + 2];
  return b;
}

__device__ __forceinline__ double p1(const double* f, const Bary& b, int off = 0) {
  return f[off + b.i0] * b.f0 + f[off + b.i1] * b.f1 + f[off + b.i2] * b.f2;
}

__device__ __forceinline__ double p1g(const double* f, const Bary& b, const double* g, int off = 0) {
  return f[off + b.i0] * g[0] + f[off + b.i1] * g[1] + f[off + b.i2] * g[2];
}

// _get_tmp(field, layer) : layer slice in 3D, the whole array in 2D
__device__ __forceinline__ const double* layer_of(const FieldsDev& f, int fid, int l) {
  const double* t = f.tmp[fid];
  return (f.dim == 2 || l == -1) ? t : t + (size_t)l * f.npoints;
}

// zlocalizer.pyx:51-71 (Cython lowers double // to floor(a / b))
__device__ __forceinline__ int lower_layer_index(int nl, double zs, double zb, double z) {
  if (nl == 1) return 0;
  double hl = (zs - zb) / (nl - 1);
  if (!(hl > 0.)) return 0;
  double fi = floor((z - zb) / hl);
  if (!(fi >= 0.)) return 0;
  if (!(fi <= (double)(nl - 2))) return nl - 2;
  return (int)fi;
}

// _interpolate_field_2d, eulerian_field_set.pyx:994-1055
__device__ __forceinline__ double field2d(const FieldsDev& f, const MeshDev& m, int fid, int k, int l,
                                          double x, double y) {
  Bary b = bary(m, k, x, y);
  return p1(layer_of(f, fid, f.dim == 2 ? -1 : l), b);
}

// _interpolate_velocity_2d, eulerian_field_set.pyx:1057-1115
__device__ __forceinline__ void vel2d(const FieldsDev& f, const MeshDev& m, int k, double x, double y,
                                      double& ux, double& uy) {
  Bary b = bary(m, k, x, y);
  ux = p1(f.tmp[f.fid_ux], b);
  uy = p1(f.tmp[f.fid_uy], b);
}

// z_localize, eulerian_field_set.pyx:1117-1182
__device__ inline int z_localize(const FieldsDev& f, const MeshDev& m, int k, double x, double y, double z) {
  if (k == -1) return PART_LOC_O;
  Bary b = bary(m, k, x, y);
  const double *zb, *zs;
  if (f.dim == 2) {
    zb = layer_of(f, f.fid_zb, -1);
    zs = layer_of(f, f.fid_zs, -1);
  } else {
    zb = layer_of(f, f.fid_zs, 0);
    zs = layer_of(f, f.fid_zs, f.nlayers - 1);
  }
  double zbp = p1(zb, b), zsp = p1(zs, b);
  int ll = f.dim == 2 ? 0 : lower_layer_index(f.nlayers, zsp, zbp, z);
  if (z < zbp) ll = PART_LOC_B;
  else if (z > zsp) ll = PART_LOC_A;
  return ll;
}

// _interpolate_field_3d (:1184-1287) with zlocalize = 0
__device__ inline double field3d(const FieldsDev& f, const MeshDev& m, int fid, int k, int ll, double x,
                                 double y, double z, int& err) {
  if (ll == PART_LOC_B) { err = E_BELOW_BOTTOM; return 0.; }
  if (ll == PART_LOC_A) { err = E_ABOVE_SURFACE; return 0.; }
  Bary b = bary(m, k, x, y);
  const double* el = f.tmp[f.fid_zs];
  const double* fd = f.tmp[fid];
  int jl = ll * f.npoints, ju = (ll + 1) * f.npoints;
  double zlp = p1(el, b, jl), zup = p1(el, b, ju);
  double dz = zlp - zup;
  if (fabs(dz) < EPS) { err = E_DEGENERATE_LAYER; return 0.; }
  double zf0 = (z - zup) / dz, zf1 = (zlp - z) / dz;
  double resl = p1(fd, b, jl), resu = p1(fd, b, ju);
  return zf0 * resl + zf1 * resu;
}

// _interpolate_velocity_3d (:1289-1394) with zlocalize = 0
__device__ inline void vel3d(const FieldsDev& f, const MeshDev& m, int k, int ll, double x, double y, double z,
                             double& ux, double& uy, double& uz, int& err) {
  if (ll == PART_LOC_B) { err = E_BELOW_BOTTOM; return; }
  if (ll == PART_LOC_A) { err = E_ABOVE_SURFACE; return; }
  Bary b = bary(m, k, x, y);
  const double* el = f.tmp[f.fid_zs];
  int jl = ll * f.npoints, ju = (ll + 1) * f.npoints;
  double zlp = p1(el, b, jl), zup = p1(el, b, ju);
  double dz = zlp - zup;
  if (fabs(dz) < EPS) { err = E_DEGENERATE_LAYER; return; }
  double zf0 = (z - zup) / dz, zf1 = (zlp - z) / dz;
  const double *u = f.tmp[f.fid_ux], *v = f.tmp[f.fid_uy], *w = f.tmp[f.fid_uz];
  ux = zf0 * p1(u, b, jl) + zf1 * p1(u, b, ju);
  uy = zf0 * p1(v, b, jl) + zf1 * p1(v, b, ju);
  uz = zf0 * p1(w, b, jl) + zf1 * p1(w, b, ju);
}

__device__ __forceinline__ double keps_K(double k, double eps, double fac) {
  return fac * k * k / fmax(eps, EPS);  // eulerian_field_set.pyx:1995-1997
}

// _interpolate_diffusivity_gradient_c, eulerian_field_set.pyx:1644-1763
__device__ inline void diffusivity_gradient(const FieldsDev& f, const MeshDev& m, int k, int ll, double x,
                                            double y, double z, double sigc, double& Kp, double& dkdx,
                                            double& dkdy, double& dkdz, int& err) {
  double fac = CMU / sigc;
  Bary b = bary(m, k, x, y);
  const double* gx = m.gx + 3 * k;
  const double* gy = m.gy + 3 * k;
  const double *kt = f.tmp[f.fid_turbeng], *ep = f.tmp[f.fid_dissip];
  if (f.dim == 2) {
    double kn0 = keps_K(kt[b.i0], ep[b.i0], fac);
    double kn1 = keps_K(kt[b.i1], ep[b.i1], fac);
    double kn2 = keps_K(kt[b.i2], ep[b.i2], fac);
    double K = kn0 * b.f0 + kn1 * b.f1 + kn2 * b.f2;
    dkdx = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
    dkdy = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
    dkdz = 0.;
    Kp = fmax(0., K);
    return;
  }
  if (ll == PART_LOC_B) { err = E_BELOW_BOTTOM; return; }
  if (ll == PART_LOC_A) { err = E_ABOVE_SURFACE; return; }
  int np = f.npoints, jl = ll * np, ju = (ll + 1) * np;
  const double* el = f.tmp[f.fid_zs];
  double kn0 = keps_K(kt[jl + b.i0], ep[jl + b.i0], fac);
  double kn1 = keps_K(kt[jl + b.i1], ep[jl + b.i1], fac);
  double kn2 = keps_K(kt[jl + b.i2], ep[jl + b.i2], fac);
  double Kl = kn0 * b.f0 + kn1 * b.f1 + kn2 * b.f2;
  double gxl = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
  double gyl = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
  double zlp = p1(el, b, jl), zlx = p1g(el, b, gx, jl), zly = p1g(el, b, gy, jl);
  kn0 = keps_K(kt[ju + b.i0], ep[ju + b.i0], fac);
  kn1 = keps_K(kt[ju + b.i1], ep[ju + b.i1], fac);
  kn2 = keps_K(kt[ju + b.i2], ep[ju + b.i2], fac);
  double Ku = kn0 * b.f0 + kn1 * b.f1 + kn2 * b.f2;
  double gxu = kn0 * gx[0] + kn1 * gx[1] + kn2 * gx[2];
  double gyu = kn0 * gy[0] + kn1 * gy[1] + kn2 * gy[2];
  double zup = p1(el, b, ju), zux = p1g(el, b, gx, ju), zuy = p1g(el, b, gy, ju);
  double dz = zup - zlp;
  if (fabs(dz) < EPS) { err = E_DEGENERATE_LAYER; return; }
  double s = (z - zlp) / dz;
  double K = (1. - s) * Kl + s * Ku;
  dkdz = (Ku - Kl) / dz;
  dkdx = (1. - s) * gxl + s * gxu - dkdz * ((1. - s) * zlx + s * zux);
  dkdy = (1. - s) * gyl + s * gyu - dkdz * ((1. - s) * zly + s * zuy);
  Kp = fmax(0., K);
}

// friction_calc, eulerian_field_set.pyx:1957-1993
__device__ inline double friction_calc(const FieldsDev& f, double h) {
  double c = f.friction_coefficient;
  switch (f.friction_model) {
    case 0: return 2. * GRAV / (c * c);
    case 1: return 2. * GRAV / (c * c) / cr_pow(fmax(EPS, h), 1. / 3.);
    case 2: return 2. * GRAV * (c * c) / cr_pow(fmax(EPS, h), 1. / 3.);
    case 3: {
      double aux = h * 11.036 / c;
      aux = aux > 1.001 ? aux : 1.001;
      double q = cr_log(aux) / KAPPA;
      return 2. / (q * q);
    }
  }
  return __longlong_as_double(0x7ff8000000000000LL);
}

// pseudo_3d_velocity_corrections, eulerian_field_set.pyx:1887-1955 (with
// _interpolate_surface_velocity, :1581-1642)
__device__ inline void pseudo3d(const FieldsDev& f, const MeshDev& m, int k, double x, double y, double z,
                                double dt, double ux, double uy, double& w, double& hor) {
  const double *zs_n = f.tmp[f.fid_zs], *zb_n = f.tmp[f.fid_zb];
  Bary b = bary(m, k, x, y);
  const double* gx = m.gx + 3 * k;
  const double* gy = m.gy + 3 * k;
  double zs = p1(zs_n, b), zb = p1(zb_n, b);
  double h = fmax(EPS, zs - zb);
  double sigma = fmin(1., fmax(0., (z - zb) / h));
  double Cf = friction_calc(f, h);
  double z0 = fmax(EPS, h * cr_exp(-(1. + KAPPA / sqrt(Cf / 2.))));
  hor = (z - zb <= z0) ? 0.0 : cr_log((z - zb) / z0) / (cr_log(h / z0) - 1.);
  double uc = ux * hor, vc = uy * hor;
  double dzsdx = p1g(zs_n, b, gx), dzsdy = p1g(zs_n, b, gy);
  double dzbdx = p1g(zb_n, b, gx), dzbdy = p1g(zb_n, b, gy);
  double dzsdt = 0.;
  if (!f.frozen && f.prev_valid) dzsdt = (zs - p1(f.prev[0], b)) / dt;
  w = (1. - sigma) * (uc * dzbdx + vc * dzbdy) + sigma * (dzsdt + uc * dzsdx + vc * dzsdy);
}

// _interpolate_acceleration_2d, eulerian_field_set.pyx:1396-1464
__device__ inline void accel2d(const FieldsDev& f, const MeshDev& m, int k, double x, double y, double dt,
                               double ua, double va, double& ax, double& ay) {
  Bary b = bary(m, k, x, y);
  const double* gx = m.gx + 3 * k;
  const double* gy = m.gy + 3 * k;
  const double *u = f.tmp[1], *v = f.tmp[2], *u0 = f.prev[1], *v0 = f.prev[2];
  double dudx = p1g(u, b, gx), dudy = p1g(u, b, gy), dvdx = p1g(v, b, gx), dvdy = p1g(v, b, gy);
  double dudt = 0., dvdt = 0.;
  if (!f.frozen && f.prev_valid) {
    dudt = ((u[b.i0] - u0[b.i0]) * b.f0 + (u[b.i1] - u0[b.i1]) * b.f1 + (u[b.i2] - u0[b.i2]) * b.f2) / dt;
    dvdt = ((v[b.i0] - v0[b.i0]) * b.f0 + (v[b.i1] - v0[b.i1]) * b.f1 + (v[b.i2] - v0[b.i2]) * b.f2) / dt;
  }
  ax = dudt + ua * dudx + va * dudy;
  ay = dvdt + ua * dvdx + va * dvdy;
}

struct PrismVG {
  double v, gx, gy, gz;
};

// _prism_value_gradient, eulerian_field_set.pyx:1999-2023
__device__ __forceinline__ PrismVG prism_vg(const double* fld, int jl, int ju, const Bary& b,
                                            const double* gx, const double* gy, double s, double dz,
                                            double zlx, double zly, double zux, double zuy) {
  double fl = p1(fld, b, jl), fu = p1(fld, b, ju);
  double gxl = p1g(fld, b, gx, jl), gyl = p1g(fld, b, gy, jl);
  double gxu = p1g(fld, b, gx, ju), gyu = p1g(fld, b, gy, ju);
  double dfdz = (fu - fl) / dz;
  PrismVG r;
  r.v = (1. - s) * fl + s * fu;
  r.gx = (1. - s) * gxl + s * gxu - dfdz * ((1. - s) * zlx + s * zux);
  r.gy = (1. - s) * gyl + s * gyu - dfdz * ((1. - s) * zly + s * zuy);
  r.gz = dfdz;
  return r;
}

// _interpolate_acceleration_3d, eulerian_field_set.pyx:1466-1579
__device__ inline void accel3d(const FieldsDev& f, const MeshDev& m, int k, int ll, double x, double y,
                               double z, double dt, double ua, double va, double wa, double& ax, double& ay,
                               double& az, int& err) {
  Bary b = bary(m, k, x, y);
  const double* gx = m.gx + 3 * k;
  const double* gy = m.gy + 3 * k;
  const double* el = f.tmp[f.fid_zs];
  int jl = ll * f.npoints, ju = (ll + 1) * f.npoints;
  double zlp = p1(el, b, jl), zup = p1(el, b, ju);
  double dz = zup - zlp;
  if (fabs(dz) < EPS) { err = E_DEGENERATE_LAYER; return; }
  double s = (z - zlp) / dz;
  double zlx = p1g(el, b, gx, jl), zly = p1g(el, b, gy, jl), zux = p1g(el, b, gx, ju), zuy = p1g(el, b, gy, ju);
  PrismVG U = prism_vg(f.tmp[1], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy);
  PrismVG V = prism_vg(f.tmp[2], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy);
  PrismVG W = prism_vg(f.tmp[3], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy);
  double dudt = 0., dvdt = 0., dwdt = 0.;
  if (!f.frozen && f.prev_valid) {
    dudt = (U.v - prism_vg(f.prev[1], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy).v) / dt;
    dvdt = (V.v - prism_vg(f.prev[2], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy).v) / dt;
    dwdt = (W.v - prism_vg(f.prev[3], jl, ju, b, gx, gy, s, dz, zlx, zly, zux, zuy).v) / dt;
  }
  ax = dudt + ua * U.gx + va * U.gy + wa * U.gz;
  ay = dvdt + ua * V.gx + va * V.gy + wa * V.gz;
  az = dwdt + ua * W.gx + va * W.gy + wa * W.gz;
}

}  // namespace cyl
