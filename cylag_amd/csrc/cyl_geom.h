// Element localization on the device: point-in-triangle predicate, the
// segment walk across edges (cylag/geom/xylocalizer.pyx:338-435) and the
// fallback search that replaces the reference's centroid KD-tree
// (xylocalizer.pyx:200-286) with a bucket grid over triangle bounding boxes.
//
// All predicates use the reference's exact-zero floating-point tests in the
// same operation order (cylag/geom/utils.pyx:228-385); the library is built
// with -ffp-contract=off so no product is fused into an FMA.
#pragma once
#include "cyl_types.h"

namespace cyl {

__device__ __forceinline__ int wrap_tri(const MeshDev& m, int k) { return k < 0 ? k + m.nt : k; }

__device__ __forceinline__ double pintri(double x, double y, double xa, double ya, double xb,
                                         double yb, double xc, double yc) {
  // utils.pyx:356-357
  return ((xb - xa) * (y - ya) - (x - xa) * (yb - ya)) * ((x - xa) * (yc - ya) - (xc - xa) * (y - ya));
}

// point_in_tri (xylocalizer.pyx:126-167): returns k (the argument, possibly -1
// which indexes the last triangle like Python) when inside, else -1.
__device__ __forceinline__ int point_in_tri(const MeshDev& m, int k, double x, double y) {
  int kk = wrap_tri(m, k);
  int v0 = m.tri[3 * kk], v1 = m.tri[3 * kk + 1], v2 = m.tri[3 * kk + 2];
  double x0 = m.x[v0], y0 = m.y[v0], x1 = m.x[v1], y1 = m.y[v1], x2 = m.x[v2], y2 = m.y[v2];
  bool in = pintri(x, y, x0, y0, x1, y1, x2, y2) >= 0 && pintri(x, y, x1, y1, x2, y2, x0, y0) >= 0 &&
            pintri(x, y, x2, y2, x0, y0, x1, y1) >= 0;
  return in ? k : -1;
}

__device__ __forceinline__ int orient(double v0x, double v0y, double v1x, double v1y, double v2x,
                                      double v2y) {
  // utils.pyx:246-252
  double d = (v0y - v1y) * (v2x - v0x) - (v0x - v1x) * (v2y - v0y);
  return d == 0. ? 0 : (d > 0. ? 1 : 2);
}

__device__ __forceinline__ bool on_seg(double v0x, double v0y, double v1x, double v1y, double v2x,
                                       double v2y) {
  // utils.pyx:300-303
  return v2x <= fmax(v1x, v0x) && v2x >= fmin(v1x, v0x) && v2y <= fmax(v1y, v0y) && v2y >= fmin(v1y, v0y);
}

__device__ __forceinline__ bool seg_intersect(double p0x, double p0y, double p1x, double p1y,
                                              double s0x, double s0y, double s1x, double s1y) {
  // intersections.pyx:106-148
  int o1 = orient(p0x, p0y, p1x, p1y, s0x, s0y);
  int o2 = orient(p0x, p0y, p1x, p1y, s1x, s1y);
  int o3 = orient(s0x, s0y, s1x, s1y, p0x, p0y);
  int o4 = orient(s0x, s0y, s1x, s1y, p1x, p1y);
  if (o1 != o2 && o3 != o4) return true;
  if (o1 == 0 && on_seg(p0x, p0y, p1x, p1y, s0x, s0y)) return true;
  if (o2 == 0 && on_seg(p0x, p0y, p1x, p1y, s1x, s1y)) return true;
  if (o3 == 0 && on_seg(s0x, s0y, s1x, s1y, p0x, p0y)) return true;
  if (o4 == 0 && on_seg(s0x, s0y, s1x, s1y, p1x, p1y)) return true;
  return false;
}

// Path walk from element k along [p0, p1] (xylocalizer.pyx:338-435).
// Returns the element, or -1 with bnd_j = crossed boundary edge, or (-1, -1).
__device__ inline int walk_path(const MeshDev& m, int k, double p0x, double p0y, double x, double y,
                                int& bnd_j) {
  int new_k = -1, niter = 0, last_edge = -1;
  bnd_j = -1;
  while (new_k == -1 && niter <= 50) {
    bool hit = false;
    int kk = wrap_tri(m, k);
    for (int i = 0; i < 3; ++i) {
      int j = m.triedge[3 * kk + i];
      if (j == last_edge) continue;
      int i0 = m.edges[2 * j], i1 = m.edges[2 * j + 1];
      if (seg_intersect(p0x, p0y, x, y, m.x[i0], m.y[i0], m.x[i1], m.y[i1])) {
        hit = true;
        int k0 = m.edgetri[2 * j], k1 = m.edgetri[2 * j + 1];
        int kn = (k0 == k) ? k1 : k0;
        if (kn == -1) {
          bnd_j = j;
          return -1;
        }
        new_k = point_in_tri(m, kn, x, y);
        if (new_k != -1) return new_k;
        k = kn;
        last_edge = j;
        break;
      }
    }
    if (!hit) return point_in_tri(m, k, x, y) == -1 ? -1 : k;
    ++niter;
  }
  return new_k;
}

// Fallback search: among the triangles that contain (x, y) the one whose
// centroid is nearest, ties to the lower index.  This is the element the
// reference's KD-tree scan (nearest 10, 20, 100, 1000 centroids in distance
// order) returns whenever that element's centroid is among the 1000 nearest.
__device__ inline int locate_grid(const MeshDev& m, double x, double y) {
  if (!(x == x) || !(y == y)) return -1;
  double fx = floor((x - m.g_x0) * m.g_inv), fy = floor((y - m.g_y0) * m.g_inv);
  if (fx < 0. || fy < 0. || fx >= (double)m.g_nx || fy >= (double)m.g_ny) {
    // outside the grid: only triangles touching the border cells can contain
    // the point, and a bounding-box test on the border would have caught it
    fx = fx < 0. ? 0. : (fx >= m.g_nx ? m.g_nx - 1 : fx);
    fy = fy < 0. ? 0. : (fy >= m.g_ny ? m.g_ny - 1 : fy);
  }
  int c = (int)fy * m.g_nx + (int)fx;
  int best = -1;
  double dbest = 0.;
  for (int q = m.g_start[c]; q < m.g_start[c + 1]; ++q) {
    int k = m.g_items[q];
    if (point_in_tri(m, k, x, y) == -1) continue;
    double dx = x - m.centroid[2 * k], dy = y - m.centroid[2 * k + 1];
    double d = dx * dx + dy * dy;
    if (best == -1 || d < dbest || (d == dbest && k < best)) {
      best = k;
      dbest = d;
    }
  }
  return best;
}

// xy_localize (xylocalizer.pyx:27-124) for methods 0, 2 and 3.
__device__ inline int xy_localize(const MeshDev& m, int k, double p0x, double p0y, double x, double y,
                                  int method, int& bnd_j, int& err) {
  bnd_j = -1;
  int nk = point_in_tri(m, k, x, y);
  if (nk != -1) return nk;
  if (method == 2) {
    if (k == -1 && x == p0x && y == p0y) method = 0;
    if (k != -1 && x == p0x && y == p0y) {
      err = E_METHOD1;
      return -1;
    }
  }
  if (method == 0) return locate_grid(m, x, y);
  nk = walk_path(m, k, p0x, p0y, x, y, bnd_j);
  if (method == 3 && nk == -1 && bnd_j == -1) nk = locate_grid(m, x, y);
  return nk;
}

}  // namespace cyl
