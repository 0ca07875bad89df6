// Fused 2D steps on packed, cache-line-sized mesh records: LSM-1 (Euler
// advection + Euler-Maruyama diffusion + walk + rebounds) and LSM-2 with the
// direct integrator and position diffusion (the velocity relaxes exactly
// towards the fluid velocity; rebounds reflect it).  These are the common
// production configurations; every other configuration runs k_step.
//
// Packed layout in HBM:
//   ElemF  (128 B, one L2 line): P1 coefficients det/gradx/grady and the three
//          vertex coordinates of an element
//   int4   element node ids, element edge ids (triedge order), edge records
//          {vertex0, vertex1, element0, element1}
//   NodeF  (32 B): time-interpolated u, v and nodal k-epsilon diffusivity of a
//          node, rewritten once per step by k_pack_nodes
// Packed node records and node ids use the nodes' Hilbert-curve order, so the
// nodes of neighbouring elements share cache lines; element and edge ids are
// the mesh's own.
// The arithmetic is the same, operation for operation, as update_state_lsm1.pyx
// (euler_lsm1 :135-177, diffusion_lsm1 :658-726), update_state_lsm2.pyx
// (di_lsm2 :1329-1543, lsm2_direct_coefficients :1240-1327), xylocalizer.pyx:338-435
// and compute_bc.pyx:35-395.
#pragma once
#include "cyl_events.h"

namespace cyl {

struct __align__(128) ElemF {
  double det[3], gx[3], gy[3];
  double vx[3], vy[3];
  double pad;
};

struct __align__(32) NodeF {
  double u, v, K, pad;
};

// Walk record of an element (32 B): its edges in increasing id order (the
// reference's triedge order), for each edge the element-local indices of the
// edge's two vertices in the edge's own (v0, v1) order, and the element across
// the edge (-1 on the boundary).  The walk then tests an edge with the vertex
// coordinates already loaded from the current element's ElemF, in exactly the
// operand order of the edge table.
struct __align__(32) WalkRec {
  int e[3];
  int vv;     // bits 4i..4i+1: local index of edge i's v0; 4i+2..4i+3: of its v1
  int nb[3];
  int pad;
};

struct FastMesh {
  const ElemF* ef;
  const WalkRec* wr;
  const int4* en;    // node ids
  const int4* te;    // edge ids of the element, increasing
  const int4* ed;    // edge: v0, v1, t0, t1
  const double2* xy; // node coordinates
  const int* edge_labels;
  int nt;
};

// tmp -> packed node records (packed node i is mesh node old[i]);
// K = (CMU/sigc) k^2 / max(eps, EPSILON) per node (eulerian_field_set.pyx:1995-1997),
// the value every particle would compute
__global__ void __launch_bounds__(BLK) k_pack_nodes(const double* u, const double* v, const double* kt,
                                                    const double* ep, double fac, int n, const int* old,
                                                    NodeF* out) {
  int i = blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  int o = old[i];
  NodeF r;
  r.u = u[o];
  r.v = v[o];
  r.K = kt ? fac * kt[o] * kt[o] / fmax(ep[o], EPS) : 0.;
  r.pad = 0.;
  out[i] = r;
}

// The step's time interpolation of every field at mesh node o (k_time_interp:
// tmp = tf f[rec] + (1 - tf) f[rec + 1]) with the node's packed record
// (k_pack_nodes, from the values just interpolated) in one pass: the frames
// and tmp are read and written in mesh order, the record goes to the node's
// packed index nnew[o].  q_u, q_v, q_k, q_e: the positions of u, v, k and eps
// among the interpolated fields (q_k < 0: no k-epsilon diffusivity, K = 0).
// TAB: a captured step, whose record and weight come from the step's entry
// of the table through the device step clock.
struct InterpPack {
  const double* const* frames;  // [ntimes][nf] frame pointers of the interpolated fields
  double* const* out;           // [nf] destinations
  int nf;
  int64_t n;                    // nodes
  int64_t rec;                  // eager steps: the record and weight
  double tf;
  const InterpStep* tab;        // captured steps
  const int* clock;
  int step0;
  int q_u, q_v, q_k, q_e;
  double fac;
  const int* nnew;
  NodeF* rec_out;
};

template <bool TAB>
__global__ void __launch_bounds__(BLK) k_interp_pack(InterpPack p) {
  const int64_t o = (int64_t)blockIdx.x * BLK + threadIdx.x;
  if (o >= p.n) return;
  int64_t rec = p.rec;
  double tf = p.tf;
  if (TAB) {
    const InterpStep e = p.tab[p.clock[0] - p.step0];
    rec = e.rec;
    tf = e.tf;
  }
  const double* const* f0 = p.frames + rec * p.nf;
  const double* const* f1 = f0 + p.nf;
  double u = 0., v = 0., k = 0., ep = 0.;
  for (int q = 0; q < p.nf; ++q) {
    const double val = tf * f0[q][o] + (1. - tf) * f1[q][o];
    p.out[q][o] = val;
    if (q == p.q_u) u = val;
    if (q == p.q_v) v = val;
    if (q == p.q_k) k = val;
    if (q == p.q_e) ep = val;
  }
  NodeF r;
  r.u = u;
  r.v = v;
  r.K = p.q_k >= 0 ? p.fac * k * k / fmax(ep, EPS) : 0.;
  r.pad = 0.;
  p.rec_out[p.nnew[o]] = r;
}

// element records from the mesh tables (built once per mesh); node ids in
// the packed numbering nnew[old]
__global__ void __launch_bounds__(BLK) k_pack_elems(MeshDev m, const int* nnew, ElemF* ef, int4* en, int4* te) {
  int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m.nt) return;
  ElemF r;
  for (int i = 0; i < 3; ++i) {
    int v = m.tri[3 * k + i];
    r.det[i] = m.det[3 * k + i];
    r.gx[i] = m.gx[3 * k + i];
    r.gy[i] = m.gy[3 * k + i];
    r.vx[i] = m.x[v];
    r.vy[i] = m.y[v];
  }
  r.pad = 0.;
  ef[k] = r;
  en[k] = make_int4(nnew[m.tri[3 * k]], nnew[m.tri[3 * k + 1]], nnew[m.tri[3 * k + 2]], 0);
  te[k] = make_int4(m.triedge[3 * k], m.triedge[3 * k + 1], m.triedge[3 * k + 2], 0);
}

__global__ void __launch_bounds__(BLK) k_pack_walk(MeshDev m, WalkRec* wr) {
  int k = blockIdx.x * BLK + threadIdx.x;
  if (k >= m.nt) return;
  WalkRec r;
  r.vv = 0;
  r.pad = 0;
  for (int i = 0; i < 3; ++i) {
    int j = m.triedge[3 * k + i];
    int v0 = m.edges[2 * j], v1 = m.edges[2 * j + 1];
    int a = 0, b = 0;
    for (int q = 0; q < 3; ++q) {
      if (m.tri[3 * k + q] == v0) a = q;
      if (m.tri[3 * k + q] == v1) b = q;
    }
    r.e[i] = j;
    r.vv |= (a | (b << 2)) << (4 * i);
    int t0 = m.edgetri[2 * j], t1 = m.edgetri[2 * j + 1];
    r.nb[i] = (t0 == k) ? t1 : t0;
  }
  wr[k] = r;
}

__global__ void __launch_bounds__(BLK) k_pack_edges(MeshDev m, const int* nnew, const int* old, int4* ed,
                                                    double2* xy) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j < m.ne)
    ed[j] = make_int4(nnew[m.edges[2 * j]], nnew[m.edges[2 * j + 1]], m.edgetri[2 * j], m.edgetri[2 * j + 1]);
  if (j < m.nv) xy[j] = make_double2(m.x[old[j]], m.y[old[j]]);
}

__device__ __forceinline__ bool in_elem(const ElemF& e, double x, double y) {
  // point_in_triangle_dotp (utils.pyx:359-385) on the element's vertices
  return pintri(x, y, e.vx[0], e.vy[0], e.vx[1], e.vy[1], e.vx[2], e.vy[2]) >= 0 &&
         pintri(x, y, e.vx[1], e.vy[1], e.vx[2], e.vy[2], e.vx[0], e.vy[0]) >= 0 &&
         pintri(x, y, e.vx[2], e.vy[2], e.vx[0], e.vy[0], e.vx[1], e.vy[1]) >= 0;
}

__device__ __forceinline__ int in_tri_fast(const FastMesh& fm, int k, double x, double y) {
  int kk = k < 0 ? k + fm.nt : k;
  return in_elem(fm.ef[kk], x, y) ? k : -1;
}

__device__ __forceinline__ double pick3(int i, double a0, double a1, double a2) {
  return i == 0 ? a0 : (i == 1 ? a1 : a2);
}

// orientation class of a determinant (utils.pyx:246-252): 0 collinear, 1, 2
__device__ __forceinline__ int orient_class(double d) { return d == 0. ? 0 : (d > 0. ? 1 : 2); }

__device__ __forceinline__ int pick3i(int i, int a0, int a1, int a2) { return i == 0 ? a0 : (i == 1 ? a1 : a2); }

// xy_localize_point_from_path on the packed tables (xylocalizer.pyx:338-435),
// starting in element k (k >= 0) whose record is e0.  Each pass tests the
// current element's edges in its triedge order, skipping the edge it entered
// by, and moves across the first one the path [p0, p1] intersects
// (intersec2d_segment_segment, intersections.pyx:106-148).
//
// The four orientations of an edge test are the reference's, value for
// value: orient(p0, p1, v) = (p0y - p1y)(vx - p0x) - (p0x - p1x)(vy - p0y)
// depends on the vertex only, so it is computed once per vertex of the
// element and shared by the two edges through that vertex; orient(a, b, p0)
// and orient(a, b, p1) share the edge's own (ay - by) and (ax - bx).  The
// collinear cases (on_segment, utils.pyx:282-303) only run when one of an
// edge's orientations is zero.  The next element's walk record is loaded with
// its vertices.
// One edge test of the walk (intersec2d_segment_segment with the collinear
// cases): edge i of walk record w, the element's vertices and their
// orientation classes c0..c2 with respect to the path.  o1, o2: the classes
// of the edge's first and second vertex.
__device__ __forceinline__ bool walk_edge_hit(const WalkRec& w, int i, double vx0, double vx1, double vx2,
                                              double vy0, double vy1, double vy2, int c0, int c1, int c2,
                                              double p0x, double p0y, double x, double y, int& o1, int& o2) {
  const int ia = (w.vv >> (4 * i)) & 3, ib = (w.vv >> (4 * i + 2)) & 3;
  const double ax = pick3(ia, vx0, vx1, vx2), ay = pick3(ia, vy0, vy1, vy2);
  const double bx = pick3(ib, vx0, vx1, vx2), by = pick3(ib, vy0, vy1, vy2);
  o1 = pick3i(ia, c0, c1, c2);
  o2 = pick3i(ib, c0, c1, c2);
  const double ey = ay - by, ex = ax - bx;
  const int o3 = orient_class(ey * (p0x - ax) - ex * (p0y - ay));
  const int o4 = orient_class(ey * (x - ax) - ex * (y - ay));
  bool hit = o1 != o2 && o3 != o4;
  if (!hit && (o1 == 0 || o2 == 0 || o3 == 0 || o4 == 0))
    hit = (o1 == 0 && on_seg(p0x, p0y, x, y, ax, ay)) || (o2 == 0 && on_seg(p0x, p0y, x, y, bx, by)) ||
          (o3 == 0 && on_seg(ax, ay, bx, by, p0x, p0y)) || (o4 == 0 && on_seg(ax, ay, bx, by, x, y));
  return hit;
}

// From the second pass on, the edge the walk entered by is skipped (the
// reference tests it and discards its result: its id is the last edge), and
// the orientation classes of its two vertices are those the previous pass
// computed (same nodes, same path, same operands), so a pass evaluates one
// vertex orientation and the two other edges, in triedge order, instead of
// three and three.  Each lane selects its two edges, so that the saving is
// real in a wavefront whose lanes entered by different edge slots.
__device__ inline int walk_fast(const FastMesh& fm, int k, const ElemF& e0, double p0x, double p0y, double x,
                                double y, int& bnd_j) {
  bnd_j = -1;
  const double ly = p0y - y, lx = p0x - x;
  double vx0 = e0.vx[0], vx1 = e0.vx[1], vx2 = e0.vx[2];
  double vy0 = e0.vy[0], vy1 = e0.vy[1], vy2 = e0.vy[2];
  WalkRec w = fm.wr[k < 0 ? k + fm.nt : k];
  // first pass: the three vertex classes; slot 2 is tested here and slots 0
  // and 1 by the loop body, as if the walk had entered by slot 2
  int c0 = orient_class(ly * (vx0 - p0x) - lx * (vy0 - p0y));
  int c1 = orient_class(ly * (vx1 - p0x) - lx * (vy1 - p0y));
  int c2 = orient_class(ly * (vx2 - p0x) - lx * (vy2 - p0y));
  int hit_i = -1, h1 = 0, h2 = 0, ie = 2;
  if (walk_edge_hit(w, 2, vx0, vx1, vx2, vy0, vy1, vy2, c0, c1, c2, p0x, p0y, x, y, h1, h2)) hit_i = 2;
  for (int niter = 0; niter <= 50; ++niter) {
    // the two slots other than ie, in triedge order: the first that is hit wins
    const int iA = ie == 0 ? 1 : 0, iB = ie == 2 ? 1 : 2;
    int o1, o2;
    if (walk_edge_hit(w, iB, vx0, vx1, vx2, vy0, vy1, vy2, c0, c1, c2, p0x, p0y, x, y, o1, o2)) {
      hit_i = iB;
      h1 = o1;
      h2 = o2;
    }
    if (walk_edge_hit(w, iA, vx0, vx1, vx2, vy0, vy1, vy2, c0, c1, c2, p0x, p0y, x, y, o1, o2)) {
      hit_i = iA;
      h1 = o1;
      h2 = o2;
    }
    if (hit_i < 0) {
      // no edge crossed: the point is in the current element or nowhere
      bool in = pintri(x, y, vx0, vy0, vx1, vy1, vx2, vy2) >= 0 && pintri(x, y, vx1, vy1, vx2, vy2, vx0, vy0) >= 0 &&
                pintri(x, y, vx2, vy2, vx0, vy0, vx1, vy1) >= 0;
      return in ? k : -1;
    }
    const int j = pick3i(hit_i, w.e[0], w.e[1], w.e[2]), kn = pick3i(hit_i, w.nb[0], w.nb[1], w.nb[2]);
    if (kn == -1) {
      bnd_j = j;
      return -1;
    }
    const ElemF& en = fm.ef[kn];
    w = fm.wr[kn];
    vx0 = en.vx[0]; vx1 = en.vx[1]; vx2 = en.vx[2];
    vy0 = en.vy[0]; vy1 = en.vy[1]; vy2 = en.vy[2];
    // point_in_triangle_dotp on the new element (in_elem)
    bool in = pintri(x, y, vx0, vy0, vx1, vy1, vx2, vy2) >= 0 && pintri(x, y, vx1, vy1, vx2, vy2, vx0, vy0) >= 0 &&
              pintri(x, y, vx2, vy2, vx0, vy0, vx1, vy1) >= 0;
    if (in) return kn;
    k = kn;
    // next pass: the entry edge (slot ie of the new record) is skipped, its
    // vertices keep the classes h1, h2 (the edge's first and second vertex)
    ie = w.e[0] == j ? 0 : (w.e[1] == j ? 1 : 2);
    const int iae = (w.vv >> (4 * ie)) & 3, ibe = (w.vv >> (4 * ie + 2)) & 3;
    const int ic = 3 - iae - ibe;
    const int cc = orient_class(ly * (pick3(ic, vx0, vx1, vx2) - p0x) - lx * (pick3(ic, vy0, vy1, vy2) - p0y));
    c0 = iae == 0 ? h1 : (ibe == 0 ? h2 : cc);
    c1 = iae == 1 ? h1 : (ibe == 1 ? h2 : cc);
    c2 = iae == 2 ? h1 : (ibe == 2 ? h2 : cc);
    hit_i = -1;
  }
  return -1;
}

// xy_localize with method 3 (xylocalizer.pyx:65-120)
__device__ __forceinline__ int localize_fast(const FastMesh& fm, const MeshDev& m, int k, double p0x, double p0y,
                                             double x, double y, int& bnd_j) {
  bnd_j = -1;
  int nk = in_tri_fast(fm, k, x, y);
  if (nk != -1) return nk;
  nk = walk_fast(fm, k, fm.ef[k < 0 ? k + fm.nt : k], p0x, p0y, x, y, bnd_j);
  if (nk == -1 && bnd_j == -1) nk = locate_grid(m, x, y);
  return nk;
}

// Per-scenario parameters of the fused kernel (one entry per scenario)
struct FastScen {
  int diffusion_model;     // 0, 1 (constant), 2 (k-epsilon K, constant vertical), 3 (k-epsilon K in 3D)
  double diff_const;       // sqrt(2*nux)*sqrtdt (model 1)
  double diff_const_z;     // sqrt(2*nuz)*sqrtdt (3D, models 1 and 2)
  int boundary_conditions, max_wall_rebounds, bnd_statistics;
  double damp, strand;
  int deposition;          // 3D: deposition on the bed instead of the bottom reflection
};

struct FastArgs {
  FastMesh fm;
  MeshDev m;               // grid for the fallback search
  const NodeF* nodes;
  const double* xA;
  const double* yA;
  const int* triA;
  const double* upA0;
  const double* upA1;
  const double* usA0;
  const double* usA1;
  double* xB;
  double* yB;
  int* triB;
  double* upB0;
  double* upB1;
  double* usB0;
  double* usB1;
  int* inactive;
  int* del_step;
  const double* diameter;  // per slot (LSM-2)
  ParticleConst pc;        // drag and added-mass constants (LSM-2)
  const int* rank;
  const int64_t* off;
  const int* pool;
  const int* scen;         // scenario of each slot (batched runs) or null
  int j0, n;               // slots [j0, j0 + n)
  int nscen;
  double dt, sqrtdt;
  const FastScen* fsc;     // [nscen]
  const int* scen_end;     // [nscen] number of steps each scenario runs, or null
  int batch_step;          // steps run since the scenarios were set
  const int* scen_base;    // [nscen] first pool index of each scenario, or null
  const int* clock;        // graph replays: {step id, batch step} in device memory, or null
  int nopen, nopen_stride; // open boundaries; stride of bnd_stats per scenario
  int general, guarded;
  int* flag;
  StreamDev* stream;       // [nscen]
  const JumpTable* jt;     // [nscen]
  const JumpTable8* jt8;   // [nscen]
  const ZigTables* zt;
  int* bnd_stats;
  int* n_deleted;
  unsigned long long* err;         // [nscen]
  unsigned long long* stream_end;  // [nscen]
  int step_id;
  int64_t* pos0;      // per slot: stream position of the first draw
  uint64_t* state0;   // per slot: generator state at that position
  double2* xi;        // the step's two normals per (scenario, active rank), see xi_index
  int64_t xi_cap;     // entries of xi
  const int64_t* nact;        // [nscen] active rows
  const int64_t* chunk_base;  // [nscen + 1] first normal-generator chunk of each scenario
  int gen_shift;              // log2 of the generator's ranks per chunk
  const int* pregen;          // [nscen] scenario's normals come from a generator this step
  int xi_by_slot;             // the generator stored each slot's pair at xi[slot] (k_gen_slots)
  int skip_us;                // LSM-1: us (= up) is not stored; downloads derive it from up
  double* rec_xyz;            // deposits (3D): positions on the bed, or null (not recorded)
  int2* rec_key;              // (step, pool index | REC_DEPOSIT)
  int* rec_count;             // [0] records so far
  int rec_cap;
};


__device__ __forceinline__ int step_of(const FastArgs& a) { return a.clock ? a.clock[0] : a.step_id; }
__device__ __forceinline__ int batch_step_of(const FastArgs& a) { return a.clock ? a.clock[1] : a.batch_step; }

// Grid-stride walk over the slots for a persistent grid (a multiple of 8
// blocks).  Blocks are dispatched round-robin over the 8 XCDs, so block b
// takes the chunks of the (b % 8)-th eighth of the slots: neighbouring slots
// (neighbouring elements once sorted) share one XCD's L2.
template <class F, class G>
__device__ __forceinline__ void for_slots_xcd(int n, F&& f, G&& g) {
  const int nx = 8;
  int xcd = blockIdx.x % nx, lb = blockIdx.x / nx, per = gridDim.x / nx;
  int64_t nch = ((int64_t)n + BLK - 1) / BLK;
  int64_t cpx = (nch + nx - 1) / nx;
  for (int64_t c = lb; c < cpx; c += per) {
    int64_t chunk = xcd * cpx + c;
    int64_t j = chunk * BLK + threadIdx.x;
    if (chunk < nch && j < n) f((int)j);
    g();  // every lane, once per iteration (the trip count is block-uniform)
  }
}

template <class F>
__device__ __forceinline__ void for_slots_xcd(int n, F&& f) {
  for_slots_xcd(n, f, [] {});
}

__device__ __forceinline__ void load_jump8(JumpTable8& t8, const JumpTable8* src) {
  const uint64_t* s = reinterpret_cast<const uint64_t*>(src);
  uint64_t* d = reinterpret_cast<uint64_t*>(&t8);
  for (int q = threadIdx.x; q < (int)(sizeof(JumpTable8) / 8); q += BLK) d[q] = s[q];
}

// the fast-path Ziggurat tables (the generator never takes the slow path)
struct ZigFast {
  uint32_t kn[128];
  double wn[128];
};

// First draw of every active slot: the serial stream position of its first
// unit (uniform: units_per_part * active rank; general: pool-order scan) and
// the generator state there.  Lookups of neighbouring slots are independent,
// so this small kernel runs at high occupancy and hides their latency.
template <int RNG>
__global__ void __launch_bounds__(BLK) k_draw_start(FastArgs a) {
  if (a.guarded && a.flag[0] == 0) return;
  __shared__ JumpTable8 t8;
  load_jump8(t8, a.jt8);
  __syncthreads();
  for_slots_xcd(a.n, [&](int jl) {
    int j = a.j0 + jl;
    if (a.inactive[j] != 0) return;
    int s = a.scen ? a.scen[j] : 0;
    const StreamDev& st = a.stream[s];
    int64_t u = a.general ? a.off[j] : (int64_t)st.units_per_part * a.rank[j];
    Draws d;
    d.st = &st;
    int64_t pos;
    if (RNG == 1) pos = u;
    else if (RNG == 2) pos = d.zig_position(u);
    else pos = d.polar_position(u);
    // a position past the scanned window is reported if the slot draws (-1)
    a.pos0[j] = (RNG != 1 && pos >= st.window) ? -1 : pos;
    // the byte-window table depends on the stream increment only: scenarios on
    // streams of scenario 0's increment (one stream advanced per scenario) use
    // the staged copy
    a.state0[j] = st.inc == a.stream[0].inc ? pcg_jump8(st.base, (uint64_t)pos, t8, a.jt[s])
                                            : pcg_jump8(st.base, (uint64_t)pos, a.jt8[s], a.jt[s]);
  });
}

// Normals of the step in pool order.  The active ranks of scenario s are cut
// into chunks of P = 2^gen_shift ranks (P = GEN_PER = 32 in 2D; 16 in 3D,
// where a rank takes three normals and 32-rank chunks leave too few waves to
// hide the draws' latency), numbered globally from chunk_base[s]; thread c
// serves one chunk, i.e. a contiguous run of the scenario's serial stream.  It
// finds the run's first position once (event table), jumps there, then draws
// sequentially: a Ziggurat normal that starts on a slow event takes the
// event's precomputed value and skips its positions; every other normal is a
// one-draw fast normal.  The values of rank r = P c_local + k (chunk c =
// chunk_base[s] + c_local) are stored in tiles of 64 chunks: xi[((c / 64) * P
// + k) * 64 + c % 64], so that each of a wave's stores is one coalesced row and
// its P stores stay within 64 P entries.  Indices stay below nchunk rounded up
// to 64, times P, where nchunk <= (active rows) / P + (scenarios): the buffer
// holds rows + GEN_PER * (scenarios + 2) + 64 * GEN_PER entries for any P <=
// GEN_PER.
constexpr int GEN_PER = 32;
constexpr int GEN_SHIFT_2D = 5, GEN_SHIFT_3D = 4;

__device__ __forceinline__ int64_t xi_tile_index(int64_t c, int64_t k, int sh) {
  return (((c >> 6) << sh) + k) * 64 + (c & 63);
}

__device__ __forceinline__ int64_t xi_index(const int64_t* chunk_base, int sh, int s, int64_t r) {
  return xi_tile_index(chunk_base[s] + (r >> sh), r & ((1 << sh) - 1), sh);
}

// Step preparation in one block, in stream order of the four small passes it
// replaces: each scenario's active total after a rank scan (only when rows
// were deleted or added: *rank_flag != 0), the clearing of that flag, the
// step's units per scenario under uniform assignment (units per row x active
// rows, none once the scenario has ended), and the first generator chunk of
// each scenario (exclusive scan of ceil(active / GEN_PER) over the scenarios
// whose rows the generator serves).
struct PrepArgs {
  const int64_t* scan;       // pool-order scan of the active flags
  const int64_t* total;      // its total
  const int* scen_base;      // [nscen + 1] first pool index of each scenario
  int nscen, npool;
  int* rank_flag;            // rows deleted or added since the last scan
  int64_t* nact;             // [nscen] active rows
  const StreamDev* st;
  const int* scen_end;       // step at which each scenario ends, or null
  int batch_step;
  const int* clock;          // device step clock of a captured period, or null
  int64_t* scen_units;       // [nscen] (uniform only)
  int uniform;
  const int* pregen;         // [nscen] generator serves the scenario, or null
  int64_t* chunk_base;       // [nscen] first generator chunk, or null (no generator)
  int gen_shift;             // log2 of the generator's ranks per chunk
  int layout;                // lay out the step's event tables (ev_layout)
  LayoutArgs ev;
};

__global__ void __launch_bounds__(BLK) k_step_prep(PrepArgs a) {
  __shared__ int64_t part[BLK];
  __shared__ int64_t carry;
  const bool rescan = *a.rank_flag != 0;
  const int step = a.clock ? a.clock[1] : a.batch_step;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();  // every thread has read the flag
  if (threadIdx.x == 0) *a.rank_flag = 0;
  for (int base = 0; base < a.nscen; base += BLK) {
    const int s = base + threadIdx.x;
    int64_t v = 0;
    if (s < a.nscen) {
      int64_t act;
      if (rescan) {
        int e = a.scen_base[s + 1];
        int64_t end = e >= a.npool ? *a.total : a.scan[e];
        act = end - a.scan[a.scen_base[s]];
        a.nact[s] = act;
      } else {
        act = a.nact[s];
      }
      const bool ended = a.scen_end && step >= a.scen_end[s];
      if (a.uniform) a.scen_units[s] = ended ? 0 : (int64_t)a.st[s].units_per_part * act;
      if (a.chunk_base && (!a.pregen || a.pregen[s]) && !ended && a.st[s].units_per_part > 0)
        v = (act + (1 << a.gen_shift) - 1) >> a.gen_shift;
    }
    if (!a.chunk_base) continue;  // uniform across the block
    part[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < BLK; o <<= 1) {
      int64_t t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
      __syncthreads();
      part[threadIdx.x] += t;
      __syncthreads();
    }
    if (s < a.nscen) a.chunk_base[s] = carry + part[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == BLK - 1) carry += part[BLK - 1];
    __syncthreads();
  }
  if (a.chunk_base && threadIdx.x == 0) a.chunk_base[a.nscen] = carry;
  if (a.layout) {
    __shared__ int64_t lds[BLK / 64];
    __syncthreads();  // every scenario's units are stored
    ev_layout(a.ev, lds);
  }
}

// The three normals of a rank in 3D (the diffusion pair, then the vertical
// single), one 32 B record per rank in the tiled layout of the pairs.
struct __align__(32) XiRec3 {
  double a, b, c, pad;
};

// The next four slow events of a generator run (normal index relative to the
// run's first unit, positions taken and value), held in registers: loaded
// before the run's loop and again only once all four are used (a run of 64
// normals holds 0.8 events on average and more than four in about one run of
// 800).  Loads inside the loop would wait, at their use, for every store the
// wave has in flight.  The run's bookkeeping is 32-bit: a run spans at most
// 3 GEN_PER normals.
struct EvQueue {
  static constexpr int32_t NONE = 0x7fffffff;  // no event, or one past the run
  int32_t k0, k1, k2, k3, l0, l1, l2, l3;
  double v0, v1, v2, v3;

  __device__ __forceinline__ void load(const int64_t* key, const int64_t* ex, const double* val, int64_t n,
                                       int64_t ev, int64_t u0) {
    int64_t prev = ev > 0 && ev - 1 < n ? ex[ev - 1] : 0, e;
    auto one = [&](int64_t i, int32_t& k, int32_t& l, double& v) {
      if (i < n) {
        const int64_t kr = key[i] - u0;  // >= 0: the queue starts at the run's first event
        k = kr < NONE ? (int32_t)kr : NONE;
        e = ex[i];
        l = (int32_t)(e - prev + 1);
        prev = e;
        v = val[i];
      } else {
        k = NONE;
        l = 0;
        v = 0.;
      }
    };
    one(ev, k0, l0, v0);
    one(ev + 1, k1, l1, v1);
    one(ev + 2, k2, l2, v2);
    one(ev + 3, k3, l3, v3);
  }

  __device__ __forceinline__ void pop() {
    k0 = k1; l0 = l1; v0 = v1;
    k1 = k2; l1 = l2; v1 = v2;
    k2 = k3; l2 = l3; v2 = v3;
    k3 = NONE;
  }
};

// Normals of ranks [r0, r1) of scenario s, drawn sequentially from the stream
// position of rank r0's first unit, in the order the fused kernels call them:
// NV = 2, the diffusion pair (generate_random_pair, random_utils.pyx:202-233);
// NV = 3, the pair then the vertical single (generate_random_single,
// random_utils.pyx:235-262; under rng_method 3 a single is one polar call of
// which the first value is kept).  store(i, v) receives the values of rank
// r0 + i.  Returns the stream position after the last draw; sets err on a
// window overflow (a value starting past the window: positions only grow, so
// the last value's start is checked) or a slow Ziggurat start missing from
// the event table.
template <int RNG, int NV, class Store>
__device__ __forceinline__ unsigned long long gen_run(const FastArgs& a, int s, int64_t r0, int64_t r1,
                                                      const JumpTable8& t8, uint64_t t8_inc, const ZigFast& zt,
                                                      int& err, Store&& store) {
  const StreamDev& st = a.stream[s];
  // the stream record in registers: the stores below may not alias it, but
  // the compiler cannot tell
  const int64_t window = st.window, n_ev = st.n_ev;
  const int64_t* ev_key = st.ev_key;
  const int64_t* ev_extra = st.ev_extra;
  const double* ev_val = st.ev_val;
  int64_t u = r0 * st.units_per_part;
  Draws d;
  d.st = &st;
  int64_t pos0, ev = 0;
  EvQueue q;
  if (RNG == 1) {
    pos0 = u;
  } else if (RNG == 2) {
    ev = d.events_before(u);
    pos0 = u + (ev > 0 ? ev_extra[ev - 1] : 0);
    q.load(ev_key, ev_extra, ev_val, n_ev, ev, u);
  } else {
    pos0 = 2 * (u + d.events_before(u));
  }
  if (RNG != 1 && pos0 >= window) err = E_RNG_WINDOW;
  Pcg g;
  g.inc = st.inc;
  g.state = st.inc == t8_inc ? pcg_jump8(st.base, (uint64_t)pos0, t8, a.jt[s])
                             : pcg_jump8(st.base, (uint64_t)pos0, a.jt8[s], a.jt[s]);
  int32_t prel = 0, plast = 0, urel = 0;  // positions taken, start of the last value, units drawn
  const int nr = (int)(r1 - r0);
  for (int i = 0; i < nr; ++i) {
    double v[NV];
    if (RNG == 1) {
#pragma unroll
      for (int k = 0; k < NV; ++k) v[k] = rect_normal(g);
      prel += NV;
    } else if (RNG == 2) {
      for (int k = 0; k < NV; ++k, ++urel) {
        plast = prel;
        if (urel == q.k0) {
          const int32_t len = q.l0;
          v[k] = q.v0;
          for (int32_t t = 0; t < len; ++t) pcg_next(g);
          prel += len;
          ++ev;
          q.pop();
          if (q.k0 == EvQueue::NONE && ev < n_ev) q.load(ev_key, ev_extra, ev_val, n_ev, ev, u);  // all four used
        } else {
          int32_t hz = (int32_t)pcg_next(g);
          uint32_t iz = (uint32_t)hz & 127u;
          uint32_t ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;
          if (ahz >= zt.kn[iz]) err = E_RNG_EVENTS;  // a slow start missing from the table
          v[k] = hz * zt.wn[iz];
          prel += 1;
        }
      }
    } else {
      for (int k = 0; k < NV; k += 2) {
        double x1 = 0., x2 = 0., w = 2.0;
        while (w >= 1.0 || w == 0.0) {
          x1 = 2.0 * u32_to_uniform(pcg_next(g)) - 1.0;
          x2 = 2.0 * u32_to_uniform(pcg_next(g)) - 1.0;
          w = x1 * x1 + x2 * x2;
          prel += 2;
        }
        w = sqrt((-2.0 * cr_log(w)) / w);
        v[k] = x1 * w;
        if (k + 1 < NV) v[k + 1] = x2 * w;
      }
    }
    store(i, v);
  }
  if (RNG == 2 && nr > 0 && pos0 + plast >= window) err = E_RNG_WINDOW;
  return (unsigned long long)(pos0 + prel);
}

// Normals of generator chunk c (scenario s): the NV values of each of its
// ranks, in the serial stream order from the chunk's first position.  Returns
// the stream position after its last draw (0 if the scenario draws nothing).
template <int RNG, int NV>
__device__ __forceinline__ unsigned long long gen_chunk(const FastArgs& a, int64_t c, int s, const JumpTable8& t8,
                                                        uint64_t t8_inc, const ZigFast& zt) {
  const StreamDev& st = a.stream[s];
  if (st.units_per_part == 0) return 0;
  const int64_t per = 1 << a.gen_shift;
  int64_t r0 = (c - a.chunk_base[s]) << a.gen_shift, r1 = r0 + per < a.nact[s] ? r0 + per : a.nact[s];
  int err = 0;
  unsigned long long pend = gen_run<RNG, NV>(a, s, r0, r1, t8, t8_inc, zt, err, [&](int i, const double* v) {
    int64_t xidx = xi_tile_index(c, i, a.gen_shift);
    if (xidx >= a.xi_cap) return;
    if (NV == 2) {
      a.xi[xidx] = make_double2(v[0], v[1]);
    } else {
      XiRec3 q;
      q.a = v[0];
      q.b = v[1];
      q.c = v[NV - 1];
      q.pad = 0.;
      reinterpret_cast<XiRec3*>(a.xi)[xidx] = q;
    }
  });
  // reported against pool index 0: the rank's slot is not known here
  if (err) atomicMin(&a.err[s], (unsigned long long)(unsigned)err);
  return pend;
}

// NV values per rank: 2 for the 2D fused kernels (pairs in a.xi), 3 for the
// 3D one (XiRec3 records in the same buffer, which then holds 32 B per entry)
template <int RNG, int NV>
__global__ void __launch_bounds__(BLK) k_gen_normals(FastArgs a) {
  __shared__ JumpTable8 t8;
  __shared__ ZigFast zt;
  __shared__ uint64_t t8_inc;
  __shared__ int t8_set;
  if (RNG == 2) {
    for (int q = threadIdx.x; q < 128; q += BLK) {
      zt.kn[q] = a.zt->kn[q];
      zt.wn[q] = a.zt->wn[q];
    }
  }
  if (threadIdx.x == 0) t8_set = 0;
  __syncthreads();
  const int64_t nchunk = a.chunk_base[a.nscen];
  // the block's BLK chunks of an iteration are consecutive, so they mostly
  // belong to one scenario: its byte-window jump table, which depends on the
  // stream increment only, is staged in LDS (reloaded when the block reaches a
  // scenario of another increment); threads on chunks of a scenario with
  // another increment read its table from memory
  for (int64_t cb = (int64_t)blockIdx.x * BLK; cb < nchunk; cb += (int64_t)gridDim.x * BLK) {
    int sb = scen_of(a.chunk_base, a.nscen, cb);
    if (!t8_set || a.stream[sb].inc != t8_inc) {  // block-uniform
      __syncthreads();
      load_jump8(t8, a.jt8 + sb);
      if (threadIdx.x == 0) {
        t8_inc = a.stream[sb].inc;
        t8_set = 1;
      }
      __syncthreads();
    }
    int64_t c = cb + threadIdx.x;
    int s = -1;
    unsigned long long pend = 0;
    if (c < nchunk) {
      s = (c < a.chunk_base[sb + 1]) ? sb : scen_of(a.chunk_base, a.nscen, c);
      if (!(a.scen_end && batch_step_of(a) >= a.scen_end[s])) pend = gen_chunk<RNG, NV>(a, c, s, t8, t8_inc, zt);
    }
    (void)pend;  // the streams' ends come from the units drawn (k_finish)
  }
}

// Normals of the step in slot order: every active slot of a scenario that
// draws (pregen[s]) jumps to the stream position of its own rank's first unit
// and draws its pair there, stored at xi[slot], so that the fused kernel reads
// them coalesced.  The pool-order generator above draws 32 ranks per jump but
// its output, read at xi[rank], is a scattered read once the slots are sorted
// by element; here each slot pays a jump (4 table lookups in LDS and 64-bit
// multiply-adds) and every access is coalesced.
template <int RNG>
__global__ void __launch_bounds__(BLK) k_gen_slots(FastArgs a) {
  __shared__ JumpTable8 t8;
  __shared__ ZigFast zt;
  if (RNG == 2)
    for (int q = threadIdx.x; q < 128; q += BLK) {
      zt.kn[q] = a.zt->kn[q];
      zt.wn[q] = a.zt->wn[q];
    }
  load_jump8(t8, a.jt8);
  __syncthreads();
  const uint64_t t8_inc = a.stream[0].inc;
  for_slots_xcd(a.n, [&](int jl) {
    int j = a.j0 + jl;
    if (a.inactive[j] != 0) return;
    int s = a.scen ? a.scen[j] : 0;
    if (!a.pregen[s] || a.stream[s].units_per_part == 0) return;
    if (a.scen_end && batch_step_of(a) >= a.scen_end[s]) return;
    int err = 0;
    int64_t r = a.rank[j];
    unsigned long long pend = gen_run<RNG, 2>(a, s, r, r + 1, t8, t8_inc, zt, err,
                                              [&](int, const double* v) {
                                                if (j < a.xi_cap) a.xi[j] = make_double2(v[0], v[1]);
                                              });
    (void)pend;  // the streams' ends come from the units drawn (k_finish)
    if (err) atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)err);
  });
}

// One particle of the step; returns the stream position after its last draw
// (0 if it drew nothing or its normals were generated beforehand).
// MODEL 1: euler_lsm1 + diffusion_lsm1.  MODEL 2: di_lsm2 without added mass
// and without velocity noise (diffusion_lsm2_option 1), then diffusion_lsm2,
// which is diffusion_lsm1 (update_state_lsm2.pyx:30-139, 1329-1543).
// PREGEN: the step's two normals of the slot are xi[rank] (uniform draw
// assignment); otherwise they are sampled here from the prepared start.
template <int MODEL, int RNG, bool PREGEN>
__device__ __forceinline__ unsigned long long fast_particle(const FastArgs& a, int j, const ZigTables& zt) {
  // The slot's loads, then the element record, node ids and normals they
  // index, are issued before any branch, so that their latencies overlap
  // instead of adding up.  The indices are clamped into their tables; rows
  // that do not use a value (inactive, no element, no draws) ignore it.
  const int ina = a.inactive[j];
  const int s = a.scen ? a.scen[j] : 0;
  const int tri0 = a.triA[j];
  const double x0 = a.xA[j], y0 = a.yA[j];
  int64_t xidx = 0;
  if (PREGEN) xidx = a.xi_by_slot ? (int64_t)j : xi_index(a.chunk_base, a.gen_shift, s, a.rank[j]);
  const int te = (unsigned)tri0 < (unsigned)a.fm.nt ? tri0 : 0;
  const ElemF& ef = a.fm.ef[te];
  const int4 en = a.fm.en[te];
  // the P1 part of the record, read here with the node ids
  double det[3], gx[3], gy[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    det[k] = ef.det[k];
    gx[k] = ef.gx[k];
    gy[k] = ef.gy[k];
  }
  const double2 xi_raw = PREGEN ? a.xi[(uint64_t)xidx < (uint64_t)a.xi_cap ? xidx : 0] : make_double2(0., 0.);
  if (ina != 0) return 0;
  if (a.scen_end && batch_step_of(a) >= a.scen_end[s]) {
    // the scenario has run all its steps: its rows carry over unchanged
    a.xB[j] = x0;
    a.yB[j] = y0;
    a.triB[j] = tri0;
    a.upB0[j] = a.upA0[j];
    a.upB1[j] = a.upA1[j];
    a.usB0[j] = a.usA0[j];
    a.usB1[j] = a.usA1[j];
    return 0;
  }
  const StreamDev& st = a.stream[s];
  const FastScen& q = a.fsc[s];
  if (tri0 < 0) {
    // the particle has no element: the reference raises in the velocity lookup
    atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)E_NOT_IN_TRI);
    a.xB[j] = x0;
    a.yB[j] = y0;
    a.triB[j] = tri0;
    return 0;
  }
  // the pair of normals of this slot's active rank
  double2 xi = make_double2(0., 0.);
  if (PREGEN && q.diffusion_model != 0 && st.units_per_part > 0 && xidx < a.xi_cap) xi = xi_raw;
  // P1 factors at the start position (eulerian_field_set.pyx:1103-1105)
  double f0 = det[0] + x0 * gx[0] + y0 * gy[0];
  double f1 = det[1] + x0 * gx[1] + y0 * gy[1];
  double f2 = det[2] + x0 * gx[2] + y0 * gy[2];
  NodeF n0 = a.nodes[en.x], n1 = a.nodes[en.y], n2 = a.nodes[en.z];
  double ux = n0.u * f0 + n1.u * f1 + n2.u * f2;
  double uy = n0.v * f0 + n1.v * f1 + n2.v * f2;
  double x, y, upx = ux, upy = uy;  // LSM-1: particle velocity = fluid velocity seen = U
  int perr = 0;
  if (MODEL == 1) {
    // euler_lsm1
    x = x0 + ux * a.dt;
    y = y0 + uy * a.dt;
  } else {
    // di_lsm2: exact Ornstein-Uhlenbeck drift of the particle velocity
    double u0x = a.upA0[j], u0y = a.upA1[j];
    double dx = ux - u0x, dy = uy - u0y;
    double a1 = relax_rate_dp(a.pc, a.diameter[j], sqrt(dx * dx + dy * dy));
    double beta, D1, D2, fxc, um_beta;
    if (!lsm2_drift_coefficients(a1, a.dt, beta, D1, D2, fxc, um_beta)) {
      perr = E_LSM2_COEF;
      beta = D1 = D2 = fxc = um_beta = 0.;
    }
    double Fx = a.pc.a2c * 0., Fy = a.pc.a2c * 0.;  // no added mass: zero fluid acceleration
    upx = u0x * beta + ux * um_beta + Fx * D1;
    upy = u0y * beta + uy * um_beta + Fy * D1;
    x = x0 + u0x * D1 + ux * D2 + Fx * fxc;
    y = y0 + u0y * D1 + uy * D2 + Fy * fxc;
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
  // diffusion_lsm1 with K and grad K at the start-of-step element
  if (q.diffusion_model != 0) {
    double diffx, driftx = 0., drifty = 0.;
    if (q.diffusion_model >= 2) {
      double K = n0.K * f0 + n1.K * f1 + n2.K * f2;
      double dkdx = n0.K * gx[0] + n1.K * gx[1] + n2.K * gx[2];
      double dkdy = n0.K * gy[0] + n1.K * gy[1] + n2.K * gy[2];
      diffx = sqrt(2. * fmax(0., K)) * a.sqrtdt;
      driftx = dkdx * a.dt;
      drifty = dkdy * a.dt;
    } else {
      diffx = q.diff_const;
    }
    double xix = 0., xiy = 0.;
    if (diffx != 0.) {
      if (PREGEN) {
        xix = xi.x;
        xiy = xi.y;
        dr.nunits = RNG == 3 ? 1 : 2;  // one pair call (polar) or two normals
      } else {
        dr.pair(xix, xiy);
      }
    }
    x += driftx + diffx * xix;
    y += drifty + diffx * xiy;
  }
  // localization from (tri0, x0) along the path (update_state_lsm1.pyx:113-124):
  // xy_localize method 3 after its point-in-triangle test of tri0
  int bnd_j = -1, tri = tri0;
  if (!in_elem(ef, x, y)) {
    tri = walk_fast(a.fm, tri0, ef, x0, y0, x, y, bnd_j);
    if (tri == -1 && bnd_j == -1) tri = locate_grid(a.m, x, y);
  }
  bool deleted = false;
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
          double coef = 1. - q.damp / 2.;
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
            // specular reflection of the particle velocity on the wall (normal2u,
            // utils.pyx:99-119; compute_bc.pyx rebound branch), damped
            double ex = v1.x - v0.x, ey = v1.y - v0.y;
            double nrm = sqrt(ex * ex + ey * ey);
            double nn0 = (v0.y - v1.y) / nrm, nn1 = (v1.x - v0.x) / nrm;
            double up0 = upx;
            upx = up0 * nn0 + upy * nn1;
            upy = -up0 * nn1 + upy * nn0;
            upx = -1. * upx;
            up0 = upx;
            upx = up0 * nn0 - upy * nn1;
            upy = up0 * nn1 + upy * nn0;
            upx *= 1. - q.damp;
            upy *= 1. - q.damp;
          }
          ++rc;
          if (tri == PART_LOC_O && bjj == -1) break;
        }
      }
      if (lab >= 1 && !deleted) {
        deleted = true;
        if (q.bnd_statistics && a.bnd_stats && lab - 1 < a.nopen) atomicAdd(&a.bnd_stats[s * a.nopen_stride + lab - 1], 1);
      }
    }
    if (!deleted && tri == PART_LOC_O) deleted = true;
  } else if (tri == PART_LOC_O) {
    deleted = true;
  }
  if (perr) atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)perr);
  if (dr.err) atomicMin(&a.err[s], ((unsigned long long)(unsigned)a.pool[j] << 8) | (unsigned)dr.err);
  if (!a.general && dr.nunits != st.units_per_part) atomicOr(&a.flag[0], 1);
  if (deleted) {
    a.inactive[j] = 2;
    a.del_step[j] = step_of(a);
    atomicAdd(a.n_deleted, 1);
    x = 0.;
    y = 0.;
    tri = -1;
    ux = 0.;
    uy = 0.;
    upx = 0.;
    upy = 0.;
  }
  a.xB[j] = x;
  a.yB[j] = y;
  a.triB[j] = tri;
  a.upB0[j] = upx;
  a.upB1[j] = upy;
  if (MODEL != 1 || !a.skip_us) {
    a.usB0[j] = ux;
    a.usB1[j] = uy;
  }
  return dr.live ? (unsigned long long)(dr.pos0 + dr.ncalls) : 0ull;
}

// One slot per thread, chunks mapped to XCDs by xcd_chunk (grid: fast_grid).
// Without the grid-stride loop of a persistent launch, the kernel's uniform
// values need not stay live across slots, which spares scalar registers.  The
// Ziggurat tables are staged in LDS per block for the inline sampler.

template <int MODEL, int RNG, bool PREGEN>
__device__ __forceinline__ void fast_step_body(const FastArgs& a) {
  if (a.guarded && a.flag[0] == 0) return;
  __shared__ ZigTables zt;
  if (RNG == 2 && !PREGEN) {
    for (int q = threadIdx.x; q < 128; q += BLK) {
      zt.kn[q] = a.zt->kn[q];
      zt.wn[q] = a.zt->wn[q];
      zt.fn[q] = a.zt->fn[q];
    }
    __syncthreads();
  }
  const int64_t chunk = xcd_chunk(a.n);
  if (chunk < 0) return;
  const int64_t jl = chunk * BLK + threadIdx.x;
  if (jl < a.n) (void)fast_particle<MODEL, RNG, PREGEN>(a, a.j0 + (int)jl, zt);
}

// 5 waves per SIMD (<= 96 VGPRs) with pre-generated normals, 4 with the inline sampler
template <int RNG, bool PREGEN>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(PREGEN ? 5 : 4))) k_lsm1_2d(FastArgs a) {
  fast_step_body<1, RNG, PREGEN>(a);
}

template <int RNG, bool PREGEN>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) k_lsm2_2d(FastArgs a) {
  fast_step_body<2, RNG, PREGEN>(a);
}

}  // namespace cyl
