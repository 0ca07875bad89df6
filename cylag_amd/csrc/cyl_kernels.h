// Kernels of the step pipeline.
//
//  k_step        fused particle update: arrival init, LSM update, walk,
//                boundary conditions; one thread per pool slot (physical order)
//  k_time_interp tmp = tf*f[rec] + (1-tf)*f[rec+1] on every node of every field
//  pool scans    scatter a per-slot value to pool order, exclusive scan with a
//                reset at every scenario, gather back (active rank, draw offsets)
//  rng events    slow Ziggurat starts / rejected polar attempts in the step's
//                stream window, resolved into the serial chain of normals
//  sort          counting sort of slots by element, then permutation of state
#pragma once
#include "cyl_particle.h"

namespace cyl {

constexpr int BLK = 256;

struct StepArgs {
  MeshDev m;
  FieldsDev f;
  ParamsDev par;
  ParticleConst pc;
  PoolDev A, B;       // input / output state
  SlotDev slot;
  const int* init_tri;
  int* del_step;      // per slot: step at which the row was deleted
  double* depth;      // per slot (compute_depth) or null
  int n;              // physical slots
  int dim;
  double dt;
  int step_id;
  StreamDev* stream;  // device copy, one per scenario
  const JumpTable* jt;
  const JumpTable8* jt8;
  const ZigTables* zt;
  int* bnd_stats;     // [nscen][nopen] of this step
  int nopen;       // stride of bnd_stats
  int nopen_real;  // open boundaries of the mesh
  int* n_deleted;
  unsigned long long* err;        // per scenario
  unsigned long long* stream_end; // per scenario: max end position
  int* flag;          // [0] mispredict, [1] general-mode request consumed
  int* units_pool;    // COUNT mode: units per global pool index
  const int* gpool;   // global pool index of each slot
  int init_only;      // 1: only initialize tri == -2 rows
  int general;        // 1: offsets come from slot.off
  int guarded;        // general-mode launch that runs only if flag[0] is set
  int j0;             // first slot of the launch (slots [j0, j0 + n))
  const ParamsDev* pars;  // per-scenario parameters (batched runs) or null: par
  const int* scen_end;    // per scenario: number of steps it runs, or null
  int batch_step;         // steps run since the scenarios were set
  const int* scen_base;   // per scenario: first pool index, or null (one scenario)
  const int* clock;       // graph replays: {step id, batch step} in device memory, or null
  double* strand_xyz;     // stranding_output: stranded positions, or null
  int2* strand_key;       // (step, pool index) of each stranded position
  int* strand_count;      // [0] records, [1] records at the start of the step
  int strand_cap;
};

__device__ __forceinline__ int step_of(const StepArgs& a) { return a.clock ? a.clock[0] : a.step_id; }
__device__ __forceinline__ int batch_step_of(const StepArgs& a) { return a.clock ? a.clock[1] : a.batch_step; }

__device__ __forceinline__ void load_part(const PoolDev& A, const SlotDev& s, int j, int dim, Part& p) {
  p.x[0] = A.x[j];
  p.x[1] = A.y[j];
  p.x[2] = dim == 3 ? A.z[j] : 0.;
  for (int c = 0; c < 3; ++c) {
    p.up[c] = c < dim ? A.up[c][j] : 0.;
    p.us[c] = c < dim ? A.us[c][j] : 0.;
    p.lx[c] = p.x[c];
    p.lup[c] = p.up[c];
  }
  p.tri = A.tri[j];
  p.ll = A.ll[j];
  p.last_tri = p.tri;
  p.inactive = A.inactive[j];
  p.deleted = 0;
  p.nrebound = 0;
  p.diameter = s.diameter ? s.diameter[j] : 0.;
  p.volume = s.volume ? s.volume[j] : 0.;
}

__device__ __forceinline__ void store_part(const PoolDev& B, int j, int dim, const Part& p, int inactive_code) {
  if (inactive_code) B.inactive[j] = inactive_code;  // per-slot flags are updated in place
  B.x[j] = p.x[0];
  B.y[j] = p.x[1];
  if (dim == 3) B.z[j] = p.x[2];
  for (int c = 0; c < dim; ++c) {
    B.up[c][j] = p.up[c];
    B.us[c][j] = p.us[c];
  }
  B.tri[j] = p.tri;
  B.ll[j] = p.ll;
}

// One pool slot; returns the stream position after the slot's last draw (0 if none).
template <int DIM, int MODEL, bool COUNT>
__device__ __forceinline__ unsigned long long step_slot(const StepArgs& a, int j, int& scen) {
  Part p;
  load_part(a.A, a.slot, j, DIM, p);
  scen = a.slot.scen ? a.slot.scen[j] : 0;
  if (p.inactive) {
    if (COUNT) a.units_pool[a.gpool[j]] = 0;
    return 0;
  }
  if (a.scen_end && batch_step_of(a) >= a.scen_end[scen] && !a.init_only) {
    // the scenario has run all its steps: its rows carry over unchanged
    if (COUNT) a.units_pool[a.gpool[j]] = 0;
    else store_part(a.B, j, DIM, p, 0);
    return 0;
  }
  const StreamDev& st = a.stream[scen];
  const ParamsDev& par = a.pars ? a.pars[scen] : a.par;
  Ctx c{a.m, a.f, par, a.pc, a.dt, DIM, 0};
  Draws dr;
  dr.st = &st;
  dr.jt = &a.jt[scen];
  dr.jt8 = a.jt8 ? &a.jt8[scen] : nullptr;
  dr.pre_pos = nullptr;
  dr.pre_state = nullptr;
  dr.slot = j;
  dr.zt = a.zt;
  dr.method = par.rng_method;
  dr.dry = COUNT;
  dr.nunits = 0;
  dr.ncalls = 0;
  dr.pos0 = 0;
  dr.live = false;
  dr.err = 0;
  dr.unit0 = 0;
  if (!COUNT) {
    if (a.general) dr.unit0 = a.slot.off[j];
    else dr.unit0 = (int64_t)st.units_per_part * a.slot.rank[j];
  }
  int pool = a.slot.pool[j];
  int bnd_j = -1;
  bool was_new = p.tri == -2;
  if (was_new) init_particle(c, p, dr, a.init_tri ? a.init_tri[j] : -3);
  if (!p.inactive && !c.err && !a.init_only) {
    if (MODEL == 0) bnd_j = update_test(c, p);
    else if (MODEL == 1) bnd_j = update_lsm1(c, p, dr);
    else if (MODEL == 2) bnd_j = update_lsm2(c, p, dr);
    else bnd_j = update_lsm3(c, p, dr);
    if (!c.err && !COUNT) {
      // draws are counted before the boundary conditions, which use a sub-stream
      if (par.boundary_conditions) {
        StrandRec sr{a.strand_xyz, a.strand_key, a.strand_count, a.strand_cap, step_of(a), pool};
        if (a.strand_xyz && !a.init_only) c.sr = &sr;
        BcDraws bd;
        bd.method = par.rng_method;
        bd.zt = a.zt;
        bd.g.inc = 0x6c62272e07bb0143ULL | 1u;
        int lpool = pool - (a.scen_base ? a.scen_base[scen] : 0);  // row of the scenario's own pool
        bd.g.state = st.base ^ ((uint64_t)(unsigned)lpool * 0x9e3779b97f4a7c15ULL);
        compute_bc(c, p, bnd_j, bd, a.bnd_stats ? a.bnd_stats + scen * a.nopen : nullptr, a.nopen_real);
      } else {
        if (p.tri == PART_LOC_O) delete_part(p, DIM);
        if (p.ll == PART_LOC_B) delete_part(p, DIM);
        if (p.ll == PART_LOC_A) delete_part(p, DIM);
      }
      if (par.compute_depth && !p.inactive && DIM == 3 && a.depth)
        a.depth[j] = field2d(a.f, a.m, a.f.fid_zs, p.tri, a.f.nlayers - 1, p.x[0], p.x[1]) - p.x[2];
    }
  }
  if (was_new && !p.inactive && par.compute_depth && DIM == 3 && a.depth && !COUNT && !c.err && a.init_only)
    a.depth[j] = field2d(a.f, a.m, a.f.fid_zs, p.tri, a.f.nlayers - 1, p.x[0], p.x[1]) - p.x[2];
  if (dr.err && !c.err) c.err = dr.err;
  if (c.err) atomicMin(&a.err[scen], ((unsigned long long)(unsigned)pool << 8) | (unsigned)c.err);
  if (COUNT) {
    a.units_pool[a.gpool[j]] = (int)dr.nunits;
    return 0;
  }
  if (!a.general && !a.init_only && dr.nunits != st.units_per_part) atomicOr(&a.flag[0], 1);
  int code = 0;
  if (p.deleted) {
    code = 2;
    atomicAdd(a.n_deleted, 1);
    if (a.del_step) a.del_step[j] = step_of(a);
  } else if (p.inactive) {
    code = 1;
  }
  store_part(a.B, j, DIM, p, code);
  return dr.live ? (unsigned long long)(dr.pos0 + dr.ncalls) : 0ull;
}

// Max of v into base[key] for the lanes of a wave (lanes with key < 0 or
// v == 0 take no part): one atomic per distinct key in the wave.  Called by
// every lane of the wave.
__device__ __forceinline__ void wave_max_by_key(int key, unsigned long long v, unsigned long long* base) {
  bool pending = key >= 0 && v != 0;
  for (;;) {
    const unsigned long long m = __ballot(pending);
    if (m == 0) break;
    const int leader = __ffsll((long long)m) - 1;
    const int k0 = __shfl(key, leader);
    const bool mine = pending && key == k0;
    unsigned long long x = mine ? v : 0ull;
    for (int o = 32; o > 0; o >>= 1) {
      unsigned long long y = __shfl_xor(x, o);
      x = y > x ? y : x;
    }
    if ((int)(threadIdx.x & 63) == leader) atomicMax(&base[k0], x);
    if (mine) pending = false;
  }
}

// Running stream end of the scenario whose rows a thread steps.  A thread
// that moves on to another scenario parks the finished one's end; flush(),
// called by every lane of the wave once per loop iteration, merges parked ends
// with one atomic per scenario in the wave, and done() merges the rest.
struct StreamEndAcc {
  unsigned long long e = 0, pe = 0;
  int es = -1, ps = -1;
  __device__ __forceinline__ void add(int s, unsigned long long q) {
    if (es >= 0 && es != s) {  // at most once per iteration: flush() follows each
      ps = es;
      pe = e;
      e = 0;
    }
    es = s;
    e = q > e ? q : e;
  }
  __device__ __forceinline__ void flush(unsigned long long* base) {
    if (__any(pe != 0)) {
      wave_max_by_key(ps, pe, base);
      pe = 0;
      ps = -1;
    }
  }
  __device__ __forceinline__ void done(unsigned long long* base) { wave_max_by_key(es, e, base); }
};

// Grid-stride over the slots of the launch (the grid is at most what fits on
// the device at once, so a guarded launch that finds nothing to do ends at once).
// The streams' ends come from the units drawn (k_finish).
template <int DIM, int MODEL, bool COUNT>
__global__ void __launch_bounds__(BLK) k_step(StepArgs a) {
  if (a.guarded && a.flag[0] == 0) return;
  for (int64_t jl = (int64_t)blockIdx.x * BLK + threadIdx.x; jl < a.n; jl += (int64_t)gridDim.x * BLK) {
    int scen = -1;
    (void)step_slot<DIM, MODEL, COUNT>(a, a.j0 + (int)jl, scen);
  }
}

// tmp[fid][i] = tf*frame_rec[i] + (1-tf)*frame_rec1[i]  (eulerian_field_set.pyx:926-938)
__global__ void __launch_bounds__(BLK) k_time_interp(const double* const* f0, const double* const* f1,
                                                     double* const* out, int nf, int64_t size, double tf) {
  int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int q = blockIdx.y;
  if (i >= size || q >= nf) return;
  out[q][i] = tf * f0[q][i] + (1. - tf) * f1[q][i];
}

// Record and weight of one step's time interpolation, computed by the host for
// the steps of a replayed graph (the record bracketing the step's time and
// tf = (t - t[rec+1]) / (t[rec] - t[rec+1]), eulerian_field_set.pyx:909-916)
struct InterpStep {
  double tf;
  int64_t rec;
};

// k_time_interp of a captured step: the step's entry of the table, found
// through the device step clock, selects the frames and the weight
__global__ void __launch_bounds__(BLK) k_time_interp_tab(const double* const* frames, double* const* out, int nf,
                                                         int64_t size, const InterpStep* tab, const int* clock,
                                                         int step0) {
  int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x;
  int q = blockIdx.y;
  if (i >= size || q >= nf) return;
  const InterpStep e = tab[clock[0] - step0];
  const double* const* f0 = frames + e.rec * nf;
  out[q][i] = e.tf * f0[q][i] + (1. - e.tf) * f0[nf + q][i];
}

// ---------------------------------------------------------------------------
// Pool-order scans
// ---------------------------------------------------------------------------
// The per-slot kernels guarded by a device flag run grid-stride on a capped
// grid (GUARD_GRID blocks): when the flag is clear, they end as fast as an
// empty launch instead of retiring one block per 256 slots.
constexpr int GUARD_GRID = 2048;

__global__ void __launch_bounds__(BLK) k_scatter_active(const int* inactive, const int* gpool, int n,
                                                        int* vals, const int* cond) {
  if (cond && *cond == 0) return;
  for (int j = blockIdx.x * BLK + threadIdx.x; j < n; j += gridDim.x * BLK) vals[gpool[j]] = inactive[j] == 0 ? 1 : 0;
}

// block-local exclusive scan into int64, block totals out; elements at or past
// the device count *ndev (if given) read as zero
// Exclusive scans (sum, max) in three passes: tiles of BLK x SCAN_ITEMS
// elements, one block over the tile totals, then the tile prefix is folded in.
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = BLK * SCAN_ITEMS;

struct OpSum {
  static constexpr int64_t ident = 0;
  __device__ __forceinline__ int64_t operator()(int64_t a, int64_t b) const { return a + b; }
};
struct OpMax {
  static constexpr int64_t ident = -1;  // reach keys are >= -1
  __device__ __forceinline__ int64_t operator()(int64_t a, int64_t b) const { return a > b ? a : b; }
};

// exclusive scan of one value per thread over a block of NT threads; *total
// receives the block's reduction.  lds holds NT/64 entries.
template <class Op, int NT>
__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t& total, int64_t* lds) {
  Op op;
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    int64_t t = __shfl_up(incl, o);
    if (lane >= o) incl = op(incl, t);
  }
  if (lane == 63) lds[w] = incl;
  __syncthreads();
  int64_t wpre = Op::ident, tot = Op::ident;
  for (int q = 0; q < NT / 64; ++q) {
    int64_t x = lds[q];
    if (q < w) wpre = op(wpre, x);
    tot = op(tot, x);
  }
  int64_t excl = __shfl_up(incl, 1);
  if (lane == 0) excl = Op::ident;
  __syncthreads();
  total = tot;
  return op(wpre, excl);
}

// tiles that hold live entries: every tile when no device count is given,
// else those covering indices 0..min(*ndev, n) (index *ndev receives the total)
__device__ __forceinline__ int64_t scan_live_tiles(int64_t n, const int64_t* ndev) {
  int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (!ndev) return nt;
  int64_t lim = *ndev < n ? *ndev : n;
  int64_t t = lim / SCAN_TILE + 1;
  return t < nt ? t : nt;
}

// grid-stride over the live tiles; entries at or past the device count read
// as the identity, so out[*ndev] is the total
template <class T, class Op>
__global__ void __launch_bounds__(BLK) k_scan_tiles(const T* in, int64_t* out, int64_t* tile_sum, int64_t n,
                                                    const int64_t* ndev, const int* cond) {
  if (cond && *cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  __shared__ int64_t buf[SCAN_TILE];
  Op op;
  const int64_t lim = ndev ? (*ndev < n ? *ndev : n) : n;
  const int64_t ntl = scan_live_tiles(n, ndev);
  for (int64_t t = blockIdx.x; t < ntl; t += gridDim.x) {
    const int64_t base = t * SCAN_TILE;
    // coalesced load of the tile, then each thread takes SCAN_ITEMS consecutive values
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      int64_t i = base + k * BLK + threadIdx.x;
      buf[k * BLK + threadIdx.x] = i < lim ? (int64_t)in[i] : Op::ident;
    }
    __syncthreads();
    int64_t v[SCAN_ITEMS];
    int64_t run = Op::ident;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      int64_t x = buf[threadIdx.x * SCAN_ITEMS + k];
      v[k] = run;  // exclusive within the thread
      run = op(run, x);
    }
    int64_t total;
    int64_t pre = block_excl_scan<Op, BLK>(run, total, lds);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) buf[threadIdx.x * SCAN_ITEMS + k] = op(pre, v[k]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      int64_t i = base + k * BLK + threadIdx.x;
      if (i < n) out[i] = buf[k * BLK + threadIdx.x];
    }
    if (threadIdx.x == 0) tile_sum[t] = total;
    __syncthreads();  // buf and lds are refilled by the next tile
  }
}

// exclusive scan of the live tile totals by one block of 1024 threads
template <class Op>
__global__ void __launch_bounds__(1024) k_scan_top(int64_t* ts, int64_t n, const int64_t* ndev, int64_t* total,
                                                   const int* cond) {
  if (cond && *cond == 0) return;
  __shared__ int64_t lds[16];
  Op op;
  const int64_t nt = scan_live_tiles(n, ndev);
  int64_t per = (nt + 1023) / 1024;
  int64_t i0 = (int64_t)threadIdx.x * per;
  int64_t run = Op::ident;
  for (int64_t k = 0; k < per; ++k)
    if (i0 + k < nt) run = op(run, ts[i0 + k]);
  int64_t tot;
  int64_t pre = block_excl_scan<Op, 1024>(run, tot, lds);
  for (int64_t k = 0; k < per; ++k) {
    if (i0 + k < nt) {
      int64_t x = ts[i0 + k];
      ts[i0 + k] = pre;
      pre = op(pre, x);
    }
  }
  if (threadIdx.x == 0 && total) *total = tot;
}

__global__ void __launch_bounds__(BLK) k_scan_fix(int64_t* out, const int64_t* ts, int64_t n, const int64_t* ndev,
                                                  const int* cond) {
  if (cond && *cond == 0) return;
  const int64_t ntl = scan_live_tiles(n, ndev);
  for (int64_t t = blockIdx.x; t < ntl; t += gridDim.x) {
    const int64_t add = ts[t];
    const int64_t base = t * SCAN_TILE;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      int64_t i = base + k * BLK + threadIdx.x;
      if (i < n) out[i] += add;
    }
  }
}

// slot value = scan[gpool] - scan[scenario base]; as int rank or int64 offset
__global__ void __launch_bounds__(BLK) k_gather_rank(const int64_t* scan, const int* gpool, const int* scen,
                                                     const int* scen_base, int n, int* rank, int64_t* off,
                                                     const int* cond) {
  if (cond && *cond == 0) return;
  for (int j = blockIdx.x * BLK + threadIdx.x; j < n; j += gridDim.x * BLK) {
    int s = scen ? scen[j] : 0;
    int64_t v = scan[gpool[j]] - scan[scen_base[s]];
    if (rank) rank[j] = (int)v;
    if (off) off[j] = v;
  }
}

// total units of each scenario after a scan: scan[end] - scan[base]
__global__ void k_scen_totals(const int64_t* scan, const int64_t* total, const int* scen_base, int nscen,
                              int npool, int64_t* scen_units, const int* cond) {
  if (cond && *cond == 0) return;
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nscen) return;
  int e = scen_base[s + 1];
  int64_t end = e >= npool ? *total : scan[e];
  scen_units[s] = end - scan[scen_base[s]];
}

// One chunk of BLK slots per block, chunks mapped to the XCDs.  Blocks are
// dispatched round-robin over the 8 XCDs, so block b takes chunk (b % 8) *
// ceil(chunks / 8) + b / 8: each XCD sweeps one contiguous eighth of the
// slots, and neighbouring slots (neighbouring elements once sorted) share its
// L2.  The grid is 8 * ceil(chunks / 8) blocks; returns -1 past the last chunk.
__device__ __forceinline__ int64_t xcd_chunk(int n) {
  int64_t nch = ((int64_t)n + BLK - 1) / BLK, cpx = (nch + 7) / 8;
  int64_t chunk = (blockIdx.x % 8) * cpx + blockIdx.x / 8;
  return chunk < nch ? chunk : -1;
}

// Element re-sort: counting sort of the slots by a key that groups 2^SORT_SHIFT
// elements neighbouring along a Hilbert curve through the element centroids
// (active rows), then rows without element, then inactive rows.  The input is
// nearly sorted after the first sort, so equal keys come in long runs of
// consecutive slots.
constexpr int SORT_SHIFT = 2;

__device__ __forceinline__ int sort_nbins(int nt) { return (nt >> SORT_SHIFT) + 3; }

// batched pools: the kernel class of the slot's scenario is the major key, so
// each class occupies one contiguous slot range
__device__ __forceinline__ int sort_key(int t, int inactive, const int* elem_rank, int nt, int cls) {
  int nb = (nt >> SORT_SHIFT) + 1;
  return cls * (nb + 2) + (inactive ? nb + 1 : (t >= 0 ? elem_rank[t] >> SORT_SHIFT : nb));
}

// Block-local histogram: a block's SORT_ITEMS * BLK slots meet in an LDS hash
// table of their keys (open addressing, SORT_PROBES probes), so a bin that
// holds many of the block's slots takes one global atomic per block -- a point
// release puts a million slots in a handful of bins.  A key that finds no free
// entry goes to the global counters directly.
constexpr int SORT_ITEMS = 8;
constexpr int SORT_TABLE = 4096;  // entries, a power of two >= 2 * SORT_ITEMS * BLK
constexpr int SORT_PROBES = 8;

__device__ __forceinline__ int sort_table_slot(int* tkey, int k) {
  unsigned h = ((unsigned)k * 2654435761u) >> 20;  // 12 bits: SORT_TABLE entries
#pragma unroll 1
  for (int p = 0; p < SORT_PROBES; ++p) {
    int idx = (int)((h + (unsigned)p) & (SORT_TABLE - 1));
    int old = atomicCAS(&tkey[idx], -1, k);
    if (old == -1 || old == k) return idx;
  }
  return -1;
}

__device__ __forceinline__ void sort_table_clear(int* tkey, int* tcnt) {
  for (int i = threadIdx.x; i < SORT_TABLE; i += BLK) {
    tkey[i] = -1;
    tcnt[i] = 0;
  }
}

__global__ void __launch_bounds__(BLK) k_sort_hist(const int* tri, const int* inactive, const int* elem_rank,
                                                   const int* scen, const int* class_of_scen, int n, int nt,
                                                   int* cnt, int* key) {
  __shared__ int tkey[SORT_TABLE];
  __shared__ int tcnt[SORT_TABLE];
  sort_table_clear(tkey, tcnt);
  __syncthreads();
  int64_t b0 = (int64_t)blockIdx.x * BLK * SORT_ITEMS;
#pragma unroll
  for (int q = 0; q < SORT_ITEMS; ++q) {
    int64_t j = b0 + q * BLK + threadIdx.x;
    if (j >= n) continue;
    int cls = scen ? class_of_scen[scen[j]] : 0;
    int k = sort_key(tri[j], inactive[j], elem_rank, nt, cls);
    key[j] = k;
    int idx = sort_table_slot(tkey, k);
    if (idx >= 0) atomicAdd(&tcnt[idx], 1);
    else atomicAdd(&cnt[k], 1);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SORT_TABLE; i += BLK)
    if (tkey[i] >= 0) atomicAdd(&cnt[tkey[i]], tcnt[i]);
}

// Each block reserves one range per key in the global cursors and places its
// slots in it (the order of slots within a bin is arbitrary).
__global__ void __launch_bounds__(BLK) k_sort_place(const int* key, int n, int64_t* cursor, int* perm) {
  __shared__ int tkey[SORT_TABLE];
  __shared__ int tcnt[SORT_TABLE];
  __shared__ int tbase[SORT_TABLE];
  sort_table_clear(tkey, tcnt);
  __syncthreads();
  int64_t b0 = (int64_t)blockIdx.x * BLK * SORT_ITEMS;
  int slot[SORT_ITEMS], lrank[SORT_ITEMS];
#pragma unroll
  for (int q = 0; q < SORT_ITEMS; ++q) {
    int64_t j = b0 + q * BLK + threadIdx.x;
    slot[q] = -1;
    lrank[q] = 0;
    if (j >= n) continue;
    int k = key[j];
    int idx = sort_table_slot(tkey, k);
    if (idx >= 0) {
      slot[q] = idx;
      lrank[q] = atomicAdd(&tcnt[idx], 1);
    } else {
      perm[atomicAdd((unsigned long long*)&cursor[k], 1ull)] = (int)j;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SORT_TABLE; i += BLK)
    if (tkey[i] >= 0) tbase[i] = (int)atomicAdd((unsigned long long*)&cursor[tkey[i]], (unsigned long long)tcnt[i]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SORT_ITEMS; ++q) {
    int64_t j = b0 + q * BLK + threadIdx.x;
    if (slot[q] >= 0) perm[tbase[slot[q]] + lrank[q]] = (int)j;
  }
}

__global__ void __launch_bounds__(BLK) k_slot_of_pool(const int* gpool, int n, int* slot_of_pool) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j < n) slot_of_pool[gpool[j]] = j;
}

// end of step: advance each scenario's stream past its last draw, reset counters
// clock (graph replays): the device-side step counters, advanced here
// End of a step: each scenario's stream advances past the step's last draw.
// Its rows drew scen_units[s] units in pool order (the uniform count, or the
// general-mode scan total), so the position after the last draw is the
// position of unit scen_units[s]: the unit itself for rectangular normals,
// shifted by the slow Ziggurat normals' extra draws or the rejected polar
// attempts before it (the step's event tables, Draws::zig_position /
// polar_position).  No kernel reduces the stream end.
__global__ void k_finish(StreamDev* st, const JumpTable* jt, const int64_t* scen_units, int method, int nscen,
                         int* flag, int* n_deleted, long long* n_deleted_total, int step_id, int* step_done,
                         unsigned long long* err_step, unsigned long long* err, int* clock, int* strand_count,
                         int* unchecked) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nscen) {
    const int64_t u = scen_units[s];
    unsigned long long e = 0;
    if (u > 0) {
      Draws d;
      d.st = &st[s];
      e = (unsigned long long)(method == 1 ? u : (method == 2 ? d.zig_position(u) : d.polar_position(u)));
    }
    if (e) st[s].base = pcg_jump(st[s].base, e, jt[s]);
    if (err_step[s] < err[s]) err[s] = err_step[s];
    err_step[s] = ~0ull;
  }
  if (s == 0) {
    if (*n_deleted) flag[1] = 1;
    *n_deleted_total += *n_deleted;
    *n_deleted = 0;
    if (unchecked && flag[0]) *unchecked = 1;  // mispredicted in a step run without the rerun
    flag[0] = 0;
    int sid = clock ? clock[0] : step_id;
    *step_done = sid + 1;
    if (strand_count) strand_count[1] = strand_count[0];
    if (clock) {
      clock[0] = sid + 1;
      clock[1] += 1;
    }
  }
}

// reset the per-step outputs before a general-mode rerun of the step
__global__ void k_rerun_reset(unsigned long long* stream_end, unsigned long long* err, int* bnd_stats, int nstat,
                              int nscen, int* n_deleted, const int* cond, int* strand_count) {
  if (cond && *cond == 0) return;
  if (strand_count && blockIdx.x == 0 && threadIdx.x == 0) strand_count[0] = strand_count[1];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nscen) {
    stream_end[i] = 0;
    err[i] = ~0ull;
  }
  if (bnd_stats && i < nstat) bnd_stats[i] = 0;
  if (i == 0) *n_deleted = 0;
}

// a general-mode rerun starts from the step's input: undo its deletion marks
__global__ void __launch_bounds__(BLK) k_undelete(int* inactive, const int* del_step, int n, int step_id,
                                                  const int* cond, const int* clock) {
  if (cond && *cond == 0) return;
  if (clock) step_id = clock[0];
  for (int j = blockIdx.x * BLK + threadIdx.x; j < n; j += gridDim.x * BLK)
    if (inactive[j] == 2 && del_step[j] == step_id) inactive[j] = 0;
}

// Control sections (solver.pyx:540-549, control_shapes.pyx:56-85): a row that
// was active at the start of the step is counted once per section any of
// whose segments its step path [last position, position] meets.  A row deleted
// during the step has both positions zeroed by delete_single, so its path is
// the origin.  seg[q] = (x0, y0, x1, y1) of segment q, sec_of[q] its section.
__global__ void __launch_bounds__(BLK) k_control_sections(const double* ax, const double* ay, const double* bx,
                                                          const double* by, const int* inactive, const int* del_step,
                                                          int step_id, int n, const double4* seg, const int* sec_of,
                                                          int nseg, int* counts) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= n) return;
  int ina = inactive[j];
  bool deleted_now = ina == 2 && del_step[j] == step_id;
  if (ina != 0 && !deleted_now) return;
  double x0 = deleted_now ? 0. : ax[j], y0 = deleted_now ? 0. : ay[j];
  double x1 = deleted_now ? 0. : bx[j], y1 = deleted_now ? 0. : by[j];
  int last = -1;
  for (int q = 0; q < nseg; ++q) {
    int sc = sec_of[q];
    if (sc == last) continue;  // this section already counted the row
    double4 g = seg[q];
    if (seg_intersect(x0, y0, x1, y1, g.x, g.y, g.z, g.w)) {
      atomicAdd(&counts[sc], 1);
      last = sc;
    }
  }
}

// fill a device int array (graph-capturable replacement of a memset)
__global__ void __launch_bounds__(BLK) k_fill_int(int* p, int n, int v) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j < n) p[j] = v;
}

__global__ void k_set_flag(int* flag, int idx, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) flag[idx] = v;
}

// dst[q][j] = src[q][perm[j]] for every listed array, one pass
constexpr int PERM_MAX = 16;
struct PermArgs {
  const int* perm;
  int n, nd, ni;
  const double* dsrc[PERM_MAX];
  double* ddst[PERM_MAX];
  const int* isrc[PERM_MAX];
  int* idst[PERM_MAX];
};

// Destinations in XCD-mapped chunks: the slots of one contiguous eighth of
// the new order come, up to the local reshuffle since the last sort, from one
// eighth of the old order, so the source lines an XCD gathers stay in its own
// L2 instead of being fetched by all eight.
__global__ void __launch_bounds__(BLK) k_permute_all(PermArgs p) {
  const int64_t chunk = xcd_chunk(p.n);
  if (chunk < 0) return;
  const int64_t jj = chunk * BLK + threadIdx.x;
  if (jj >= p.n) return;
  const int j = (int)jj;
  int s = p.perm[j];
#pragma unroll
  for (int q = 0; q < PERM_MAX; ++q)
    if (q < p.nd) p.ddst[q][j] = p.dsrc[q][s];
#pragma unroll
  for (int q = 0; q < PERM_MAX; ++q)
    if (q < p.ni) p.idst[q][j] = p.isrc[q][s];
}

}  // namespace cyl
