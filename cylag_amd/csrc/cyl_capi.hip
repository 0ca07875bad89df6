// Host driver and C ABI of libcylag_hip.so (see include/cylag_hip.h).
//
// One context = one HIP stream on one GPU, holding the whole mesh, every field
// record and the particle pool in HBM.  cyl_step enqueues, per time step:
//   time interpolation -> (active-rank scan if rows were deleted) ->
//   random-stream event tables (Ziggurat / polar) -> fused step kernel ->
//   guarded general-mode rerun (count, pool-order scan, events, step) that
//   only does work when the fused kernel saw a draw count other than the
//   uniform prediction -> end-of-step stream advance.
// Nothing in a multi-step call waits on the host; results, statistics and the
// random state are read back once at the end.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cylag_hip.h"
#include "cyl_events.h"
#include "cyl_fast.h"
#include "cyl_fast3d.h"

using namespace cyl;

namespace {

// steps_since_sort value that makes the next step re-sort the slots
constexpr int SORT_NOW = 1 << 28;

struct Bufs {
  std::vector<void*> ptrs;
  void free_all() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
  // free one buffer of the set
  hipError_t release(void* p) {
    auto it = std::find(ptrs.begin(), ptrs.end(), p);
    if (it == ptrs.end()) return hipSuccess;
    ptrs.erase(it);
    return hipFree(p);
  }
};

}  // namespace

struct cyl_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int errcode = 0, errparticle = -1;
  std::string errmsg;

  // mesh
  MeshDev mesh{};
  bool has_mesh = false;
  Bufs mesh_bufs;
  // Hilbert rank of each element (re-sort key); node renumbering of the
  // packed records (new -> old, old -> new)
  int *d_elem_rank = nullptr, *d_node_old = nullptr, *d_node_new = nullptr;
  // packed records of the fused 2D LSM-1 kernel (built on first use)
  bool fast_ready = false;
  Bufs fast_bufs;
  FastMesh fm{};
  NodeF* d_nodes = nullptr;
  NodeL* d_nodes3 = nullptr;  // 3D: per node and sigma plane
  // packed node records are valid for the frozen fields of stamp pack_stamp,
  // with k-epsilon K on or off (packed_keps) and factor CMU / packed_sigc
  int64_t field_stamp = 0, pack_stamp = -1;
  int packed_keps = -1;
  double packed_sigc = 0.;

  // fields
  FieldsDev fields{};
  bool has_fields = false;
  Bufs field_bufs;
  std::vector<double> times;
  int ntimes = 0;
  int64_t fsize = 0;
  std::vector<int> fids_present;            // field ids with data
  std::vector<double*> frame_ptr;           // [fid*ntimes + rec]
  double* d_tmp[4 + MAX_OPT] = {};
  double* d_prev[4] = {};
  int record = 0, interp_done = 0, prev_valid = 0, frozen = 1;
  double interp_hist[2] = {0., 0.};         // times of the last two interpolations (most recent last)
  // graph replays of time-varying fields: the (record, weight) of each step of
  // the call, indexed by step id - interp_step0 (host-computed)
  InterpStep* d_interp_tab = nullptr;
  int64_t interp_tab_cap = 0;
  int interp_step0 = 0;
  int interp_count = 0;
  const double** d_ti_src = nullptr;        // [ntimes][nti] frame pointers of the interpolated fields
  double** d_ti_dst = nullptr;              // [2][nti] destination per buffer parity
  int nti = 0, tmp_parity = 0;
  std::vector<int> ti_fid;                  // field id of each interpolated field
  double* d_tmpbuf2[4] = {};                // second buffer of the main fields (previous values)

  // particles
  int n = 0, dim = 2;
  PoolDev pool[2]{};
  int cur = 0;
  Bufs part_bufs;
  int *d_gpool = nullptr, *d_rank = nullptr, *d_init_tri = nullptr, *d_del_step = nullptr;
  double2* d_xi = nullptr;            // the step's normals per active rank (pairs in 2D, XiRec3 in 3D)
  int64_t xi_cap = 0;
  int64_t* d_off = nullptr;
  double *d_diam = nullptr, *d_vol = nullptr, *d_depth = nullptr;
  ParticleConst pc{};
  bool pending_init = false;   // rows with tri == -2 exist
  bool have_last = false;
  // 2D LSM-1 in a plain context: the fused kernel leaves us unwritten, as
  // fluid_velocity_seen equals velocity for every 2D LSM-1 row after a step
  // (update_state_lsm1.pyx:169-172 and the other schemes; rebounds reflect no
  // LSM-1 velocity, compute_bc.pyx:195); downloads then derive it
  bool us_from_up = false;

  // scenarios (a single one for the plain solver; several after cyl_scenarios_set)
  bool batched = false;
  std::vector<cyl_params> scen_par;     // per scenario (batched)
  std::vector<int> scen_nsteps;         // steps each scenario runs (batched)
  int batch_steps = 0;                  // steps run since the scenarios were set
  ParamsDev* d_pars = nullptr;          // [nscen] device copy of the parameters
  FastScen* d_fsc = nullptr;            // [nscen] fused-kernel parameters
  int* d_scen_end = nullptr;            // [nscen] number of steps a scenario runs
  int* d_pregen = nullptr;              // [nscen] 1: scenario stepped by the fused kernel
  int64_t* d_gen_base = nullptr;        // [nscen + 1] normal-generator chunks
  int* d_class_of_scen = nullptr;       // [nscen] kernel class (physical slot group)
  struct KClass { bool fast; int model; int lo, hi; };
  std::vector<KClass> classes;          // batched: slot ranges per kernel class
  std::vector<int> class_of_scen_h;     // batched: class of each scenario
  std::vector<int> pregen_h;            // last uploaded d_pregen
  bool classes_fast_of_scen(int s) const {
    return !classes.empty() && s < (int)class_of_scen_h.size() && classes[class_of_scen_h[s]].fast;
  }
  int nscen = 1;
  std::vector<int> scen_base_h{0, 0};
  int* d_scen_base = nullptr;
  int* d_scen = nullptr;
  StreamDev* d_stream = nullptr;
  JumpTable* d_jt = nullptr;
  JumpTable8* d_jt8 = nullptr;
  int ncu = 256;
  unsigned long long *d_stream_end = nullptr, *d_err = nullptr, *d_err_step = nullptr;
  int64_t *d_scen_units = nullptr, *d_scen_active = nullptr;
  int* d_flag = nullptr;
  // bounds of the nodal k-epsilon diffusivity over every frame (k > 0 at every
  // node, or keps_ok false): K = (CMU/sigc) k^2 / max(eps, EPSILON) lies in
  // [fac * keps_lo, fac * keps_hi]
  bool keps_ok = false;
  double keps_lo = 0., keps_hi = 0.;
  bool skip_checked_rerun = true;  // tuning: leave out the rerun chain when draws are proven fixed
  int* d_n_deleted = nullptr;
  long long* d_n_deleted_total = nullptr;
  int* d_step_done = nullptr;
  std::vector<uint64_t> rng_state, rng_inc;
  bool rng_dirty = true;
  ZigTables* d_zt = nullptr;

  // pool-order scan scratch
  int* d_vals = nullptr;
  int64_t *d_scan = nullptr, *d_bsum = nullptr, *d_total = nullptr;

  // random-stream events
  int64_t ev_units_cap = 0, nchunks = 0, ev_cap = 0;
  Bufs ev_bufs;
  int64_t* d_chunk_base = nullptr;
  int64_t nspan = 0;                               // spans of BLK chunks
  int *d_chunk_loc = nullptr, *d_span_cnt = nullptr;
  int64_t *d_span_off = nullptr, *d_span_max = nullptr, *d_span_carry = nullptr;
  int64_t *d_span_pk = nullptr, *d_span_pkoff = nullptr, *d_scen_c = nullptr, *d_scen_c0 = nullptr;
  int64_t* d_lb_base = nullptr;                    // the step's live bucket entries per scenario
  int64_t *d_reg_pos = nullptr, *d_reg_len = nullptr, *d_reg_key = nullptr;
  double* d_reg_val = nullptr;                     // span regions of the candidate records
  int64_t *d_ev_pos = nullptr, *d_ev_len = nullptr, *d_reach = nullptr;
  int64_t *d_E = nullptr, *d_key = nullptr, *d_xtra = nullptr;
  double* d_ev_val = nullptr;                      // compacted slow-normal values
  double* d_cand_val = nullptr;                    // value per candidate
  int64_t *d_ev_begin = nullptr, *d_nev = nullptr;
  int64_t* d_chunk_live = nullptr;                 // end of the step's live chunks
  int* d_bucket = nullptr;          // bucket index of the event keys, all scenarios
  int64_t* d_bucket_base = nullptr;
  int64_t nbucket = 0;
  int bshift = 6;

  // element re-sort
  int sort_interval = 8, steps_since_sort = 0;  // SORT_NOW: re-sort at the next step
  int blocks_per_cu = 0;  // persistent grids: resident blocks per CU (0 = as many as fit)
  int inline_normals = 0; // 2D fused kernels' normals: 0 pool-order generator, 1 inline sampling, 2 slot-order generator
  int use_graph = 0;      // replay multi-step loops from a captured HIP graph
  int* d_clock = nullptr; // graph replays: {step id, batch step} on the device
  int last_graph_steps = 0;
  // control sections: segments (sections ordered), per-step counts of the last call
  int nsec = 0, nsecseg = 0;
  double4* d_sec_seg = nullptr;
  int* d_sec_of = nullptr;
  int* d_sec_counts = nullptr;
  size_t sec_cap = 0;
  std::vector<int> last_sec_counts;
  // stranded-particle records of the last call (stranding_output)
  double* d_strand_xyz = nullptr;
  int2* d_strand_key = nullptr;
  int* d_strand_count = nullptr;
  int strand_cap = 0;
  std::vector<double> last_strand;  // 3 per record, in (step, pool index) order
  std::vector<double> last_deposit; // deposits on the bed: 3 per record, same order
  std::vector<int> last_deposit_step;  // step of the call (0-based) of each deposit
  int *d_perm = nullptr, *d_skey = nullptr, *d_scnt = nullptr;
  int64_t* d_cursor = nullptr;
  int64_t* d_sort_tiles = nullptr;
  int sort_nt = 0, sort_ncls = 0;
  int* sp_scen = nullptr;
  // spare per-slot buffers the sort gathers into (pointers swapped after)
  int *sp_gpool = nullptr, *sp_init_tri = nullptr, *sp_del_step = nullptr, *sp_inactive = nullptr;
  int* sp_rank = nullptr;
  double *sp_diam = nullptr, *sp_vol = nullptr, *sp_depth = nullptr;
  void* d_tmpbuf = nullptr;

  int step_counter = 0;
  int* d_bnd = nullptr;
  size_t bnd_cap = 0;
  int nopen = 0;

  double last_kernel_ms = 0., last_total_ms = 0., last_step_kernel_ms = 0., last_prep_ms = 0.;
  int64_t* d_pos0 = nullptr;     // per slot: first-draw stream position (fused 2D step)
  uint64_t* d_state0 = nullptr;  // per slot: generator state there
  int start_cap = 0;
  int last_step_launches = 0;
  std::vector<hipEvent_t> kev;              // start/stop pairs around the fused step kernel
  int last_launches = 0;
  int last_fast_launches = 0, last_generic_launches = 0;  // main-kernel launches that did the step
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  int set_err(int code, int particle, const std::string& msg) {
    errcode = code;
    errparticle = particle;
    errmsg = msg;
    return code;
  }
};

#define HIPCHK(x)                                                                             \
  do {                                                                                        \
    hipError_t e_ = (x);                                                                      \
    if (e_ != hipSuccess)                                                                     \
      return ctx->set_err(CYL_E_HIP, -1, std::string(#x) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <class T>
static hipError_t dalloc(Bufs& b, T** p, size_t count) {
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e == hipSuccess) b.ptrs.push_back(q);
  *p = (T*)q;
  return e;
}

static inline unsigned nblk(int64_t n, int b = BLK) { return (unsigned)((n + b - 1) / b); }

// distance along the Hilbert curve of order 16 of the cell (x, y), x, y < 2^16
// (the classic rotate-and-accumulate conversion)
static uint64_t hilbert_index(uint32_t x, uint32_t y) {
  const uint32_t n = 1u << 16;
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) {
        x = n - 1 - x;
        y = n - 1 - y;
      }
      std::swap(x, y);
    }
  }
  return d;
}

// ---------------------------------------------------------------------------
// Staging of cyl_batch_collect_noalloc: one file-scope device array per GPU
// (the code object's own copy on each device), sized for CYL_NOALLOC_MAX_ROWS
// rows of a 3D scenario.  Regions, each CYL_NOALLOC_MAX_ROWS entries long:
// position, velocity, seen velocity (3 doubles per row each), element id,
// layer id, inactive flag (int), deletion key (uint64).  Host code never
// passes `scratch` itself to a kernel or copy: it uses the device address
// hipGetSymbolAddress returns for the current device, fetched once per device
// when a context is created there and kept in noalloc_ptr[].
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_ROWS = CYL_NOALLOC_MAX_ROWS;
constexpr size_t NOALLOC_BYTES = NOALLOC_ROWS * (9 * sizeof(double) + 3 * sizeof(int) + sizeof(uint64_t));
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;                     // guards noalloc_ptr and serialises the noalloc calls
static void* noalloc_ptr[NOALLOC_MAX_DEVICES];    // device address of `scratch` on each device
static uint64_t noalloc_keys[NOALLOC_ROWS];       // host side of the deletion keys

// With `device` current: fetch (once) the device address of this device's
// copy of `scratch`.
static hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (noalloc_ptr[device]) return hipSuccess;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  noalloc_ptr[device] = p;
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int cyl_ctx_create(int device, int flags, cyl_ctx** out) {
  (void)flags;
  if (!out) return CYL_E_ARG;
  cyl_ctx* ctx = new cyl_ctx();
  ctx->device = device;
  *out = ctx;
  HIPCHK(hipSetDevice(device));
  HIPCHK(noalloc_scratch_init(device));
  HIPCHK(hipStreamCreate(&ctx->stream));
  HIPCHK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device));
  HIPCHK(hipEventCreate(&ctx->ev0));
  HIPCHK(hipEventCreate(&ctx->ev1));
  // Ziggurat tables with the host libm (random_utils.pyx:118-146)
  ZigTables z;
  {
    double m1 = 2147483648.0, dn = 3.442619855899, tn = dn, vn = 9.91256303526217e-3;
    double q = vn / std::exp(-0.5 * dn * dn);
    z.kn[0] = (uint32_t)((dn / q) * m1);
    z.kn[1] = 0;
    z.wn[0] = q / m1;
    z.wn[127] = dn / m1;
    z.fn[0] = 1.0;
    z.fn[127] = std::exp(-0.5 * dn * dn);
    for (int i = 126; i > 0; --i) {
      dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
      z.kn[i + 1] = (uint32_t)((dn / tn) * m1);
      tn = dn;
      z.fn[i] = std::exp(-0.5 * dn * dn);
      z.wn[i] = dn / m1;
    }
  }
  HIPCHK(hipMalloc(&ctx->d_zt, sizeof(ZigTables)));
  HIPCHK(hipMemcpy(ctx->d_zt, &z, sizeof(ZigTables), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&ctx->d_flag, 8 * sizeof(int)));
  HIPCHK(hipMemset(ctx->d_flag, 0, 8 * sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_n_deleted, sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_n_deleted_total, sizeof(long long)));
  HIPCHK(hipMalloc(&ctx->d_step_done, sizeof(int)));
  HIPCHK(hipMalloc(&ctx->d_clock, 2 * sizeof(int)));
  HIPCHK(hipMemset(ctx->d_n_deleted, 0, sizeof(int)));
  HIPCHK(hipMemset(ctx->d_n_deleted_total, 0, sizeof(long long)));
  ctx->rng_state.assign(1, 0);
  ctx->rng_inc.assign(1, 1);
  // module default seed, cylag/lsm/__init__.py:6-10
  Pcg g;
  pcg_seed(g, 42, 0xda3e06d11bb2b203ULL);
  ctx->rng_state[0] = g.state;
  ctx->rng_inc[0] = g.inc;
  return CYL_OK;
}

extern "C" int cyl_ctx_destroy(cyl_ctx* ctx) {
  if (!ctx) return CYL_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->pool[1].inactive = nullptr;
  ctx->mesh_bufs.free_all();
  ctx->fast_bufs.free_all();
  ctx->fast_ready = false;
  if (ctx->d_pos0) (void)hipFree(ctx->d_pos0);
  if (ctx->d_state0) (void)hipFree(ctx->d_state0);
  if (ctx->d_interp_tab) (void)hipFree(ctx->d_interp_tab);
  ctx->field_bufs.free_all();
  ctx->part_bufs.free_all();
  ctx->ev_bufs.free_all();
  for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
  void* singles[] = {ctx->d_zt,        ctx->d_flag, ctx->d_n_deleted, ctx->d_n_deleted_total, ctx->d_step_done,
                     ctx->d_bnd,       ctx->d_clock, ctx->d_sec_seg, ctx->d_sec_of,       ctx->d_sec_counts,
                     ctx->d_strand_xyz, ctx->d_strand_key, ctx->d_strand_count};
  for (void* p : singles)
    if (p) (void)hipFree(p);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CYL_OK;
}

extern "C" int cyl_last_error(cyl_ctx* ctx, int* code, int* particle, char* msg, size_t n) {
  if (!ctx) return CYL_E_ARG;
  if (code) *code = ctx->errcode;
  if (particle) *particle = ctx->errparticle;
  if (msg && n) {
    std::strncpy(msg, ctx->errmsg.c_str(), n - 1);
    msg[n - 1] = 0;
  }
  return CYL_OK;
}

extern "C" int cyl_device_info(cyl_ctx* ctx, char* name, size_t n, size_t* free_b, size_t* total_b) {
  if (!ctx) return CYL_E_ARG;
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
  if (name && n) {
    std::snprintf(name, n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  }
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  if (free_b) *free_b = f;
  if (total_b) *total_b = t;
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------------
extern "C" int cyl_mesh_upload(cyl_ctx* ctx, int nv, const double* x, const double* y, int nt, const int* tri,
                               int ne, const int* edges, const int* edgetri, const int* triedge,
                               const int* edge_labels, const double* det, const double* gradx,
                               const double* grady, int nopen) {
  if (!ctx || nv <= 0 || nt <= 0 || ne <= 0) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->mesh_bufs.free_all();
  ctx->fast_bufs.free_all();
  ctx->fast_ready = false;
  MeshDev m{};
  m.nv = nv;
  m.nt = nt;
  m.ne = ne;
  m.nopen = nopen;
  ctx->nopen = nopen;
  auto up = [&](auto** dst, const auto* src, size_t cnt) -> hipError_t {
    hipError_t e = dalloc(ctx->mesh_bufs, dst, cnt);
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, src, cnt * sizeof(**dst), hipMemcpyHostToDevice);
  };
  double *dx, *dy, *ddet, *dgx, *dgy, *dcen;
  int *dtri, *ded, *det2, *dte, *dlab;
  HIPCHK(up(&dx, x, nv));
  HIPCHK(up(&dy, y, nv));
  HIPCHK(up(&dtri, tri, (size_t)nt * 3));
  HIPCHK(up(&ded, edges, (size_t)ne * 2));
  HIPCHK(up(&det2, edgetri, (size_t)ne * 2));
  HIPCHK(up(&dte, triedge, (size_t)nt * 3));
  HIPCHK(up(&dlab, edge_labels, ne));
  HIPCHK(up(&ddet, det, (size_t)nt * 3));
  HIPCHK(up(&dgx, gradx, (size_t)nt * 3));
  HIPCHK(up(&dgy, grady, (size_t)nt * 3));
  // centroids (triangular_mesh.pyx:621-626) and the bucket grid for fallback search
  std::vector<double> cen((size_t)nt * 2);
  double xmin = x[0], xmax = x[0], ymin = y[0], ymax = y[0];
  for (int i = 0; i < nv; ++i) {
    xmin = std::min(xmin, x[i]);
    xmax = std::max(xmax, x[i]);
    ymin = std::min(ymin, y[i]);
    ymax = std::max(ymax, y[i]);
  }
  for (int k = 0; k < nt; ++k) {
    int a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
    cen[2 * k] = (x[a] + x[b] + x[c]) / 3.0;
    cen[2 * k + 1] = (y[a] + y[b] + y[c]) / 3.0;
  }
  HIPCHK(up(&dcen, cen.data(), cen.size()));
  double w = std::max(xmax - xmin, 1e-300), h = std::max(ymax - ymin, 1e-300);
  // Hilbert-curve ranks of the element centroids and of the nodes: the
  // element re-sort key, and the node order of the packed node records
  {
    double sc = 65535.0 / std::max(w, h);
    auto hkey = [&](double px, double py) {
      double fx = std::floor((px - xmin) * sc), fy = std::floor((py - ymin) * sc);
      uint32_t hx = (uint32_t)std::min(65535.0, std::max(0.0, fx));
      uint32_t hy = (uint32_t)std::min(65535.0, std::max(0.0, fy));
      return hilbert_index(hx, hy);
    };
    auto ranks = [&](int cnt, auto&& coord, std::vector<int>& order) {
      std::vector<uint64_t> key(cnt);
      for (int i = 0; i < cnt; ++i) {
        double px, py;
        coord(i, px, py);
        key[i] = hkey(px, py);
      }
      order.resize(cnt);
      for (int i = 0; i < cnt; ++i) order[i] = i;
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    };
    std::vector<int> eord, nord;
    ranks(nt, [&](int k, double& px, double& py) { px = cen[2 * k]; py = cen[2 * k + 1]; }, eord);
    ranks(nv, [&](int i, double& px, double& py) { px = x[i]; py = y[i]; }, nord);
    std::vector<int> erank(nt), nnew(nv);
    for (int i = 0; i < nt; ++i) erank[eord[i]] = i;
    for (int i = 0; i < nv; ++i) nnew[nord[i]] = i;
    HIPCHK(up(&ctx->d_elem_rank, erank.data(), erank.size()));
    HIPCHK(up(&ctx->d_node_new, nnew.data(), nnew.size()));
    HIPCHK(up(&ctx->d_node_old, nord.data(), nord.size()));
  }
  double cell = std::sqrt(w * h / nt) * 1.5;
  if (!(cell > 0.)) cell = std::max(w, h);
  int gnx = (int)std::min<double>(std::max(1.0, std::ceil(w / cell)), 1 << 14);
  int gny = (int)std::min<double>(std::max(1.0, std::ceil(h / cell)), 1 << 14);
  double inv = 1.0 / cell;
  auto cidx = [&](double v, double v0, int nmax) {
    double f = std::floor((v - v0) * inv);
    if (f < 0.) f = 0.;
    if (f > nmax - 1) f = nmax - 1;
    return (int)f;
  };
  std::vector<int> start((size_t)gnx * gny + 1, 0);
  for (int pass = 0; pass < 2; ++pass) {
    std::vector<int> fill;
    if (pass == 1) {
      for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
      fill.assign(start.begin(), start.end() - 1);
    }
    static std::vector<int> items;
    if (pass == 1) items.assign(start.back(), 0);
    for (int k = 0; k < nt; ++k) {
      int a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
      double bx0 = std::min(x[a], std::min(x[b], x[c])), bx1 = std::max(x[a], std::max(x[b], x[c]));
      double by0 = std::min(y[a], std::min(y[b], y[c])), by1 = std::max(y[a], std::max(y[b], y[c]));
      int ix0 = cidx(bx0, xmin, gnx), ix1 = cidx(bx1, xmin, gnx);
      int iy0 = cidx(by0, ymin, gny), iy1 = cidx(by1, ymin, gny);
      for (int iy = iy0; iy <= iy1; ++iy)
        for (int ix = ix0; ix <= ix1; ++ix) {
          size_t cc = (size_t)iy * gnx + ix;
          if (pass == 0) start[cc + 1]++;
          else items[fill[cc]++] = k;
        }
    }
    if (pass == 1) {
      int *dst, *dit;
      HIPCHK(up(&dst, start.data(), start.size()));
      HIPCHK(up(&dit, items.data(), std::max<size_t>(items.size(), 1)));
      m.g_start = dst;
      m.g_items = dit;
      items.clear();
      items.shrink_to_fit();
    }
  }
  m.x = dx; m.y = dy; m.tri = dtri; m.edges = ded; m.edgetri = det2; m.triedge = dte; m.edge_labels = dlab;
  m.det = ddet; m.gx = dgx; m.gy = dgy; m.centroid = dcen;
  m.g_x0 = xmin; m.g_y0 = ymin; m.g_inv = inv; m.g_nx = gnx; m.g_ny = gny;
  ctx->mesh = m;
  ctx->has_mesh = true;
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// fields
// ---------------------------------------------------------------------------
extern "C" int cyl_fields_upload(cyl_ctx* ctx, int dim, int nlayers, int npoints, int ntimes, const double* times,
                                 const double* elev, const double* vx, const double* vy, const double* vz, int nopt,
                                 const double* opt, const int* fids, int friction_model, double friction_coef) {
  if (ctx) ++ctx->field_stamp;
  if (!ctx || (dim != 2 && dim != 3) || nlayers < 1 || npoints < 1 || ntimes < 1 || nopt < 0 || nopt > MAX_OPT)
    return CYL_E_ARG;
  if (!elev || !vx || !vy || (dim == 3 && !vz) || (nopt > 0 && !opt)) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->field_bufs.free_all();
  // the packed records depend on the fields' planes: rebuilt on the next step
  ctx->fast_bufs.free_all();
  ctx->fast_ready = false;
  ctx->d_nodes3 = nullptr;
  int64_t size = (int64_t)npoints * nlayers;
  ctx->fsize = size;
  ctx->ntimes = ntimes;
  ctx->times.assign(times, times + ntimes);
  int nf = 4 + nopt;
  ctx->frame_ptr.assign((size_t)nf * ntimes, nullptr);
  ctx->fids_present.clear();
  const double* src[4 + MAX_OPT] = {elev, vx, vy, vz};
  for (int q = 0; q < nopt; ++q) src[4 + q] = opt + (size_t)q * ntimes * size;
  FieldsDev f{};
  f.dim = dim;
  f.nlayers = nlayers;
  f.npoints = npoints;
  f.nopt = nopt;
  f.fid_zs = fids[0]; f.fid_ux = fids[1]; f.fid_uy = fids[2]; f.fid_uz = fids[3];
  f.fid_turbeng = fids[4]; f.fid_dissip = fids[5]; f.fid_zb = fids[6];
  f.friction_model = friction_model;
  f.friction_coefficient = friction_coef;
  for (int fid = 0; fid < nf; ++fid) {
    if (!src[fid]) continue;
    ctx->fids_present.push_back(fid);
    double* d;
    HIPCHK(dalloc(ctx->field_bufs, &d, (size_t)ntimes * size));
    HIPCHK(hipMemcpy(d, src[fid], (size_t)ntimes * size * sizeof(double), hipMemcpyHostToDevice));
    for (int r = 0; r < ntimes; ++r) ctx->frame_ptr[(size_t)fid * ntimes + r] = d + (size_t)r * size;
    double* t;
    HIPCHK(dalloc(ctx->field_bufs, &t, size));
    // _init_tmp: record 0 (eulerian_field_set.pyx:743-776)
    HIPCHK(hipMemcpy(t, d, size * sizeof(double), hipMemcpyDeviceToDevice));
    ctx->d_tmp[fid] = t;
    f.tmp[fid] = t;
    if (fid < 4) {
      double* pv;
      HIPCHK(dalloc(ctx->field_bufs, &pv, size));
      HIPCHK(hipMemcpy(pv, t, size * sizeof(double), hipMemcpyDeviceToDevice));
      ctx->d_tmpbuf2[fid] = pv;
      ctx->d_prev[fid] = pv;
      f.prev[fid] = pv;
    }
  }
  // k-epsilon bounds: the range of k^2 / max(eps, EPSILON) over every node
  // and frame (nodal K without its constant factor), with k strictly positive
  // and finite everywhere
  ctx->keps_ok = false;
  if (f.fid_turbeng >= 0 && f.fid_dissip >= 0 && src[f.fid_turbeng] && src[f.fid_dissip]) {
    const double* kt = src[f.fid_turbeng];
    const double* ep = src[f.fid_dissip];
    double lo = HUGE_VAL, hi = 0.;
    bool ok = true;
    for (int64_t i = 0; i < (int64_t)ntimes * size && ok; ++i) {
      ok = kt[i] > 0. && std::isfinite(kt[i]) && std::isfinite(ep[i]);
      double r = kt[i] * kt[i] / std::max(ep[i], EPS);
      lo = std::min(lo, r);
      hi = std::max(hi, r);
    }
    if (ok && std::isfinite(hi)) {
      ctx->keps_ok = true;
      ctx->keps_lo = lo;
      ctx->keps_hi = hi;
    }
  }
  // time-interpolation pointer tables: fields 0..dim (main) and every optional field
  std::vector<int> ti;
  int ntmp = dim == 2 ? 3 : 4;
  for (int fid : ctx->fids_present)
    if (!(fid < 4 && fid >= ntmp)) ti.push_back(fid);
  int nti = (int)ti.size();
  ctx->nti = nti;
  ctx->ti_fid = ti;
  std::vector<const double*> srct((size_t)ntimes * nti);
  for (int r = 0; r < ntimes; ++r)
    for (int q = 0; q < nti; ++q) srct[(size_t)r * nti + q] = ctx->frame_ptr[(size_t)ti[q] * ntimes + r];
  std::vector<double*> dstt((size_t)2 * nti);
  for (int q = 0; q < nti; ++q) {
    int fid = ti[q];
    dstt[q] = ctx->d_tmp[fid];
    dstt[nti + q] = fid < 4 ? ctx->d_tmpbuf2[fid] : ctx->d_tmp[fid];
  }
  HIPCHK(dalloc(ctx->field_bufs, &ctx->d_ti_src, srct.size()));
  HIPCHK(hipMemcpy(ctx->d_ti_src, srct.data(), srct.size() * sizeof(void*), hipMemcpyHostToDevice));
  HIPCHK(dalloc(ctx->field_bufs, &ctx->d_ti_dst, dstt.size()));
  HIPCHK(hipMemcpy(ctx->d_ti_dst, dstt.data(), dstt.size() * sizeof(void*), hipMemcpyHostToDevice));
  ctx->tmp_parity = 0;
  f.frozen = ntimes == 1;
  ctx->frozen = f.frozen;
  ctx->record = 0;
  ctx->interp_done = 0;
  ctx->prev_valid = 0;
  ctx->fields = f;
  ctx->has_fields = true;
  return CYL_OK;
}

static int lower_index(const std::vector<double>& a, double v, int last) {
  // zlocalizer.pyx:125-151
  int n = (int)a.size();
  for (int i = last; i < n - 1; ++i)
    if (v >= a[i] && v < a[i + 1]) return i;
  if (n > 1 && v >= a[n - 1]) return n - 2;
  return 0;
}

extern "C" int cyl_fields_reset(cyl_ctx* ctx, int frozen, double t0) {
  if (!ctx || !ctx->has_fields) return CYL_E_STATE;
  ++ctx->field_stamp;
  ctx->frozen = frozen ? 1 : 0;
  ctx->fields.frozen = ctx->frozen;
  if (!ctx->frozen) ctx->record = lower_index(ctx->times, t0, 0);
  ctx->interp_count = 0;
  ctx->interp_done = 0;
  ctx->prev_valid = 0;
  ctx->fields.prev_valid = 0;
  return CYL_OK;
}

extern "C" int cyl_deleted_count(cyl_ctx* ctx, long long* n) {
  if (!ctx || !n) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(n, ctx->d_n_deleted_total, sizeof(long long), hipMemcpyDeviceToHost));
  return CYL_OK;
}

// enqueue _time_interpolation(time) (eulerian_field_set.pyx:866-938).  The
// main fields are double-buffered: the previous interpolation stays in the
// other buffer, which is what the reference copies into _tmp_*0.
// Host state of one time interpolation at `time` (record, weight, history,
// buffer parity); returns CYL_OK or the reference's error.
static int time_interp_state(cyl_ctx* ctx, double time, double& tf) {
  ctx->prev_valid = ctx->interp_done;
  ctx->interp_done = 1;
  ctx->interp_hist[0] = ctx->interp_hist[1];
  ctx->interp_hist[1] = time;
  ++ctx->interp_count;
  ctx->record = lower_index(ctx->times, time, ctx->record);
  int rec = ctx->record;
  if (rec + 1 >= ctx->ntimes) return ctx->set_err(CYL_E_STEP, -1, "time interpolation beyond the last record");
  double dtt = ctx->times[rec] - ctx->times[rec + 1];
  if (std::fabs(dtt) < EPS)
    return ctx->set_err(CYL_E_STEP, -1, "Time interpolation failed: duplicate or too-close time values detected");
  tf = (time - ctx->times[rec + 1]) / dtt;
  ctx->tmp_parity = 1 - ctx->tmp_parity;  // the new values go into the buffer that held the older ones
  for (int fid = 0; fid < 4; ++fid) {
    if (!ctx->d_tmp[fid] || !ctx->d_tmpbuf2[fid]) continue;
    double* cur = ctx->tmp_parity == 0 ? ctx->d_tmp[fid] : ctx->d_tmpbuf2[fid];
    double* old = ctx->tmp_parity == 0 ? ctx->d_tmpbuf2[fid] : ctx->d_tmp[fid];
    ctx->fields.tmp[fid] = cur;
    ctx->fields.prev[fid] = old;
  }
  ctx->fields.prev_valid = ctx->prev_valid;
  return CYL_OK;
}

// enqueue _time_interpolation(time) (eulerian_field_set.pyx:866-938).  The
// main fields are double-buffered: the previous interpolation stays in the
// other buffer, which is what the reference copies into _tmp_*0.  clock: the
// step is being captured into a graph; the record and weight then come from
// the call's table through the device step clock.
// pack: also write the 2D fused kernels' node records from the interpolated
// values (k_interp_pack), fac = C_mu / sigma_c, keps: with the k-epsilon K.
struct PackReq {
  bool on = false, keps = false;
  double fac = 0.;
};

static int enqueue_time_interp(cyl_ctx* ctx, double time, int& launches, const int* clock = nullptr,
                               const PackReq& pack = PackReq()) {
  if (ctx->frozen) return CYL_OK;
  double tf = 0.;
  int e = time_interp_state(ctx, time, tf);
  if (e) return e;
  const int rec = ctx->record, par = ctx->tmp_parity;
  if (pack.on) {
    const FieldsDev& fd = ctx->fields;
    auto qof = [&](int fid) {
      for (int q = 0; q < ctx->nti; ++q)
        if (ctx->ti_fid[q] == fid) return q;
      return -1;
    };
    InterpPack ip{};
    ip.frames = ctx->d_ti_src;
    ip.out = ctx->d_ti_dst + (size_t)par * ctx->nti;
    ip.nf = ctx->nti;
    ip.n = fd.npoints;
    ip.rec = rec;
    ip.tf = tf;
    ip.tab = ctx->d_interp_tab;
    ip.clock = clock;
    ip.step0 = ctx->interp_step0;
    ip.q_u = qof(fd.fid_ux);
    ip.q_v = qof(fd.fid_uy);
    ip.q_k = pack.keps ? qof(fd.fid_turbeng) : -1;
    ip.q_e = pack.keps ? qof(fd.fid_dissip) : -1;
    ip.fac = pack.fac;
    ip.nnew = ctx->d_node_new;
    ip.rec_out = ctx->d_nodes;
    if (ip.q_u < 0 || ip.q_v < 0 || (pack.keps && (ip.q_k < 0 || ip.q_e < 0)))
      return ctx->set_err(CYL_E_STATE, -1, "node records: a field of the fused kernel is not interpolated");
    if (clock) k_interp_pack<true><<<nblk(fd.npoints), BLK, 0, ctx->stream>>>(ip);
    else k_interp_pack<false><<<nblk(fd.npoints), BLK, 0, ctx->stream>>>(ip);
    ++launches;
    return CYL_OK;
  }
  dim3 g(nblk(ctx->fsize), ctx->nti);
  if (clock)
    k_time_interp_tab<<<g, BLK, 0, ctx->stream>>>(ctx->d_ti_src, ctx->d_ti_dst + (size_t)par * ctx->nti, ctx->nti,
                                                 ctx->fsize, ctx->d_interp_tab, clock, ctx->interp_step0);
  else
    k_time_interp<<<g, BLK, 0, ctx->stream>>>(ctx->d_ti_src + (size_t)rec * ctx->nti,
                                             ctx->d_ti_src + (size_t)(rec + 1) * ctx->nti,
                                             ctx->d_ti_dst + (size_t)par * ctx->nti, ctx->nti, ctx->fsize, tf);
  ++launches;
  return CYL_OK;
}

extern "C" int cyl_fields_interpolate(cyl_ctx* ctx, double time) {
  if (!ctx || !ctx->has_fields) return CYL_E_STATE;
  HIPCHK(hipSetDevice(ctx->device));
  int l = 0;
  int e = enqueue_time_interp(ctx, time, l);
  if (e) return e;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CYL_OK;
}

extern "C" int cyl_fields_download_tmp(cyl_ctx* ctx, int fid, double* out) {
  if (!ctx || !ctx->has_fields || fid < 0 || fid >= 4 + MAX_OPT || !ctx->fields.tmp[fid]) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(out, ctx->fields.tmp[fid], ctx->fsize * sizeof(double), hipMemcpyDeviceToHost));
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// particles
// ---------------------------------------------------------------------------
static int alloc_pool(cyl_ctx* ctx, PoolDev& P) {
  Bufs& b = ctx->part_bufs;
  int n = ctx->n;
  HIPCHK(dalloc(b, &P.x, n));
  HIPCHK(dalloc(b, &P.y, n));
  HIPCHK(dalloc(b, &P.z, ctx->dim == 3 ? n : 1));
  for (int c = 0; c < 3; ++c) {
    HIPCHK(dalloc(b, &P.up[c], c < ctx->dim ? n : 1));
    HIPCHK(dalloc(b, &P.us[c], c < ctx->dim ? n : 1));
  }
  HIPCHK(dalloc(b, &P.tri, n));
  HIPCHK(dalloc(b, &P.ll, n));
  if (&P == &ctx->pool[0]) HIPCHK(dalloc(b, &P.inactive, n));
  return CYL_OK;
}

static int ensure_scen_tables(cyl_ctx* ctx) {
  Bufs& b = ctx->part_bufs;
  int ns = ctx->nscen;
  HIPCHK(dalloc(b, &ctx->d_scen_base, ns + 1));
  HIPCHK(hipMemcpy(ctx->d_scen_base, ctx->scen_base_h.data(), (ns + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(dalloc(b, &ctx->d_stream, ns));
  HIPCHK(dalloc(b, &ctx->d_jt, ns));
  HIPCHK(dalloc(b, &ctx->d_jt8, ns));
  HIPCHK(dalloc(b, &ctx->d_stream_end, ns));
  HIPCHK(dalloc(b, &ctx->d_err, ns));
  HIPCHK(dalloc(b, &ctx->d_err_step, ns));
  HIPCHK(hipMemset(ctx->d_err_step, 0xff, ns * sizeof(unsigned long long)));
  HIPCHK(dalloc(b, &ctx->d_scen_units, ns));
  HIPCHK(dalloc(b, &ctx->d_scen_active, ns));
  HIPCHK(dalloc(b, &ctx->d_pars, ns));
  HIPCHK(dalloc(b, &ctx->d_fsc, ns));
  HIPCHK(dalloc(b, &ctx->d_scen_end, ns));
  HIPCHK(dalloc(b, &ctx->d_pregen, ns));
  HIPCHK(dalloc(b, &ctx->d_class_of_scen, ns));
  HIPCHK(hipMemset(ctx->d_class_of_scen, 0, ns * sizeof(int)));
  HIPCHK(dalloc(b, &ctx->d_gen_base, ns + 1));
  HIPCHK(hipMemset(ctx->d_stream_end, 0, ns * sizeof(unsigned long long)));
  HIPCHK(hipMemset(ctx->d_err, 0xff, ns * sizeof(unsigned long long)));
  std::vector<StreamDev> st(ns);
  std::memset(st.data(), 0, ns * sizeof(StreamDev));
  HIPCHK(hipMemcpy(ctx->d_stream, st.data(), ns * sizeof(StreamDev), hipMemcpyHostToDevice));
  ctx->rng_dirty = true;
  return CYL_OK;
}

// Every per-slot buffer of a pool of n slots (both state copies, slot
// tables, scans, sort), in ctx->part_bufs; nscen_pad scenarios of padding in
// the normals buffer.
static int alloc_slot_bufs(cyl_ctx* ctx, int n, int nscen_pad) {
  ctx->n = n;
  int dim = ctx->dim;
  for (int b = 0; b < 2; ++b) {
    int e = alloc_pool(ctx, ctx->pool[b]);
    if (e) return e;
  }
  // the pool flags are per slot, not double-buffered: a row's flag is only
  // changed by its own thread (deletion), in place
  ctx->pool[1].inactive = ctx->pool[0].inactive;
  Bufs& bb = ctx->part_bufs;
  HIPCHK(dalloc(bb, &ctx->d_gpool, n));
  HIPCHK(dalloc(bb, &ctx->d_rank, n));
  // normals in tiles of 64 chunks x GEN_PER: chunks <= n/GEN_PER + 1 per scenario,
  // rounded up to a whole tile; an entry is a pair (16 B) in 2D, a 32 B
  // record of three values in 3D
  ctx->xi_cap = (int64_t)n + GEN_PER * (int64_t)nscen_pad + 64 * GEN_PER;
  HIPCHK(dalloc(bb, &ctx->d_xi, (size_t)ctx->xi_cap * (dim == 3 ? 2 : 1)));
  HIPCHK(dalloc(bb, &ctx->d_off, n));
  HIPCHK(dalloc(bb, &ctx->d_init_tri, n));
  HIPCHK(dalloc(bb, &ctx->d_del_step, n));
  HIPCHK(dalloc(bb, &ctx->d_diam, n));
  HIPCHK(dalloc(bb, &ctx->d_vol, n));
  HIPCHK(dalloc(bb, &ctx->d_depth, dim == 3 ? n : 1));
  HIPCHK(dalloc(bb, &ctx->d_vals, n + 1));
  HIPCHK(dalloc(bb, &ctx->d_scan, n + 1));
  HIPCHK(dalloc(bb, &ctx->d_bsum, nblk(n + 1) + 1));
  HIPCHK(dalloc(bb, &ctx->d_total, 2));
  HIPCHK(dalloc(bb, &ctx->d_perm, n));
  HIPCHK(dalloc(bb, &ctx->d_skey, n));
  {
    double* tb;
    HIPCHK(dalloc(bb, &tb, (size_t)n));
    ctx->d_tmpbuf = tb;
  }
  // sort buffers and spares are allocated on the first sort
  ctx->d_scnt = nullptr;
  ctx->d_cursor = nullptr;
  ctx->d_sort_tiles = nullptr;
  ctx->sort_nt = 0;
  ctx->sort_ncls = 0;
  ctx->sp_gpool = ctx->sp_init_tri = ctx->sp_del_step = ctx->sp_inactive = ctx->sp_rank = nullptr;
  ctx->sp_diam = ctx->sp_vol = ctx->sp_depth = nullptr;
  ctx->sp_scen = nullptr;
  ctx->d_scen = nullptr;
  return CYL_OK;
}

extern "C" int cyl_particles_upload(cyl_ctx* ctx, int npart, int dim, const double* pos, const double* vel,
                                    const double* us, const int* inactive, const int* tri, const int* ll,
                                    const double* diameter, const double* volume, const int* init_tri,
                                    const cyl_pconst* pconst) {
  if (!ctx || npart <= 0 || (dim != 2 && dim != 3) || !pos || !inactive || !tri) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->part_bufs.free_all();
  ctx->d_scnt = nullptr;
  ctx->d_cursor = nullptr;
  ctx->d_sort_tiles = nullptr;
  ctx->sort_nt = 0;
  ctx->sp_gpool = ctx->sp_init_tri = ctx->sp_del_step = ctx->sp_inactive = ctx->sp_rank = nullptr;
  ctx->sp_diam = ctx->sp_vol = ctx->sp_depth = nullptr;
  ctx->ev_bufs.free_all();
  ctx->ev_units_cap = 0;
  ctx->batched = false;
  ctx->classes.clear();
  ctx->class_of_scen_h.clear();
  ctx->pregen_h.clear();
  ctx->sp_scen = nullptr;
  ctx->sort_ncls = 0;
  ctx->n = npart;
  ctx->dim = dim;
  ctx->cur = 0;
  if (pconst) {
    ParticleConst& p = ctx->pc;
    p.rho0 = pconst->rho0; p.density = pconst->density; p.drag_coef = pconst->drag_coef;
    p.added_mass_coef = pconst->added_mass_coef; p.buoyancy_velocity_value = pconst->buoyancy_velocity_value;
    p.drag_coef_model = pconst->drag_coef_model; p.added_mass_force = pconst->added_mass_force;
    p.buoyancy_velocity_model = pconst->buoyancy_velocity_model;
    p.a0c = pconst->a0c; p.a1c = pconst->a1c; p.a2c = pconst->a2c; p.a3c = pconst->a3c;
  }
  int e0 = alloc_slot_bufs(ctx, npart, 3);
  if (e0) return e0;
  int n = npart;
  ctx->nscen = 1;
  ctx->scen_base_h = {0, n};
  ctx->d_scen = nullptr;
  int e = ensure_scen_tables(ctx);
  if (e) return e;
  // SoA conversion on the host, physical order = pool order
  std::vector<double> col(n);
  std::vector<int> icol(n);
  PoolDev& A = ctx->pool[0];
  auto putd = [&](double* dst, const double* aos, int c, int stride) -> hipError_t {
    for (int i = 0; i < n; ++i) col[i] = aos ? aos[(size_t)i * stride + c] : 0.;
    return hipMemcpy(dst, col.data(), n * sizeof(double), hipMemcpyHostToDevice);
  };
  HIPCHK(putd(A.x, pos, 0, dim));
  HIPCHK(putd(A.y, pos, 1, dim));
  if (dim == 3) HIPCHK(putd(A.z, pos, 2, dim));
  for (int c = 0; c < dim; ++c) {
    HIPCHK(putd(A.up[c], vel, c, dim));
    HIPCHK(putd(A.us[c], us, c, dim));
  }
  HIPCHK(putd(ctx->d_diam, diameter, 0, 1));
  HIPCHK(putd(ctx->d_vol, volume, 0, 1));
  bool pend = false;
  for (int i = 0; i < n; ++i) {
    icol[i] = tri[i];
    if (!inactive[i] && tri[i] == -2) pend = true;
  }
  ctx->pending_init = pend;
  HIPCHK(hipMemcpy(A.tri, icol.data(), n * sizeof(int), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i) icol[i] = ll ? ll[i] : 0;
  HIPCHK(hipMemcpy(A.ll, icol.data(), n * sizeof(int), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i) icol[i] = inactive[i] ? 1 : 0;
  HIPCHK(hipMemcpy(A.inactive, icol.data(), n * sizeof(int), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i) icol[i] = init_tri ? init_tri[i] : -3;
  HIPCHK(hipMemcpy(ctx->d_init_tri, icol.data(), n * sizeof(int), hipMemcpyHostToDevice));
  for (int i = 0; i < n; ++i) icol[i] = i;
  HIPCHK(hipMemcpy(ctx->d_gpool, icol.data(), n * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(ctx->d_del_step, 0xff, n * sizeof(int)));
  HIPCHK(hipMemset(ctx->d_depth, 0, (dim == 3 ? n : 1) * sizeof(double)));
  // the second buffer starts as a copy (inactive rows are never rewritten there)
  PoolDev& B = ctx->pool[1];
  HIPCHK(hipMemcpy(B.tri, A.tri, n * sizeof(int), hipMemcpyDeviceToDevice));
  HIPCHK(hipMemcpy(B.ll, A.ll, n * sizeof(int), hipMemcpyDeviceToDevice));
  HIPCHK(hipMemset(ctx->d_flag, 0, 8 * sizeof(int)));
  int one = 1;
  HIPCHK(hipMemcpy(ctx->d_flag + 1, &one, sizeof(int), hipMemcpyHostToDevice));  // rank dirty
  HIPCHK(hipMemcpy(ctx->d_flag + 3, &one, sizeof(int), hipMemcpyHostToDevice));  // constant 1
  HIPCHK(hipMemset(ctx->d_n_deleted, 0, sizeof(int)));
  HIPCHK(hipMemset(ctx->d_n_deleted_total, 0, sizeof(long long)));
  ctx->have_last = false;
  ctx->us_from_up = false;
  // the uploaded rows are in pool order, not element order: the first step re-sorts
  ctx->steps_since_sort = SORT_NOW;
  return CYL_OK;
}

extern "C" int cyl_particles_download(cyl_ctx* ctx, int npart, double* pos, double* vel, double* us, int* inactive,
                                      int* tri, int* ll, double* lpos, double* lvel, int* ltri, double* depth) {
  if (!ctx || npart != ctx->n) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = ctx->n, dim = ctx->dim;
  std::vector<int> gp(n), ina(n);
  HIPCHK(hipMemcpy(gp.data(), ctx->d_gpool, n * sizeof(int), hipMemcpyDeviceToHost));
  const PoolDev& P = ctx->pool[ctx->cur];
  const PoolDev& L = ctx->pool[1 - ctx->cur];
  HIPCHK(hipMemcpy(ina.data(), P.inactive, n * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<double> col(n);
  std::vector<int> icol(n);
  auto getd = [&](const double* src, double* aos, int c, int stride, bool zero_inactive) -> hipError_t {
    hipError_t e = hipMemcpy(col.data(), src, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int j = 0; j < n; ++j) aos[(size_t)gp[j] * stride + c] = (zero_inactive && ina[j]) ? 0. : col[j];
    return hipSuccess;
  };
  const double* px[3] = {P.x, P.y, P.z};
  const double* lx[3] = {L.x, L.y, L.z};
  for (int c = 0; c < dim; ++c) {
    if (pos) HIPCHK(getd(px[c], pos, c, dim, true));
    if (vel) HIPCHK(getd(P.up[c], vel, c, dim, true));
    if (us) HIPCHK(getd(ctx->us_from_up ? P.up[c] : P.us[c], us, c, dim, true));
    if (lpos) HIPCHK(getd(ctx->have_last ? lx[c] : px[c], lpos, c, dim, true));
    if (lvel) HIPCHK(getd(ctx->have_last ? L.up[c] : P.up[c], lvel, c, dim, true));
  }
  if (depth && dim == 3) HIPCHK(getd(ctx->d_depth, depth, 0, 1, true));
  if (inactive)
    for (int j = 0; j < n; ++j) inactive[gp[j]] = ina[j] ? 1 : 0;
  auto geti = [&](const int* src, int* out, int dflt) -> hipError_t {
    hipError_t e = hipMemcpy(icol.data(), src, n * sizeof(int), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    for (int j = 0; j < n; ++j) out[gp[j]] = ina[j] ? dflt : icol[j];
    return hipSuccess;
  };
  if (tri) HIPCHK(geti(P.tri, tri, -1));
  if (ll) HIPCHK(geti(P.ll, ll, 0));
  if (ltri) HIPCHK(geti(ctx->have_last ? L.tri : P.tri, ltri, -1));
  return CYL_OK;
}

extern "C" int cyl_particles_deleted(cyl_ctx* ctx, int* pool_index, int cap, int* nout) {
  if (!ctx || !nout) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = ctx->n;
  std::vector<int> gp(n), ina(n), ds(n);
  const PoolDev& P = ctx->pool[ctx->cur];
  HIPCHK(hipMemcpy(gp.data(), ctx->d_gpool, n * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ina.data(), P.inactive, n * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ds.data(), ctx->d_del_step, n * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<std::pair<int, int>> del;
  for (int j = 0; j < n; ++j)
    if (ina[j] == 2) {
      del.push_back({ds[j], gp[j]});
      ina[j] = 1;
    }
  std::sort(del.begin(), del.end());
  *nout = (int)del.size();
  for (int q = 0; q < (int)del.size() && q < cap; ++q) pool_index[q] = del[q].second;
  if (!del.empty()) HIPCHK(hipMemcpy(P.inactive, ina.data(), n * sizeof(int), hipMemcpyHostToDevice));
  return CYL_OK;
}

extern "C" int cyl_particles_add(cyl_ctx* ctx, int nadd, const int* slots, const double* pos, const double* diam,
                                 const double* vol, const int* init_tri) {
  if (!ctx || nadd < 0 || (nadd > 0 && (!slots || !pos))) return CYL_E_ARG;
  if (nadd == 0) return CYL_OK;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  int n = ctx->n, dim = ctx->dim;
  std::vector<int> gp(n), slot_of(n);
  HIPCHK(hipMemcpy(gp.data(), ctx->d_gpool, n * sizeof(int), hipMemcpyDeviceToHost));
  for (int j = 0; j < n; ++j) slot_of[gp[j]] = j;
  PoolDev& P = ctx->pool[ctx->cur];
  for (int q = 0; q < nadd; ++q) {
    if (slots[q] < 0 || slots[q] >= n) return CYL_E_ARG;
    int j = slot_of[slots[q]];
    double* px[3] = {P.x, P.y, P.z};
    for (int c = 0; c < dim; ++c) {
      double v = pos[(size_t)q * dim + c], zero = 0.;
      HIPCHK(hipMemcpy(px[c] + j, &v, sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(P.up[c] + j, &zero, sizeof(double), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(P.us[c] + j, &zero, sizeof(double), hipMemcpyHostToDevice));
    }
    int m2 = -2, z0 = 0, it = init_tri ? init_tri[q] : -3;
    HIPCHK(hipMemcpy(P.tri + j, &m2, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.ll + j, &z0, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(P.inactive + j, &z0, sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_init_tri + j, &it, sizeof(int), hipMemcpyHostToDevice));
    if (diam) HIPCHK(hipMemcpy(ctx->d_diam + j, diam + q, sizeof(double), hipMemcpyHostToDevice));
    if (vol) HIPCHK(hipMemcpy(ctx->d_vol + j, vol + q, sizeof(double), hipMemcpyHostToDevice));
  }
  ctx->pending_init = true;
  // a large arrival (an eighth of the pool or more) re-sorts right after the
  // step that initializes it, when the new rows have their elements
  if (ctx->sort_interval > 0 && (int64_t)nadd * 8 >= n)
    ctx->steps_since_sort = std::max(ctx->steps_since_sort, ctx->sort_interval - 1);
  int one = 1;
  HIPCHK(hipMemcpy(ctx->d_flag + 1, &one, sizeof(int), hipMemcpyHostToDevice));
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// random stream
// ---------------------------------------------------------------------------
extern "C" int cyl_rng_set(cyl_ctx* ctx, uint64_t state, uint64_t inc) {
  if (!ctx) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->rng_state.assign(ctx->nscen, state);
  ctx->rng_inc.assign(ctx->nscen, inc);
  // scenario s gets the stream jumped by s<<48 (solver_parallel.pyx:113-115 idiom)
  for (int s = 1; s < ctx->nscen; ++s) {
    Pcg g{state, inc};
    pcg_advance(g, (uint64_t)s << 48);
    ctx->rng_state[s] = g.state;
  }
  ctx->rng_dirty = true;
  return CYL_OK;
}

extern "C" int cyl_rng_get(cyl_ctx* ctx, uint64_t* state, uint64_t* inc) {
  if (!ctx || !state || !inc) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (!ctx->rng_dirty && ctx->d_stream) {
    StreamDev st;
    HIPCHK(hipMemcpy(&st, ctx->d_stream, sizeof(StreamDev), hipMemcpyDeviceToHost));
    ctx->rng_state[0] = st.base;
  }
  *state = ctx->rng_state[0];
  *inc = ctx->rng_inc[0];
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// batched scenarios
// ---------------------------------------------------------------------------
__global__ void k_scen_of_slot(const int* gpool, const int* base, int nscen, int n, int* scen) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= n) return;
  int g = gpool[j], lo = 0, hi = nscen - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  scen[j] = lo;
}

static bool fast_params(const cyl_ctx* ctx, const cyl_params* p);
static int enqueue_sort(cyl_ctx* ctx, bool need_us, int& launches);

// Kernel classes: a fused kernel per model for the scenarios it covers that
// share the first one's Schmidt number (the packed nodal K), the generic
// kernel per model for the others; each class gets the slot range its rows
// occupy once sorted (class is the major sort key).
static int set_classes(cyl_ctx* ctx) {
  int nscen = ctx->nscen;
  const cyl_params* par = ctx->scen_par.data();
  const std::vector<int>& base = ctx->scen_base_h;
  ctx->classes.clear();
  ctx->class_of_scen_h.assign(nscen, 0);
  double sigc = 0.;
  bool have_fast = false;
  for (int s = 0; s < nscen; ++s)
    if (fast_params(ctx, &par[s]) && (!have_fast || par[s].schmidt_number == sigc)) {
      if (!have_fast) sigc = par[s].schmidt_number;
      have_fast = true;
    }
  std::vector<int64_t> count;
  auto class_of = [&](bool fast, int model) {
    for (size_t c = 0; c < ctx->classes.size(); ++c)
      if (ctx->classes[c].fast == fast && ctx->classes[c].model == model) return (int)c;
    ctx->classes.push_back({fast, model, 0, 0});
    count.push_back(0);
    return (int)ctx->classes.size() - 1;
  };
  for (int s = 0; s < nscen; ++s) {
    bool fast = have_fast && fast_params(ctx, &par[s]) && par[s].schmidt_number == sigc;
    int c = class_of(fast, par[s].model);
    ctx->class_of_scen_h[s] = c;
    count[c] += base[s + 1] - base[s];
  }
  int lo = 0;
  for (size_t c = 0; c < ctx->classes.size(); ++c) {
    ctx->classes[c].lo = lo;
    lo += (int)count[c];
    ctx->classes[c].hi = lo;
  }
  HIPCHK(hipMemcpy(ctx->d_class_of_scen, ctx->class_of_scen_h.data(), nscen * sizeof(int), hipMemcpyHostToDevice));
  return CYL_OK;
}

extern "C" int cyl_scenarios_set(cyl_ctx* ctx, int nscen, const int* base, const cyl_params* par,
                                 const uint64_t* state, const uint64_t* inc, const int* nsteps) {
  if (!ctx || nscen < 1 || !base || !par || !state || !inc || !nsteps) return CYL_E_ARG;
  if (ctx->n == 0 || !ctx->has_mesh || !ctx->has_fields) return ctx->set_err(CYL_E_STATE, -1, "upload mesh, fields and the pool first");
  if (base[0] != 0 || base[nscen] != ctx->n) return ctx->set_err(CYL_E_ARG, -1, "scenario ranges must cover the pool");
  for (int s = 0; s < nscen; ++s) {
    if (base[s + 1] < base[s]) return ctx->set_err(CYL_E_ARG, -1, "scenario ranges must be increasing");
    if (nsteps[s] < 0) return ctx->set_err(CYL_E_ARG, -1, "negative scenario duration");
  }
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->nscen = nscen;
  ctx->scen_base_h.assign(base, base + nscen + 1);
  int e = ensure_scen_tables(ctx);
  if (e) return e;
  if (!ctx->d_scen) HIPCHK(dalloc(ctx->part_bufs, &ctx->d_scen, ctx->n));
  // normals buffer with the per-scenario chunk padding (see k_gen_normals):
  // the upload's buffer when it is large enough, else a larger one
  const int64_t need_xi = (int64_t)ctx->n + GEN_PER * ((int64_t)nscen + 2) + 64 * GEN_PER;
  if (need_xi > ctx->xi_cap) {
    HIPCHK(ctx->part_bufs.release(ctx->d_xi));
    ctx->d_xi = nullptr;
    HIPCHK(dalloc(ctx->part_bufs, &ctx->d_xi, (size_t)need_xi * (ctx->dim == 3 ? 2 : 1)));  // 32 B entries in 3D
    ctx->xi_cap = need_xi;
  }
  k_scen_of_slot<<<nblk(ctx->n), BLK, 0, ctx->stream>>>(ctx->d_gpool, ctx->d_scen_base, nscen, ctx->n, ctx->d_scen);
  ctx->rng_state.assign(state, state + nscen);
  ctx->rng_inc.assign(inc, inc + nscen);
  ctx->rng_dirty = true;
  ctx->scen_par.assign(par, par + nscen);
  ctx->scen_nsteps.assign(nsteps, nsteps + nscen);
  ctx->batch_steps = 0;
  e = set_classes(ctx);
  if (e) return e;
  ctx->batched = true;
  ctx->pregen_h.clear();
  ctx->ev_bufs.free_all();
  ctx->ev_units_cap = 0;
  // ranks restart at every scenario; slots are grouped by kernel class now
  int launches = 0;
  k_set_flag<<<1, 1, 0, ctx->stream>>>(ctx->d_flag, 1, 1);
  e = enqueue_sort(ctx, true, launches);
  if (e) return e;
  ctx->steps_since_sort = 0;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CYL_OK;
}

extern "C" int cyl_scenario_rng_get(cyl_ctx* ctx, int s, uint64_t* state, uint64_t* inc) {
  if (!ctx || !state || !inc || s < 0 || s >= ctx->nscen) return CYL_E_ARG;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (!ctx->rng_dirty && ctx->d_stream) {
    StreamDev st;
    HIPCHK(hipMemcpy(&st, ctx->d_stream + s, sizeof(StreamDev), hipMemcpyDeviceToHost));
    ctx->rng_state[s] = st.base;
  }
  *state = ctx->rng_state[s];
  *inc = ctx->rng_inc[s];
  return CYL_OK;
}

// ---------------------------------------------------------------------------
// continuous batching: scenarios leave and join a running batch on the device
// ---------------------------------------------------------------------------
template <class T>
static void enqueue_scan(cyl_ctx* ctx, const T* in, int64_t* out, int64_t n, int64_t* bsum, int64_t* total,
                         const int64_t* ndev, const int* cond, int& launches);
static int run_steps(cyl_ctx* ctx, const cyl_params* p, double* time_inout, double dt, int nsteps, int* bnd_out,
                     bool init_only);

struct ReshapeArgs {
  PoolDev A;                  // current state of the old pool
  PoolDev N0, N1;             // both state copies of the new pool
  const int *gpool, *scen, *init_tri, *del_step;
  const double *diam, *vol, *depth;
  int *ngpool, *nscen_of, *ninit_tri, *ndel_step;
  double *ndiam, *nvol, *ndepth;
  const int* map;             // old scenario -> new scenario, or -1 (leaves)
  const int* obase;           // old scenario -> first pool index
  const int* nbase;           // new scenario -> first pool index
  const int64_t* pos;         // old slot -> new slot (exclusive scan of the kept flags)
  int n, dim;
};

__global__ void k_reshape_flags(const int* scen, const int* map, int n, int* keep) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j < n) keep[j] = map[scen[j]] >= 0 ? 1 : 0;
}

// kept rows move, in their physical order, to the front of the new pool
__global__ void k_reshape_copy(ReshapeArgs r) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= r.n) return;
  int s = r.scen[j], s2 = r.map[s];
  if (s2 < 0) return;
  int64_t q = r.pos[j];
  double v;
  v = r.A.x[j]; r.N0.x[q] = v; r.N1.x[q] = v;
  v = r.A.y[j]; r.N0.y[q] = v; r.N1.y[q] = v;
  if (r.dim == 3) { v = r.A.z[j]; r.N0.z[q] = v; r.N1.z[q] = v; }
  for (int c = 0; c < r.dim; ++c) {
    v = r.A.up[c][j]; r.N0.up[c][q] = v; r.N1.up[c][q] = v;
    v = r.A.us[c][j]; r.N0.us[c][q] = v; r.N1.us[c][q] = v;
  }
  int t = r.A.tri[j];
  r.N0.tri[q] = t; r.N1.tri[q] = t;
  t = r.A.ll[j];
  r.N0.ll[q] = t; r.N1.ll[q] = t;
  r.N0.inactive[q] = r.A.inactive[j];
  r.ngpool[q] = r.nbase[s2] + (r.gpool[j] - r.obase[s]);
  r.nscen_of[q] = s2;
  r.ninit_tri[q] = r.init_tri[j];
  r.ndel_step[q] = r.del_step[j];
  r.ndiam[q] = r.diam[j];
  r.nvol[q] = r.vol[j];
  if (r.dim == 3) r.ndepth[q] = r.depth[j];
}

struct CollectArgs {
  PoolDev P;
  const int *gpool, *scen, *del_step;
  int s, base, n, dim;
  double* out_d;   // [2 * dim + dim][npart]: position, velocity, fluid velocity seen
  int* out_i;      // [4][npart]: tri, lowerlayer, inactive, deletion step
  int npart;
};

// rows of scenario s, by their index in the scenario's own pool; reported
// deletions become plain inactive rows
__global__ void k_collect(CollectArgs a) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= a.n || a.scen[j] != a.s) return;
  int i = a.gpool[j] - a.base;
  if (i < 0 || i >= a.npart) return;
  const double* px[3] = {a.P.x, a.P.y, a.P.z};
  int np = a.npart;
  for (int c = 0; c < a.dim; ++c) {
    a.out_d[(size_t)c * np + i] = px[c][j];
    a.out_d[(size_t)(a.dim + c) * np + i] = a.P.up[c][j];
    a.out_d[(size_t)(2 * a.dim + c) * np + i] = a.P.us[c][j];
  }
  int ina = a.P.inactive[j];
  a.out_i[i] = a.P.tri[j];
  a.out_i[np + i] = a.P.ll[j];
  a.out_i[2 * np + i] = ina;
  a.out_i[3 * np + i] = a.del_step[j];
  if (ina == 2) a.P.inactive[j] = 1;
}

extern "C" int cyl_batch_collect(cyl_ctx* ctx, int s, int npart, double* pos, double* vel, double* us, int* inactive,
                                 int* tri, int* ll, int* deleted, int cap, int* ndeleted) {
  if (!ctx || !ctx->batched || s < 0 || s >= ctx->nscen || !ndeleted) return CYL_E_ARG;
  int base = ctx->scen_base_h[s], np = ctx->scen_base_h[s + 1] - base, dim = ctx->dim;
  if (npart != np) return ctx->set_err(CYL_E_ARG, -1, "npart differs from the scenario's pool size");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *ndeleted = 0;
  if (np == 0) return CYL_OK;
  Bufs tmp;
  CollectArgs a{};
  HIPCHK(dalloc(tmp, &a.out_d, (size_t)3 * dim * np));
  HIPCHK(dalloc(tmp, &a.out_i, (size_t)4 * np));
  a.P = ctx->pool[ctx->cur];
  a.gpool = ctx->d_gpool;
  a.scen = ctx->d_scen;
  a.del_step = ctx->d_del_step;
  a.s = s;
  a.base = base;
  a.n = ctx->n;
  a.dim = dim;
  a.npart = np;
  k_collect<<<nblk(ctx->n), BLK, 0, ctx->stream>>>(a);
  HIPCHK(hipGetLastError());
  std::vector<double> d((size_t)3 * dim * np);
  std::vector<int> iv((size_t)4 * np);
  HIPCHK(hipMemcpyAsync(d.data(), a.out_d, d.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(iv.data(), a.out_i, iv.size() * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  tmp.free_all();
  const int* ina = iv.data() + 2 * (size_t)np;
  for (int c = 0; c < dim; ++c)
    for (int i = 0; i < np; ++i) {
      bool z = ina[i] != 0;
      if (pos) pos[(size_t)i * dim + c] = z ? 0. : d[(size_t)c * np + i];
      if (vel) vel[(size_t)i * dim + c] = z ? 0. : d[(size_t)(dim + c) * np + i];
      if (us) us[(size_t)i * dim + c] = z ? 0. : d[(size_t)(2 * dim + c) * np + i];
    }
  std::vector<std::pair<int, int>> del;
  for (int i = 0; i < np; ++i) {
    bool z = ina[i] != 0;
    if (tri) tri[i] = z ? -1 : iv[i];
    if (ll) ll[i] = z ? 0 : iv[(size_t)np + i];
    if (inactive) inactive[i] = z ? 1 : 0;
    if (ina[i] == 2) del.push_back({iv[(size_t)3 * np + i], i});
  }
  std::sort(del.begin(), del.end());
  *ndeleted = (int)del.size();
  for (int q = 0; q < (int)del.size() && q < cap; ++q)
    if (deleted) deleted[q] = del[q].second;
  return CYL_OK;
}

// ---- cyl_batch_collect_noalloc: the same rows through the static staging array
struct CollectOut {
  PoolDev P;
  const int *gpool, *scen, *del_step;
  int s, base, n, dim, npart;
  double *pos, *vel, *us;   // (npart, dim) each, in the scenario's pool order
  int *tri, *ll, *inactive;
  unsigned long long* key;  // deleted since the last collect: (deletion step << 32) | row; else ~0
};

// every output row cleared to an inactive row: the array keeps the previous call's contents
__global__ void k_collect_clear(CollectOut a) {
  for (int i = blockIdx.x * BLK + threadIdx.x; i < a.npart; i += gridDim.x * BLK) {
    for (int c = 0; c < a.dim; ++c) {
      a.pos[(size_t)i * a.dim + c] = 0.;
      a.vel[(size_t)i * a.dim + c] = 0.;
      a.us[(size_t)i * a.dim + c] = 0.;
    }
    a.tri[i] = -1;
    a.ll[i] = 0;
    a.inactive[i] = 1;
    a.key[i] = ~0ull;
  }
}

// the rows of scenario s in the output layout of cyl_batch_collect (inactive
// rows zeroed); reported deletions become plain inactive rows
__global__ void k_collect_rows(CollectOut a) {
  int j = blockIdx.x * BLK + threadIdx.x;
  if (j >= a.n || a.scen[j] != a.s) return;
  int i = a.gpool[j] - a.base;
  if (i < 0 || i >= a.npart) return;
  int ina = a.P.inactive[j];
  if (ina == 0) {
    const double* px[3] = {a.P.x, a.P.y, a.P.z};
    for (int c = 0; c < a.dim; ++c) {
      a.pos[(size_t)i * a.dim + c] = px[c][j];
      a.vel[(size_t)i * a.dim + c] = a.P.up[c][j];
      a.us[(size_t)i * a.dim + c] = a.P.us[c][j];
    }
    a.tri[i] = a.P.tri[j];
    a.ll[i] = a.P.ll[j];
    a.inactive[i] = 0;
  } else if (ina == 2) {
    a.key[i] = ((unsigned long long)(unsigned)a.del_step[j] << 32) | (unsigned)i;
    a.P.inactive[j] = 1;
  }
}

extern "C" int cyl_batch_collect_noalloc(cyl_ctx* ctx, int s, int npart, double* pos, double* vel, double* us,
                                         int* inactive, int* tri, int* ll, int* deleted, int cap, int* ndeleted) {
  if (!ctx || !ctx->batched || s < 0 || s >= ctx->nscen || !ndeleted) return CYL_E_ARG;
  int base = ctx->scen_base_h[s], np = ctx->scen_base_h[s + 1] - base, dim = ctx->dim;
  if (npart != np) return ctx->set_err(CYL_E_ARG, -1, "npart differs from the scenario's pool size");
  if ((size_t)np > NOALLOC_ROWS)
    return ctx->set_err(CYL_E_ARG, -1,
                        "scenario of " + std::to_string(np) + " rows exceeds CYL_NOALLOC_MAX_ROWS (" +
                            std::to_string(NOALLOC_ROWS) + "); use cyl_batch_collect");
  if (ctx->device < 0 || ctx->device >= NOALLOC_MAX_DEVICES || !noalloc_ptr[ctx->device])
    return ctx->set_err(CYL_E_STATE, -1, "staging array of this device not initialised");
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  *ndeleted = 0;
  if (np == 0) return CYL_OK;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  unsigned char* base_ptr = (unsigned char*)noalloc_ptr[ctx->device];
  const size_t R = NOALLOC_ROWS;
  CollectOut a{};
  a.pos = (double*)base_ptr;
  a.vel = a.pos + 3 * R;
  a.us = a.vel + 3 * R;
  a.tri = (int*)(a.us + 3 * R);
  a.ll = a.tri + R;
  a.inactive = a.ll + R;
  a.key = (unsigned long long*)(a.inactive + R);
  a.P = ctx->pool[ctx->cur];
  a.gpool = ctx->d_gpool;
  a.scen = ctx->d_scen;
  a.del_step = ctx->d_del_step;
  a.s = s;
  a.base = base;
  a.n = ctx->n;
  a.dim = dim;
  a.npart = np;
  k_collect_clear<<<std::min(nblk(np), (unsigned)(4 * ctx->ncu)), BLK, 0, ctx->stream>>>(a);
  HIPCHK(hipGetLastError());
  k_collect_rows<<<nblk(ctx->n), BLK, 0, ctx->stream>>>(a);
  HIPCHK(hipGetLastError());
  const size_t nd = (size_t)np * dim;
  if (pos) HIPCHK(hipMemcpyAsync(pos, a.pos, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (vel) HIPCHK(hipMemcpyAsync(vel, a.vel, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (us) HIPCHK(hipMemcpyAsync(us, a.us, nd * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  if (tri) HIPCHK(hipMemcpyAsync(tri, a.tri, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (ll) HIPCHK(hipMemcpyAsync(ll, a.ll, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  if (inactive)
    HIPCHK(hipMemcpyAsync(inactive, a.inactive, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(noalloc_keys, a.key, (size_t)np * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // deletion order: by step, then by row (the serial loop's order within a step)
  size_t nk = 0;
  for (int i = 0; i < np; ++i)
    if (noalloc_keys[i] != ~0ull) noalloc_keys[nk++] = noalloc_keys[i];
  std::sort(noalloc_keys, noalloc_keys + nk);
  *ndeleted = (int)nk;
  if (deleted)
    for (size_t q = 0; q < nk && q < (size_t)std::max(cap, 0); ++q) deleted[q] = (int)(noalloc_keys[q] & 0xffffffffu);
  return CYL_OK;
}

extern "C" int cyl_batch_reshape(cyl_ctx* ctx, int nkeep, const int* keep, int nnew, const int* new_npart,
                                 const cyl_params* new_par, const uint64_t* new_state, const uint64_t* new_inc,
                                 const int* new_nsteps, const double* pos, const double* vel, const double* us,
                                 const int* inactive, const int* tri, const int* ll, const double* diameter,
                                 const double* volume, const int* init_tri, double time, double dt) {
  if (!ctx || !ctx->batched || nkeep < 0 || nnew < 0 || (nkeep > 0 && !keep)) return CYL_E_ARG;
  if (nnew > 0 && (!new_npart || !new_par || !new_state || !new_inc || !new_nsteps)) return CYL_E_ARG;
  if (nkeep + nnew < 1) return ctx->set_err(CYL_E_ARG, -1, "a batch keeps at least one scenario");
  const int ons = ctx->nscen, on = ctx->n, dim = ctx->dim;
  for (int q = 0; q < nkeep; ++q)
    if (keep[q] < 0 || keep[q] >= ons || (q > 0 && keep[q] <= keep[q - 1]))
      return ctx->set_err(CYL_E_ARG, -1, "kept scenarios must be increasing scenario ids");
  int64_t nadd = 0;
  for (int q = 0; q < nnew; ++q) {
    if (new_npart[q] < 0 || new_nsteps[q] < 0) return ctx->set_err(CYL_E_ARG, -1, "negative size or duration");
    if (new_par[q].rng_method != ctx->scen_par[0].rng_method)
      return ctx->set_err(CYL_E_ARG, -1, "the scenarios of a batch share one rng_method");
    nadd += new_npart[q];
  }
  if (nadd > 0 && (!pos || !inactive || !tri)) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // the kept scenarios' random streams as they stand now
  if (!ctx->rng_dirty) {
    std::vector<StreamDev> st(ons);
    HIPCHK(hipMemcpy(st.data(), ctx->d_stream, ons * sizeof(StreamDev), hipMemcpyDeviceToHost));
    for (int s = 0; s < ons; ++s) ctx->rng_state[s] = st[s].base;
  }
  // new scenario table: kept scenarios in order, then the new ones
  std::vector<int> map(ons, -1), nbase(1, 0);
  std::vector<cyl_params> par;
  std::vector<int> nsteps;
  std::vector<uint64_t> rs, ri;
  for (int q = 0; q < nkeep; ++q) {
    int s = keep[q];
    map[s] = q;
    nbase.push_back(nbase.back() + ctx->scen_base_h[s + 1] - ctx->scen_base_h[s]);
    par.push_back(ctx->scen_par[s]);
    nsteps.push_back(std::max(0, ctx->scen_nsteps[s] - ctx->batch_steps));
    rs.push_back(ctx->rng_state[s]);
    ri.push_back(ctx->rng_inc[s]);
  }
  const int nkept_rows = nbase.back();
  for (int q = 0; q < nnew; ++q) {
    nbase.push_back(nbase.back() + new_npart[q]);
    par.push_back(new_par[q]);
    nsteps.push_back(new_nsteps[q]);
    rs.push_back(new_state[q]);
    ri.push_back(new_inc[q]);
  }
  const int nn = nbase.back(), nns = nkeep + nnew;
  if (nn < 1) return ctx->set_err(CYL_E_ARG, -1, "the new batch has no rows");
  // compaction of the kept rows (old buffers stay alive until the copy is done)
  Bufs oldb, tmp;
  oldb.ptrs.swap(ctx->part_bufs.ptrs);
  PoolDev A = ctx->pool[ctx->cur];
  const int *o_gpool = ctx->d_gpool, *o_scen = ctx->d_scen, *o_init = ctx->d_init_tri, *o_del = ctx->d_del_step;
  const double *o_diam = ctx->d_diam, *o_vol = ctx->d_vol, *o_depth = ctx->d_depth;
  int *d_map, *d_obase, *d_nbase, *d_keepf;
  int64_t *d_pos, *d_bs, *d_tot;
  HIPCHK(dalloc(tmp, &d_map, ons));
  HIPCHK(dalloc(tmp, &d_obase, ons + 1));
  HIPCHK(dalloc(tmp, &d_nbase, nns + 1));
  HIPCHK(dalloc(tmp, &d_keepf, on));
  HIPCHK(dalloc(tmp, &d_pos, on + 1));
  HIPCHK(dalloc(tmp, &d_bs, nblk(on + 1, SCAN_TILE) + 1));
  HIPCHK(dalloc(tmp, &d_tot, 2));
  HIPCHK(hipMemcpy(d_map, map.data(), ons * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_obase, ctx->scen_base_h.data(), (ons + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_nbase, nbase.data(), (nns + 1) * sizeof(int), hipMemcpyHostToDevice));
  int launches = 0;
  k_reshape_flags<<<nblk(on), BLK, 0, ctx->stream>>>(o_scen, d_map, on, d_keepf);
  enqueue_scan(ctx, (const int*)d_keepf, d_pos, (int64_t)on, d_bs, d_tot, (const int64_t*)nullptr, (const int*)nullptr,
               launches);
  int e = alloc_slot_bufs(ctx, nn, nns + 2);
  if (e) return e;
  HIPCHK(dalloc(ctx->part_bufs, &ctx->d_scen, nn));
  ReshapeArgs r{};
  r.A = A;
  r.N0 = ctx->pool[0];
  r.N1 = ctx->pool[1];
  r.gpool = o_gpool; r.scen = o_scen; r.init_tri = o_init; r.del_step = o_del;
  r.diam = o_diam; r.vol = o_vol; r.depth = o_depth;
  r.ngpool = ctx->d_gpool; r.nscen_of = ctx->d_scen; r.ninit_tri = ctx->d_init_tri; r.ndel_step = ctx->d_del_step;
  r.ndiam = ctx->d_diam; r.nvol = ctx->d_vol; r.ndepth = ctx->d_depth;
  r.map = d_map; r.obase = d_obase; r.nbase = d_nbase; r.pos = d_pos;
  r.n = on;
  r.dim = dim;
  k_reshape_copy<<<nblk(on), BLK, 0, ctx->stream>>>(r);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  oldb.free_all();
  tmp.free_all();
  // the new scenarios' rows, in their pool order, after the kept ones
  bool pend = false;
  if (nadd > 0) {
    size_t m = (size_t)nadd, o = (size_t)nkept_rows;
    std::vector<double> col(m);
    std::vector<int> icol(m);
    auto putd = [&](double* d0, double* d1, const double* aos, int c, int stride) -> hipError_t {
      for (size_t i = 0; i < m; ++i) col[i] = aos ? aos[i * stride + c] : 0.;
      hipError_t er = hipMemcpy(d0 + o, col.data(), m * sizeof(double), hipMemcpyHostToDevice);
      if (er == hipSuccess && d1) er = hipMemcpy(d1 + o, col.data(), m * sizeof(double), hipMemcpyHostToDevice);
      return er;
    };
    auto puti = [&](int* d0, int* d1) -> hipError_t {
      hipError_t er = hipMemcpy(d0 + o, icol.data(), m * sizeof(int), hipMemcpyHostToDevice);
      if (er == hipSuccess && d1) er = hipMemcpy(d1 + o, icol.data(), m * sizeof(int), hipMemcpyHostToDevice);
      return er;
    };
    PoolDev &P0 = ctx->pool[0], &P1 = ctx->pool[1];
    HIPCHK(putd(P0.x, P1.x, pos, 0, dim));
    HIPCHK(putd(P0.y, P1.y, pos, 1, dim));
    if (dim == 3) HIPCHK(putd(P0.z, P1.z, pos, 2, dim));
    for (int c = 0; c < dim; ++c) {
      HIPCHK(putd(P0.up[c], P1.up[c], vel, c, dim));
      HIPCHK(putd(P0.us[c], P1.us[c], us, c, dim));
    }
    HIPCHK(putd(ctx->d_diam, nullptr, diameter, 0, 1));
    HIPCHK(putd(ctx->d_vol, nullptr, volume, 0, 1));
    if (dim == 3) HIPCHK(putd(ctx->d_depth, nullptr, nullptr, 0, 1));
    for (size_t i = 0; i < m; ++i) {
      icol[i] = tri[i];
      if (!inactive[i] && tri[i] == -2) pend = true;
    }
    HIPCHK(puti(P0.tri, P1.tri));
    for (size_t i = 0; i < m; ++i) icol[i] = ll ? ll[i] : 0;
    HIPCHK(puti(P0.ll, P1.ll));
    for (size_t i = 0; i < m; ++i) icol[i] = inactive[i] ? 1 : 0;
    HIPCHK(puti(P0.inactive, nullptr));
    for (size_t i = 0; i < m; ++i) icol[i] = init_tri ? init_tri[i] : -3;
    HIPCHK(puti(ctx->d_init_tri, nullptr));
    for (size_t i = 0; i < m; ++i) icol[i] = -1;
    HIPCHK(puti(ctx->d_del_step, nullptr));
    size_t i = 0;
    for (int q = 0; q < nnew; ++q)
      for (int k = 0; k < new_npart[q]; ++k, ++i) icol[i] = nbase[nkeep + q] + k;
    HIPCHK(puti(ctx->d_gpool, nullptr));
    i = 0;
    for (int q = 0; q < nnew; ++q)
      for (int k = 0; k < new_npart[q]; ++k, ++i) icol[i] = nkeep + q;
    HIPCHK(puti(ctx->d_scen, nullptr));
  }
  // scenario tables
  ctx->nscen = nns;
  ctx->scen_base_h = nbase;
  e = ensure_scen_tables(ctx);
  if (e) return e;
  ctx->rng_state = rs;
  ctx->rng_inc = ri;
  ctx->rng_dirty = true;
  ctx->scen_par = par;
  ctx->scen_nsteps = nsteps;
  ctx->batch_steps = 0;
  ctx->cur = 0;
  ctx->have_last = false;
  ctx->pending_init = ctx->pending_init || pend;
  e = set_classes(ctx);
  if (e) return e;
  ctx->pregen_h.clear();
  ctx->ev_bufs.free_all();
  ctx->ev_units_cap = 0;
  HIPCHK(hipMemsetAsync(ctx->d_n_deleted, 0, sizeof(int), ctx->stream));
  // ranks restart at every scenario; slots are grouped by kernel class again
  k_set_flag<<<1, 1, 0, ctx->stream>>>(ctx->d_flag, 1, 1);
  e = enqueue_sort(ctx, true, launches);
  if (e) return e;
  ctx->steps_since_sort = 0;
  if (pend) {
    // initialize_particle_state of the new rows with the fields at `time`, as a
    // Solver built at that time does; the interpolation state of the running
    // batch (the last two steps' fields) is then rebuilt
    ctx->fields.frozen = ctx->frozen;
    int nh = ctx->interp_count;
    double h0 = ctx->interp_hist[0], h1 = ctx->interp_hist[1];
    if (!ctx->frozen) {
      ctx->record = 0;
      e = enqueue_time_interp(ctx, time, launches);
      if (e) return e;
    }
    double t = time;
    e = run_steps(ctx, &ctx->scen_par[0], &t, dt, 1, nullptr, true);
    if (e) return e;
    ctx->have_last = false;
    if (!ctx->frozen && nh >= 1) {
      ctx->record = 0;
      if (nh >= 2) {
        e = enqueue_time_interp(ctx, h0, launches);
        if (e) return e;
      }
      ctx->record = 0;
      e = enqueue_time_interp(ctx, h1, launches);
      if (e) return e;
    }
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return CYL_OK;
}

extern "C" uint32_t cyl_pcg32_next(uint64_t* state, uint64_t inc) {
  Pcg g{*state, inc};
  uint32_t r = pcg_next(g);
  *state = g.state;
  return r;
}

extern "C" void cyl_pcg32_seed(uint64_t* state, uint64_t* inc, uint64_t seed, uint64_t stream) {
  Pcg g;
  pcg_seed(g, seed, stream);
  *state = g.state;
  *inc = g.inc;
}

extern "C" void cyl_pcg32_advance(uint64_t* state, uint64_t inc, uint64_t delta) {
  Pcg g{*state, inc};
  pcg_advance(g, delta);
  *state = g.state;
}

static ZigTables& host_zig() {
  static ZigTables z;
  static bool ready = false;
  if (!ready) {
    double m1 = 2147483648.0, dn = 3.442619855899, tn = dn, vn = 9.91256303526217e-3;
    double q = vn / std::exp(-0.5 * dn * dn);
    z.kn[0] = (uint32_t)((dn / q) * m1);
    z.kn[1] = 0;
    z.wn[0] = q / m1;
    z.wn[127] = dn / m1;
    z.fn[0] = 1.0;
    z.fn[127] = std::exp(-0.5 * dn * dn);
    for (int i = 126; i > 0; --i) {
      dn = std::sqrt(-2.0 * std::log(vn / dn + std::exp(-0.5 * dn * dn)));
      z.kn[i + 1] = (uint32_t)((dn / tn) * m1);
      tn = dn;
      z.fn[i] = std::exp(-0.5 * dn * dn);
      z.wn[i] = dn / m1;
    }
    ready = true;
  }
  return z;
}

extern "C" double cyl_random_normal(uint64_t* state, uint64_t inc, int method) {
  Pcg g{*state, inc};
  double r;
  auto ex = [](double v) { return std::exp(v); };
  auto lg = [](double v) { return std::log(v); };
  auto sq = [](double v) { return std::sqrt(v); };
  if (method == 1) r = rect_normal(g);
  else if (method == 2) r = zig_normal_t(g, host_zig(), ex, lg);
  else {
    double a, b;
    polar_pair_t(g, a, b, lg, sq);
    r = a;
  }
  *state = g.state;
  return r;
}

extern "C" void cyl_random_pair(uint64_t* state, uint64_t inc, int method, double* a, double* b) {
  Pcg g{*state, inc};
  auto ex = [](double v) { return std::exp(v); };
  auto lg = [](double v) { return std::log(v); };
  auto sq = [](double v) { return std::sqrt(v); };
  if (method == 1) {
    *a = rect_normal(g);
    *b = rect_normal(g);
  } else if (method == 2) {
    *a = zig_normal_t(g, host_zig(), ex, lg);
    *b = zig_normal_t(g, host_zig(), ex, lg);
  } else {
    polar_pair_t(g, *a, *b, lg, sq);
  }
  *state = g.state;
}

// ---------------------------------------------------------------------------
// step
// ---------------------------------------------------------------------------
static ParamsDev to_dev(const cyl_params* p) {
  ParamsDev d{};
  d.model = p->model; d.time_scheme = p->time_scheme; d.particle_velocity_init = p->particle_velocity_init;
  d.rng_method = p->rng_method; d.diffusion_model = p->diffusion_model;
  d.diffusion_lsm2_option = p->diffusion_lsm2_option; d.horizontal_diffusivity = p->horizontal_diffusivity;
  d.vertical_diffusivity = p->vertical_diffusivity; d.schmidt_number = p->schmidt_number;
  d.boundary_conditions = p->boundary_conditions; d.boundary_conditions_type = p->boundary_conditions_type;
  d.max_wall_rebounds = p->max_wall_rebounds; d.bnd_statistics = p->bnd_statistics;
  d.rebound_damping_coef = p->rebound_damping_coef; d.wall_roughness_angle = p->wall_roughness_angle;
  d.stranding_probability = p->stranding_probability; d.compute_depth = p->compute_depth;
  d.lsm3_edge_opt = p->lsm3_edge_opt; d.lsm3_tl_horizontal = p->diffusion_lsm3_tl_horizontal;
  d.lsm3_tl_vertical = p->diffusion_lsm3_tl_vertical;
  d.stranding_output = p->stranding_output;
  d.bottom_deposition = p->bottom_deposition;
  return d;
}

// Draw units one active row takes in a step when nothing is data-dependent;
// -1 when no uniform prediction is made.  `umax` bounds the units of any row.
static int predict_units(const cyl_params* p, int dim, double dt, int& umax) {
  int pair = p->rng_method == 3 ? 1 : 2, single = 1;
  double sqrtdt = std::sqrt(dt);
  int dm = p->diffusion_model;
  umax = 2 * pair + 2 * single + 3 * pair;
  if (p->model == 0) return 0;
  if (p->model == 1 || (p->model == 2 && (p->diffusion_lsm2_option == 1 || dm == 0))) {
    if (dm == 0) return 0;
    int u = 0;
    if (dm == 2 || dm == 3) u += pair;  // K > 0 predicted
    else if (std::sqrt(2. * p->horizontal_diffusivity) * sqrtdt != 0.) u += pair;
    if (dim == 3) {
      double dz = (dm == 3) ? 1. : std::sqrt(2. * p->vertical_diffusivity) * sqrtdt;
      if (dm == 3 || dz != 0.) u += single;
    }
    return u;
  }
  if (p->model == 2) {
    if (p->time_scheme == 1) return pair + (dim == 3 ? single : 0);
    return dim * pair;  // direct integrator: one pair per component with nonzero amplitude
  }
  if (p->model == 3) {
    // seen-velocity fluctuations at arrival (pair + single) and the direct
    // integrator's pair + single per component
    umax = 4 * pair + 4 * single;
    if (p->time_scheme == 1) return pair + (dim == 3 ? single : 0);
    return -1;  // direct integrator: draws depend on the covariance factor
  }
  return -1;
}

static int ensure_events(cyl_ctx* ctx, int64_t units_per_scen_cap, int method) {
  if (method == 1) return CYL_OK;
  int64_t need = units_per_scen_cap;
  // event counts and extra positions are packed as two 32-bit halves
  if (need > (1LL << 30)) return ctx->set_err(CYL_E_ARG, -1, "too many draws in one step for the event tables");
  if (ctx->ev_units_cap >= need && ctx->ev_bufs.ptrs.size()) return CYL_OK;
  ctx->ev_bufs.free_all();
  ctx->ev_units_cap = need;
  int ns = ctx->nscen;
  int64_t nch = 0;
  // chunk capacity per scenario, proportional to its pool size
  std::vector<int64_t> cap_units(ns);
  for (int s = 0; s < ns; ++s) {
    int64_t rows = ctx->scen_base_h[s + 1] - ctx->scen_base_h[s];
    cap_units[s] = std::max<int64_t>(rows, 1) * (need / std::max<int64_t>(ctx->n, 1) + 1);
  }
  int64_t evcap = 0;
  for (int s = 0; s < ns; ++s) {
    int64_t u = cap_units[s];
    int64_t w = method == 2 ? u + u / 8 + 4096 : 2 * (u + u / 2 + 4096);
    nch += (w + ZCHUNK - 1) / ZCHUNK + 1;
    evcap += method == 2 ? w / 24 + 4096 : w / 4 + 4096;
  }
  ctx->nchunks = nch;
  ctx->ev_cap = evcap;
  // about one Ziggurat slow event (rate ~1.2%) per 64 normals, and about two
  // rejected polar attempts (rate 1 - pi/4) per 8 calls
  ctx->bshift = method == 2 ? 6 : 3;
  std::vector<int64_t> bb(ns + 1, 0);
  for (int s = 0; s < ns; ++s) bb[s + 1] = bb[s] + (cap_units[s] >> ctx->bshift) + 2;
  ctx->nbucket = bb[ns];
  Bufs& b = ctx->ev_bufs;
  HIPCHK(dalloc(b, &ctx->d_bucket, ctx->nbucket));
  HIPCHK(dalloc(b, &ctx->d_bucket_base, ns + 1));
  HIPCHK(hipMemcpy(ctx->d_bucket_base, bb.data(), (ns + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(dalloc(b, &ctx->d_chunk_base, ns + 1));  // the step's layout (ev_layout)
  HIPCHK(dalloc(b, &ctx->d_lb_base, ns + 1));
  ctx->nspan = (nch + BLK - 1) / BLK;
  const int64_t nreg = ctx->nspan * (method == 2 ? EV_SPAN_CAP_ZIG : EV_SPAN_CAP_POLAR);
  HIPCHK(dalloc(b, &ctx->d_chunk_loc, nch + 1));
  HIPCHK(dalloc(b, &ctx->d_span_cnt, ctx->nspan + 1));
  HIPCHK(dalloc(b, &ctx->d_span_off, ctx->nspan + 1));
  HIPCHK(dalloc(b, &ctx->d_reg_pos, nreg));
  if (method == 2) {
    HIPCHK(dalloc(b, &ctx->d_span_max, ctx->nspan + 1));
    HIPCHK(dalloc(b, &ctx->d_span_carry, ctx->nspan + 1));
    HIPCHK(dalloc(b, &ctx->d_span_pk, ctx->nspan + 1));
    HIPCHK(dalloc(b, &ctx->d_span_pkoff, ctx->nspan + 1));
    HIPCHK(dalloc(b, &ctx->d_scen_c, ns + 1));
    HIPCHK(dalloc(b, &ctx->d_scen_c0, ns + 1));
    HIPCHK(dalloc(b, &ctx->d_reg_len, nreg));
    HIPCHK(dalloc(b, &ctx->d_reg_val, nreg));
    HIPCHK(dalloc(b, &ctx->d_reg_key, nreg));
  }
  HIPCHK(dalloc(b, &ctx->d_ev_pos, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_ev_len, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_reach, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_E, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_key, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_xtra, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_ev_val, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_cand_val, evcap + 1));
  HIPCHK(dalloc(b, &ctx->d_ev_begin, ns + 1));
  HIPCHK(dalloc(b, &ctx->d_nev, 1));
  HIPCHK(dalloc(b, &ctx->d_chunk_live, 1));
  return CYL_OK;
}

// blocks of a scan pass: one per tile, at most 8 per CU
static unsigned scan_grid(const cyl_ctx* ctx, int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(n, SCAN_TILE), 8LL * ctx->ncu));
}

// exclusive scan of n values (int or int64) into out (int64, n entries) with total
template <class T>
static void enqueue_scan(cyl_ctx* ctx, const T* in, int64_t* out, int64_t n, int64_t* bsum, int64_t* total,
                         const int64_t* ndev, const int* cond, int& launches) {
  // tiles are visited grid-stride, only the live ones when a device count bounds the input
  unsigned nb = scan_grid(ctx, n);
  k_scan_tiles<T, OpSum><<<nb, BLK, 0, ctx->stream>>>(in, out, bsum, n, ndev, cond);
  k_scan_top<OpSum><<<1, 1024, 0, ctx->stream>>>(bsum, n, ndev, total, cond);
  k_scan_fix<<<nb, BLK, 0, ctx->stream>>>(out, bsum, n, ndev, cond);
  launches += 3;
}

// active rank of every row within its scenario (pool order); the scenario
// totals follow in k_step_prep
static void enqueue_rank(cyl_ctx* ctx, const int* cond, int& launches) {
  int n = ctx->n;
  const PoolDev& A = ctx->pool[ctx->cur];
  const unsigned gg = std::min<unsigned>(nblk(n), GUARD_GRID);
  k_scatter_active<<<gg, BLK, 0, ctx->stream>>>(A.inactive, ctx->d_gpool, n, ctx->d_vals, cond);
  enqueue_scan(ctx, ctx->d_vals, ctx->d_scan, (int64_t)n, ctx->d_bsum, ctx->d_total, (const int64_t*)nullptr, cond,
               launches);
  k_gather_rank<<<gg, BLK, 0, ctx->stream>>>(ctx->d_scan, ctx->d_gpool, ctx->d_scen, ctx->d_scen_base, n,
                                                  ctx->d_rank, nullptr, cond);
  launches += 2;
}

// event tables of the step's stream window (rng_method 2 and 3)
// the step's event table layout (ev_layout)
static LayoutArgs layout_args(cyl_ctx* ctx, int method) {
  LayoutArgs l{};
  l.st = ctx->d_stream;
  l.units = ctx->d_scen_units;
  l.nscen = ctx->nscen;
  l.method = method;
  l.bshift = ctx->bshift;
  l.nchunks = ctx->nchunks;
  l.bucket_base = ctx->d_bucket_base;
  l.base = ctx->d_chunk_base;
  l.live = ctx->d_chunk_live;
  l.lb_base = ctx->d_lb_base;
  l.overflow = ctx->d_flag + 2;
  return l;
}

// layout: lay the tables out here (the step's preparation has not: the units
// come from a counting run)
static void enqueue_events(cyl_ctx* ctx, int method, const int* cond, bool layout, int& launches) {
  if (method == 1) return;
  if (layout) {
    k_ev_layout<<<1, BLK, 0, ctx->stream>>>(layout_args(ctx, method), cond);
    ++launches;
  }
  EventArgs e{};
  e.stream = ctx->d_stream;
  e.jt = ctx->d_jt;
  e.jt8 = ctx->d_jt8;
  e.zt = ctx->d_zt;
  e.chunk_base = ctx->d_chunk_base;
  e.nscen = ctx->nscen;
  e.nchunks = ctx->nchunks;
  e.chunk_loc = ctx->d_chunk_loc;
  e.span_cnt = ctx->d_span_cnt;
  e.span_off = ctx->d_span_off;
  e.span_max = ctx->d_span_max;
  e.span_carry = ctx->d_span_carry;
  e.span_cap = method == 2 ? EV_SPAN_CAP_ZIG : EV_SPAN_CAP_POLAR;
  e.reg_pos = ctx->d_reg_pos;
  e.reg_len = ctx->d_reg_len;
  e.reg_val = ctx->d_reg_val;
  e.reg_key = ctx->d_reg_key;
  e.nev = ctx->d_nev;
  e.ev_begin = ctx->d_ev_begin;
  e.ev_pos = ctx->d_ev_pos;
  e.ev_len = ctx->d_ev_len;
  e.ev_val = ctx->d_cand_val;
  e.reach = ctx->d_reach;
  e.chunk_live = ctx->d_chunk_live;
  e.ev_cap = ctx->ev_cap;
  e.overflow = ctx->d_flag + 2;
  e.cond = cond;
  hipStream_t s = ctx->stream;
  auto pgrid = [&](const void* fn) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BLK, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(ctx->nchunks), (int64_t)ctx->ncu * per_cu));
  };
  // the gather visits the spans of the live chunks, a fraction of the capacity
  const unsigned ngat = (unsigned)std::max<int64_t>(1, std::min<int64_t>(ctx->nspan, 4LL * ctx->ncu));
  if (method == 2) {
    k_ev_write<2><<<pgrid((const void*)k_ev_write<2>), BLK, 0, s>>>(e);
    k_ev_span_scan<2><<<1, BLK, 0, s>>>(e);
    k_ev_gather<2><<<ngat, BLK, 0, s>>>(e);
  } else {
    k_ev_write<3><<<pgrid((const void*)k_ev_write<3>), BLK, 0, s>>>(e);
    k_ev_span_scan<3><<<1, BLK, 0, s>>>(e);
    k_ev_gather<3><<<ngat, BLK, 0, s>>>(e);
  }
  launches += 3;
  // the candidate kernels loop over the step's candidates (device count), far
  // fewer than the capacity: 4 blocks per CU
  unsigned nb = std::max(std::min<unsigned>(nblk(ctx->ev_cap), 4u * (unsigned)ctx->ncu), nblk(ctx->nscen));
  if (method == 2) {
    ChainArgs ca{};
    ca.pos = ctx->d_ev_pos;
    ca.len = ctx->d_ev_len;
    ca.val = ctx->d_cand_val;
    ca.reach = ctx->d_reach;
    ca.ev_begin = ctx->d_ev_begin;
    ca.span_cnt = ctx->d_span_cnt;
    ca.span_off = ctx->d_span_off;
    ca.chunk_live = ctx->d_chunk_live;
    ca.nscen = ctx->nscen;
    ca.ev_cap = ctx->ev_cap;
    ca.pk = ctx->d_E;
    ca.span_pk = ctx->d_span_pk;
    ca.span_pkoff = ctx->d_span_pkoff;
    ca.scen_c = ctx->d_scen_c;
    ca.scen_c0 = ctx->d_scen_c0;
    ca.st = ctx->d_stream;
    ca.key = ctx->d_key;
    ca.xtra = ctx->d_xtra;
    ca.cval = ctx->d_ev_val;
    ca.bucket = ctx->d_bucket;
    ca.bucket_base = ctx->d_bucket_base;
    ca.lb_base = ctx->d_lb_base;
    ca.bshift = ctx->bshift;
    ca.cond = cond;
    k_ev_resolve<<<ngat, BLK, 0, s>>>(ca);
    k_ev_chain_scan<<<1, BLK, 0, s>>>(ca);
    k_ev_compact<<<ngat, BLK, 0, s>>>(ca);
    launches += 3;
  } else {
    k_ev_polar<<<nb, BLK, 0, s>>>(ctx->d_stream, ctx->nscen, ctx->d_ev_begin, ctx->d_nev, ctx->d_ev_pos, ctx->d_key,
                                  ctx->d_bucket, ctx->d_bucket_base, ctx->d_lb_base, ctx->bshift, cond);
    launches += 1;
  }
  const unsigned nbb = std::min<unsigned>(nblk(ctx->nbucket), 4u * (unsigned)ctx->ncu);
  k_ev_buckets<<<nbb, BLK, 0, s>>>(ctx->d_stream, ctx->nscen, ctx->d_lb_base,
                                   method == 2 ? ctx->d_scen_c0 : ctx->d_ev_begin, ctx->d_key, ctx->d_bucket,
                                   ctx->d_bucket_base, cond);
  launches += 1;
}

static unsigned persistent_grid(cyl_ctx* ctx, const void* fn, int64_t nblocks, bool mult8);

template <int DIM, int MODEL, bool COUNT>
static void launch_step_t(cyl_ctx* ctx, const StepArgs& a) {
  unsigned g = persistent_grid(ctx, (const void*)k_step<DIM, MODEL, COUNT>, nblk(a.n, BLK), false);
  k_step<DIM, MODEL, COUNT><<<g, BLK, 0, ctx->stream>>>(a);
}

static void launch_step(cyl_ctx* ctx, const StepArgs& a, int model, bool count) {
  int d = ctx->dim;
  if (d == 2) {
    if (model == 0) count ? launch_step_t<2, 0, true>(ctx, a) : launch_step_t<2, 0, false>(ctx, a);
    else if (model == 1) count ? launch_step_t<2, 1, true>(ctx, a) : launch_step_t<2, 1, false>(ctx, a);
    else if (model == 2) count ? launch_step_t<2, 2, true>(ctx, a) : launch_step_t<2, 2, false>(ctx, a);
    else count ? launch_step_t<2, 3, true>(ctx, a) : launch_step_t<2, 3, false>(ctx, a);
  } else {
    if (model == 0) count ? launch_step_t<3, 0, true>(ctx, a) : launch_step_t<3, 0, false>(ctx, a);
    else if (model == 1) count ? launch_step_t<3, 1, true>(ctx, a) : launch_step_t<3, 1, false>(ctx, a);
    else if (model == 2) count ? launch_step_t<3, 2, true>(ctx, a) : launch_step_t<3, 2, false>(ctx, a);
    else count ? launch_step_t<3, 3, true>(ctx, a) : launch_step_t<3, 3, false>(ctx, a);
  }
}

static int upload_streams(cyl_ctx* ctx, const std::vector<int>& units_per_part) {
  std::vector<StreamDev> st(ctx->nscen);
  HIPCHK(hipMemcpy(st.data(), ctx->d_stream, ctx->nscen * sizeof(StreamDev), hipMemcpyDeviceToHost));
  for (int s = 0; s < ctx->nscen; ++s) {
    if (ctx->rng_dirty) {
      st[s].base = ctx->rng_state[s];
      st[s].inc = ctx->rng_inc[s];
    }
    st[s].units_per_part = units_per_part[s];
  }
  if (ctx->rng_dirty) {
    std::vector<JumpTable> jt(ctx->nscen);
    std::vector<JumpTable8> jt8(ctx->nscen);
    for (int s = 0; s < ctx->nscen; ++s) {
      make_jump_table(ctx->rng_inc[s], jt[s]);
      make_jump_table8(ctx->rng_inc[s], jt8[s]);
    }
    HIPCHK(hipMemcpy(ctx->d_jt, jt.data(), ctx->nscen * sizeof(JumpTable), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_jt8, jt8.data(), ctx->nscen * sizeof(JumpTable8), hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(ctx->d_stream, st.data(), ctx->nscen * sizeof(StreamDev), hipMemcpyHostToDevice));
  ctx->rng_dirty = false;
  return CYL_OK;
}

// The fused kernels cover LSM-1 with the Euler scheme, and LSM-2 with the
// direct integrator and no velocity noise (diffusion_lsm2_option 1, or
// diffusion model 0), without added mass and without wall roughness, with
// diffusion models 0-3, on 2D fields with 2D particles or on sigma-layer 3D
// fields with 3D particles (without depth output), once every row is
// initialized; every other configuration runs the generic k_step.
static bool fast_params(const cyl_ctx* ctx, const cyl_params* p) {
  const FieldsDev& f = ctx->fields;
  if (ctx->dim != f.dim) return false;  // pseudo-3D runs the generic kernel
  if (f.dim == 3 && (p->compute_depth || f.fid_zs < 0 || f.fid_uz < 0 || f.nlayers < 2)) return false;
  if (p->diffusion_model < 0 || p->diffusion_model > 3) return false;
  if (p->stranding_output && p->stranding_probability > 0.) return false;  // the generic kernel records them
  if (p->model == 1) {
    if (p->time_scheme != 1) return false;
  } else if (p->model == 2) {
    int opt = p->diffusion_lsm2_option;
    if (p->time_scheme != 5 || !(opt == 1 || (opt >= 2 && p->diffusion_model == 0))) return false;
    if (ctx->pc.added_mass_force) return false;
    if (p->boundary_conditions_type == 2 && p->wall_roughness_angle > 0.) return false;
  } else {
    return false;
  }
  if (f.npoints != ctx->mesh.nv || f.fid_ux < 0 || f.fid_uy < 0) return false;
  if (p->diffusion_model >= 2 && (f.fid_turbeng < 0 || f.fid_dissip < 0)) return false;
  return true;
}

static bool fast_eligible(const cyl_ctx* ctx, const cyl_params* p, bool init_only) {
  if (init_only || ctx->pending_init || ctx->nscen != 1) return false;
  return fast_params(ctx, p);
}

static int ensure_fast(cyl_ctx* ctx) {
  if (ctx->fast_ready) return CYL_OK;
  const MeshDev& m = ctx->mesh;
  ElemF* ef;
  int4 *en, *te, *ed;
  double2* xy;
  HIPCHK(dalloc(ctx->fast_bufs, &ef, m.nt));
  HIPCHK(dalloc(ctx->fast_bufs, &en, m.nt));
  HIPCHK(dalloc(ctx->fast_bufs, &te, m.nt));
  HIPCHK(dalloc(ctx->fast_bufs, &ed, m.ne));
  HIPCHK(dalloc(ctx->fast_bufs, &xy, m.nv));
  WalkRec* wr;
  HIPCHK(dalloc(ctx->fast_bufs, &wr, m.nt));
  HIPCHK(dalloc(ctx->fast_bufs, &ctx->d_nodes, m.nv));
  if (ctx->fields.dim == 3) HIPCHK(dalloc(ctx->fast_bufs, &ctx->d_nodes3, (size_t)m.nv * ctx->fields.nlayers));
  k_pack_elems<<<nblk(m.nt), BLK, 0, ctx->stream>>>(m, ctx->d_node_new, ef, en, te);
  k_pack_edges<<<nblk(std::max(m.ne, m.nv)), BLK, 0, ctx->stream>>>(m, ctx->d_node_new, ctx->d_node_old, ed, xy);
  k_pack_walk<<<nblk(m.nt), BLK, 0, ctx->stream>>>(m, wr);
  HIPCHK(hipGetLastError());
  ctx->fm.ef = ef;
  ctx->fm.wr = wr;
  ctx->fm.en = en;
  ctx->fm.te = te;
  ctx->fm.ed = ed;
  ctx->fm.xy = xy;
  ctx->fm.edge_labels = m.edge_labels;
  ctx->fm.nt = m.nt;
  ctx->fast_ready = true;
  ctx->pack_stamp = -1;
  return CYL_OK;
}

static int ensure_start_bufs(cyl_ctx* ctx) {
  if (ctx->start_cap >= ctx->n) return CYL_OK;
  if (ctx->d_pos0) HIPCHK(hipFree(ctx->d_pos0));
  if (ctx->d_state0) HIPCHK(hipFree(ctx->d_state0));
  ctx->d_pos0 = nullptr;
  ctx->d_state0 = nullptr;
  ctx->start_cap = 0;
  HIPCHK(hipMalloc(&ctx->d_pos0, (size_t)ctx->n * sizeof(int64_t)));
  HIPCHK(hipMalloc(&ctx->d_state0, (size_t)ctx->n * sizeof(uint64_t)));
  ctx->start_cap = ctx->n;
  return CYL_OK;
}

// arguments of the fused kernel over slots [j0, j0 + n)
static FastArgs make_fast_args(cyl_ctx* ctx, const StepArgs& a, int j0, int n) {
  FastArgs f{};
  f.fm = ctx->fm;
  f.m = ctx->mesh;
  f.nodes = ctx->d_nodes;
  f.xA = a.A.x;
  f.yA = a.A.y;
  f.triA = a.A.tri;
  f.upA0 = a.A.up[0];
  f.upA1 = a.A.up[1];
  f.usA0 = a.A.us[0];
  f.usA1 = a.A.us[1];
  f.xB = a.B.x;
  f.yB = a.B.y;
  f.triB = a.B.tri;
  f.upB0 = a.B.up[0];
  f.upB1 = a.B.up[1];
  f.usB0 = a.B.us[0];
  f.usB1 = a.B.us[1];
  f.inactive = ctx->pool[0].inactive;
  f.del_step = ctx->d_del_step;
  f.diameter = ctx->d_diam;
  f.pc = ctx->pc;
  f.rank = ctx->d_rank;
  f.off = ctx->d_off;
  f.pool = ctx->d_gpool;
  f.scen = ctx->d_scen;
  f.j0 = j0;
  f.n = n;
  f.nscen = ctx->nscen;
  f.dt = a.dt;
  f.sqrtdt = sqrt(a.dt);
  f.fsc = ctx->d_fsc;
  f.scen_end = a.scen_end;
  f.scen_base = a.scen_base;
  f.clock = a.clock;
  f.nopen = ctx->nopen;
  f.nopen_stride = std::max(ctx->nopen, 1);
  f.general = a.general;
  f.guarded = a.guarded;
  f.flag = a.flag;
  f.stream = a.stream;
  f.jt = a.jt;
  f.jt8 = a.jt8;
  f.zt = a.zt;
  f.bnd_stats = a.bnd_stats;
  f.n_deleted = a.n_deleted;
  f.err = a.err;
  f.stream_end = a.stream_end;
  f.step_id = a.step_id;
  f.batch_step = a.batch_step;
  f.pos0 = ctx->d_pos0;
  f.state0 = ctx->d_state0;
  f.xi = ctx->d_xi;
  f.xi_cap = ctx->xi_cap;
  f.nact = ctx->d_scen_active;
  f.chunk_base = ctx->d_gen_base;
  f.gen_shift = ctx->dim == 3 ? GEN_SHIFT_3D : GEN_SHIFT_2D;
  f.pregen = ctx->d_pregen;
  f.xi_by_slot = ctx->inline_normals == 2 ? 1 : 0;
  f.skip_us = !ctx->batched && ctx->dim == 2 ? 1 : 0;  // read by the LSM-1 kernel only
  f.rec_xyz = a.strand_xyz;  // set for steps that record deposits (or stranded rows)
  f.rec_key = a.strand_key;
  f.rec_count = a.strand_count;
  f.rec_cap = a.strand_cap;
  return f;
}

// persistent grid: as many blocks as fit on the device at once (a multiple of 8), at most the work
static unsigned persistent_grid(cyl_ctx* ctx, const void* fn, int64_t nblocks, bool mult8) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, BLK, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  if (ctx->blocks_per_cu > 0) per_cu = std::min(per_cu, ctx->blocks_per_cu);
  int64_t g = std::min<int64_t>(std::max<int64_t>(nblocks, 1), (int64_t)ctx->ncu * per_cu);
  return mult8 ? (unsigned)std::max<int64_t>(8, g / 8 * 8) : (unsigned)g;
}

// the fused kernels run one slot per thread: 8 * ceil(chunks / 8) blocks (xcd_chunk)
static unsigned fast_grid(int64_t nblocks) { return (unsigned)(std::max<int64_t>(1, (nblocks + 7) / 8) * 8); }

// normals of the step for every scenario of the fused kernel (uniform assignment)
static void launch_gen(cyl_ctx* ctx, const FastArgs& f0, int m) {
  if (ctx->inline_normals == 2) {
    // slot order: every slot of the pool (the scenarios' pregen flags select the rows)
    FastArgs f = f0;
    f.j0 = 0;
    f.n = ctx->n;
    int64_t nb = nblk(f.n);
    if (m == 1) k_gen_slots<1><<<persistent_grid(ctx, (const void*)k_gen_slots<1>, nb, true), BLK, 0, ctx->stream>>>(f);
    else if (m == 2) k_gen_slots<2><<<persistent_grid(ctx, (const void*)k_gen_slots<2>, nb, true), BLK, 0, ctx->stream>>>(f);
    else k_gen_slots<3><<<persistent_grid(ctx, (const void*)k_gen_slots<3>, nb, true), BLK, 0, ctx->stream>>>(f);
    return;
  }
  const FastArgs& f = f0;
  const int gsh = ctx->dim == 3 ? GEN_SHIFT_3D : GEN_SHIFT_2D;
  int64_t nb = nblk((((int64_t)ctx->n + (1 << gsh) - 1) >> gsh) + ctx->nscen);
  hipStream_t s = ctx->stream;
#define CYL_GEN_LAUNCH(R, NV) \
  k_gen_normals<R, NV><<<persistent_grid(ctx, (const void*)k_gen_normals<R, NV>, nb, false), BLK, 0, s>>>(f)
  // three values per rank for the 3D kernel (pair and vertical single), a pair in 2D
  if (ctx->dim == 3) {
    if (m == 1) CYL_GEN_LAUNCH(1, 3);
    else if (m == 2) CYL_GEN_LAUNCH(2, 3);
    else CYL_GEN_LAUNCH(3, 3);
  } else {
    if (m == 1) CYL_GEN_LAUNCH(1, 2);
    else if (m == 2) CYL_GEN_LAUNCH(2, 2);
    else CYL_GEN_LAUNCH(3, 2);
  }
#undef CYL_GEN_LAUNCH
}

// stream position and generator state of every slot's first draw (the inline
// samplers of the fused kernels)
static void launch_draw_start(cyl_ctx* ctx, const FastArgs& f, int m) {
  int64_t nb = nblk(f.n);
  hipStream_t s = ctx->stream;
  if (m == 1) k_draw_start<1><<<persistent_grid(ctx, (const void*)k_draw_start<1>, nb, true), BLK, 0, s>>>(f);
  else if (m == 2) k_draw_start<2><<<persistent_grid(ctx, (const void*)k_draw_start<2>, nb, true), BLK, 0, s>>>(f);
  else k_draw_start<3><<<persistent_grid(ctx, (const void*)k_draw_start<3>, nb, true), BLK, 0, s>>>(f);
}

// the fused kernel of `model` over f's slot range: pre-generated normals
// (uniform), or prepared stream positions and the inline sampler (general)
template <int MODEL>
static void launch_fast_model(cyl_ctx* ctx, const FastArgs& f, int m, bool general, bool start_done) {
  int64_t nb = nblk(f.n);
  hipStream_t s = ctx->stream;
#define CYL_FAST_LAUNCH(R, P)                                                                           \
  do {                                                                                                  \
    unsigned g = fast_grid(nb);                                                                         \
    if (MODEL == 1) k_lsm1_2d<R, P><<<g, BLK, 0, s>>>(f);                                               \
    else k_lsm2_2d<R, P><<<g, BLK, 0, s>>>(f);                                                          \
  } while (0)
  if (!general) {
    if (m == 1) CYL_FAST_LAUNCH(1, true);
    else if (m == 2) CYL_FAST_LAUNCH(2, true);
    else CYL_FAST_LAUNCH(3, true);
    return;
  }
  if (!start_done) launch_draw_start(ctx, f, m);
  if (m == 1) CYL_FAST_LAUNCH(1, false);
  else if (m == 2) CYL_FAST_LAUNCH(2, false);
  else CYL_FAST_LAUNCH(3, false);
#undef CYL_FAST_LAUNCH
}

// the fused 3D kernel: pre-generated normals (uniform assignment), or the
// inline sampler (its slots' stream positions prepared by launch_draw_start,
// here unless start_done)
template <int MODEL>
static void launch_fast3(cyl_ctx* ctx, const Fast3Args& f3, int m, bool inl, bool start_done) {
  int64_t nb = nblk(f3.f.n);
  hipStream_t s = ctx->stream;
  if (!inl) {
    if (m == 1) k_fast_3d<MODEL, 1, true><<<fast_grid(nb), BLK, 0, s>>>(f3);
    else if (m == 2) k_fast_3d<MODEL, 2, true><<<fast_grid(nb), BLK, 0, s>>>(f3);
    else k_fast_3d<MODEL, 3, true><<<fast_grid(nb), BLK, 0, s>>>(f3);
    return;
  }
  if (!start_done) launch_draw_start(ctx, f3.f, m);
  if (m == 1) k_fast_3d<MODEL, 1, false><<<fast_grid(nb), BLK, 0, s>>>(f3);
  else if (m == 2) k_fast_3d<MODEL, 2, false><<<fast_grid(nb), BLK, 0, s>>>(f3);
  else k_fast_3d<MODEL, 3, false><<<fast_grid(nb), BLK, 0, s>>>(f3);
}

static Fast3Args make_fast3_args(cyl_ctx* ctx, const FastArgs& f, const StepArgs& a) {
  Fast3Args r{};
  r.f = f;
  r.nodes3 = ctx->d_nodes3;
  r.nv = ctx->fields.npoints;
  r.nl = ctx->fields.nlayers;
  r.zA = a.A.z;
  r.zB = a.B.z;
  r.llA = a.A.ll;
  r.llB = a.B.ll;
  r.upA2 = a.A.up[2];
  r.upB2 = a.B.up[2];
  r.usA2 = a.A.us[2];
  r.usB2 = a.B.us[2];
  return r;
}

// whether the fused kernels of this context sample their normals inline
// (from stream positions prepared by launch_draw_start) under the uniform
// assignment: when the generator is switched off (the slot-order generator
// serves the 2D kernels only)
static bool samples_inline(const cyl_ctx* ctx) {
  return ctx->inline_normals == 1 || (ctx->dim == 3 && ctx->inline_normals == 2);
}

// start_done: the uniform assignment's draw starts were prepared before the
// timed kernel window (general-mode reruns prepare their own)
static void launch_fast_main(cyl_ctx* ctx, const FastArgs& f, int model, int m, bool general,
                             const StepArgs* a = nullptr, bool start_done = false) {
  const bool inl = general || samples_inline(ctx);
  if (ctx->dim == 3) {
    Fast3Args f3 = make_fast3_args(ctx, f, *a);
    if (model == 1) launch_fast3<1>(ctx, f3, m, inl, start_done);
    else launch_fast3<2>(ctx, f3, m, inl, start_done);
    return;
  }
  if (model == 1) launch_fast_model<1>(ctx, f, m, inl, start_done);
  else launch_fast_model<2>(ctx, f, m, inl, start_done);
}

static int enqueue_sort(cyl_ctx* ctx, bool need_us, int& launches);
static int ensure_sort_bufs(cyl_ctx* ctx);

static int run_steps(cyl_ctx* ctx, const cyl_params* p, double* time_inout, double dt, int nsteps, int* bnd_out,
                     bool init_only) {
  if (!ctx->has_mesh || !ctx->has_fields || ctx->n == 0) return ctx->set_err(CYL_E_STATE, -1, "context not ready");
  const int ns = ctx->nscen;
  std::vector<cyl_params> P(ns);
  if (ctx->batched) P = ctx->scen_par;
  else P[0] = *p;
  for (int s = 0; s < ns; ++s) {
    if (P[s].model < 0 || P[s].model > 3) return ctx->set_err(CYL_E_ARG, -1, "Wrong model");
    if (P[s].model == 3 && (P[s].lsm3_edge_opt < 1 || P[s].lsm3_edge_opt > 2))
      return ctx->set_err(CYL_E_ARG, -1, "lsm3_edge_opt must be 1 or 2");
    if (P[s].rng_method < 1 || P[s].rng_method > 3) return ctx->set_err(CYL_E_ARG, -1, "Unknown RNG method");
    if (P[s].rng_method != P[0].rng_method)
      return ctx->set_err(CYL_E_ARG, -1, "the scenarios of a batch share one rng_method");
  }
  const int method = P[0].rng_method;
  HIPCHK(hipSetDevice(ctx->device));
  auto t_start = std::chrono::steady_clock::now();
  // uniform draw prediction per scenario
  int umax = 0;
  bool can_uniform = !init_only;
  std::vector<int> D(ns, 0);
  for (int s = 0; s < ns; ++s) {
    int um = 0;
    D[s] = predict_units(&P[s], ctx->dim, dt, um);
    umax = std::max(umax, um);
    if (D[s] < 0) can_uniform = false;
  }
  if (!can_uniform) std::fill(D.begin(), D.end(), 0);
  int launches = 0;
  int e = upload_streams(ctx, D);
  if (e) return e;
  e = ensure_events(ctx, (int64_t)umax * ctx->n, method);
  if (e) return e;
  e = ensure_start_bufs(ctx);
  if (e) return e;
  // per-scenario device parameters
  {
    std::vector<ParamsDev> pd(ns);
    std::vector<FastScen> fs(ns);
    std::vector<int> send(ns);
    for (int s = 0; s < ns; ++s) {
      pd[s] = to_dev(&P[s]);
      fs[s].diffusion_model = P[s].diffusion_model;
      fs[s].diff_const = sqrt(2. * P[s].horizontal_diffusivity) * sqrt(dt);  // diffusion_lsm1, update_state_lsm1.pyx:700
      fs[s].diff_const_z = sqrt(2. * P[s].vertical_diffusivity) * sqrt(dt);
      fs[s].boundary_conditions = P[s].boundary_conditions;
      fs[s].max_wall_rebounds = P[s].max_wall_rebounds;
      fs[s].bnd_statistics = P[s].bnd_statistics;
      fs[s].damp = fmin(P[s].rebound_damping_coef, 1.0 - EPS_REBOUND);
      fs[s].strand = P[s].stranding_probability;
      fs[s].deposition = P[s].bottom_deposition;
      send[s] = ctx->batched ? ctx->scen_nsteps[s] : 0;
    }
    HIPCHK(hipMemcpy(ctx->d_pars, pd.data(), ns * sizeof(ParamsDev), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_fsc, fs.data(), ns * sizeof(FastScen), hipMemcpyHostToDevice));
    if (ctx->batched) HIPCHK(hipMemcpy(ctx->d_scen_end, send.data(), ns * sizeof(int), hipMemcpyHostToDevice));
  }
  int nstat = ns * std::max(ctx->nopen, 1);
  size_t need_bnd = (size_t)std::max(nsteps, 1) * nstat;
  if (need_bnd > ctx->bnd_cap) {
    if (ctx->d_bnd) HIPCHK(hipFree(ctx->d_bnd));
    HIPCHK(hipMalloc(&ctx->d_bnd, need_bnd * sizeof(int)));
    ctx->bnd_cap = need_bnd;
  }
  HIPCHK(hipMemsetAsync(ctx->d_bnd, 0, need_bnd * sizeof(int), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_err, 0xff, ns * sizeof(unsigned long long), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_flag + 2, 0, sizeof(int), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_flag + 4, 0, sizeof(int), ctx->stream));
  if (ctx->nsecseg > 0) {
    size_t need = (size_t)std::max(nsteps, 1) * ctx->nsec;
    if (need > ctx->sec_cap) {
      if (ctx->d_sec_counts) HIPCHK(hipFree(ctx->d_sec_counts));
      HIPCHK(hipMalloc(&ctx->d_sec_counts, need * sizeof(int)));
      ctx->sec_cap = need;
    }
    HIPCHK(hipMemsetAsync(ctx->d_sec_counts, 0, need * sizeof(int), ctx->stream));
  }
  // stranded positions and deposits on the bed (one record per row at most:
  // a stranded or deposited row is deleted)
  const bool deposit_rec = !init_only && !ctx->batched && ctx->dim == 3 && P[0].bottom_deposition &&
                           P[0].boundary_conditions;
  bool strand_rec =
      (!init_only && !ctx->batched && P[0].stranding_output && P[0].stranding_probability > 0.) || deposit_rec;
  const int step0 = ctx->step_counter;
  ctx->last_strand.clear();
  ctx->last_deposit.clear();
  ctx->last_deposit_step.clear();
  if (strand_rec) {
    if (ctx->strand_cap < ctx->n) {
      if (ctx->d_strand_xyz) HIPCHK(hipFree(ctx->d_strand_xyz));
      if (ctx->d_strand_key) HIPCHK(hipFree(ctx->d_strand_key));
      ctx->d_strand_xyz = nullptr;
      ctx->d_strand_key = nullptr;
      HIPCHK(hipMalloc(&ctx->d_strand_xyz, (size_t)ctx->n * 3 * sizeof(double)));
      HIPCHK(hipMalloc(&ctx->d_strand_key, (size_t)ctx->n * sizeof(int2)));
      ctx->strand_cap = ctx->n;
    }
    if (!ctx->d_strand_count) HIPCHK(hipMalloc(&ctx->d_strand_count, 2 * sizeof(int)));
    HIPCHK(hipMemsetAsync(ctx->d_strand_count, 0, 2 * sizeof(int), ctx->stream));
  }
  bool any_stats = false;
  for (int s = 0; s < ns; ++s) any_stats |= P[s].bnd_statistics != 0;
  // per-step records (statistics, control sections, stranded rows): eager steps only
  const bool per_step_records = any_stats || ctx->nsecseg > 0 || strand_rec;
  StepArgs a{};
  a.m = ctx->mesh;
  a.par = to_dev(&P[0]);
  a.pars = ctx->d_pars;
  a.scen_end = ctx->batched ? ctx->d_scen_end : nullptr;
  a.scen_base = ctx->batched ? ctx->d_scen_base : nullptr;
  a.pc = ctx->pc;
  a.dim = ctx->dim;
  a.dt = dt;
  a.stream = ctx->d_stream;
  a.jt = ctx->d_jt;
  a.jt8 = ctx->d_jt8;
  a.zt = ctx->d_zt;
  a.nopen = std::max(ctx->nopen, 1);
  a.nopen_real = ctx->nopen;
  a.n_deleted = ctx->d_n_deleted;
  a.err = ctx->d_err_step;
  a.stream_end = ctx->d_stream_end;
  a.flag = ctx->d_flag;
  a.units_pool = ctx->d_vals;
  a.init_only = init_only ? 1 : 0;
  a.slot.off = ctx->d_off;
  a.slot.scen = ctx->d_scen;
  double t = *time_inout;
  int* one = ctx->d_flag + 3;
  float kms = 0.f;
  int n_timed = 0, nfast = 0, ngeneric = 0;
  // Enqueue one step.  cap: the stream is being captured into a graph (no
  // event records, no host synchronisation, step counters read on the device).
  auto one_step = [&](int step, bool cap) -> int {
    // kernel classes of this step: slot ranges stepped by the fused or the generic kernel
    std::vector<cyl_ctx::KClass> cls;
    if (!ctx->batched) {
      cls.push_back({fast_eligible(ctx, p, init_only), p->model, 0, ctx->n});
    } else {
      cls = ctx->classes;
      if (init_only || ctx->pending_init)
        for (auto& c : cls) c.fast = false;
    }
    bool any_fast = false, all_fast = true;
    for (auto& c : cls) {
      any_fast |= c.fast;
      all_fast &= c.fast;
    }
    // draws_fixed: every active row draws exactly the predicted units, so the
    // general-mode rerun can never be needed.  Holds when all rows run the
    // fused kernels (their only draws are the diffusion pair, and the vertical
    // single in 3D, skipped only when the amplitude is zero) and each amplitude
    // is a constant or sqrt(2 K dt) with K bounded away from zero: nodal k > 0
    // in every frame, and nodal K spread by less than 1e6, so that a P1 value
    // inside the start element (weights >= -1e-10 after rounding) stays positive.
    bool draws_fixed = all_fast && !init_only && !ctx->pending_init && ctx->skip_checked_rerun;
    for (int s = 0; s < ns && draws_fixed; ++s) {
      if (P[s].diffusion_model >= 2) {
        double fac = CMU / P[s].schmidt_number;
        double klo = fac * ctx->keps_lo, khi = fac * ctx->keps_hi;
        draws_fixed = ctx->keps_ok && std::isfinite(khi) && klo > 1e-250 && khi < 1e6 * klo && dt > 1e-100;
      }
    }
    {
      std::vector<int> pregen(ns, 0);
      // the normal generator serves the fused kernels unless they sample inline
      const bool gen = !samples_inline(ctx);
      if (!ctx->batched) pregen[0] = cls[0].fast && gen ? 1 : 0;
      else
        for (int s = 0; s < ns; ++s) pregen[s] = ctx->classes_fast_of_scen(s) && any_fast && gen ? 1 : 0;
      if (pregen != ctx->pregen_h) {  // changes only when arrivals switch the fused kernel off or on
        if (cap) return ctx->set_err(CYL_E_STATE, -1, "generator flags changed inside a captured step");
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipMemcpy(ctx->d_pregen, pregen.data(), ns * sizeof(int), hipMemcpyHostToDevice));
        ctx->pregen_h = pregen;
      }
    }
    // us left unstored by 2D LSM-1 steps is materialized (= up) before a step
    // of another kind, which may read it
    const bool lsm1_fused_2d = !ctx->batched && !init_only && cls[0].fast && p->model == 1 && ctx->dim == 2;
    if (ctx->us_from_up && !ctx->batched && !lsm1_fused_2d && !(p->model == 1 && ctx->dim == 2)) {
      const PoolDev& C = ctx->pool[ctx->cur];
      for (int c = 0; c < ctx->dim; ++c)
        HIPCHK(hipMemcpyAsync(C.us[c], C.up[c], (size_t)ctx->n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      ctx->us_from_up = false;
    }
    bool do_sort = !init_only && ctx->sort_interval > 0 && ctx->steps_since_sort >= ctx->sort_interval;
    if (do_sort) {
      // the fused kernels rewrite us for every row they step; rows of finished
      // scenarios and generic classes carry it over
      e = enqueue_sort(ctx, !all_fast || ctx->batched, launches);
      if (e) return e;
      ctx->steps_since_sort = 0;
    }
    // the fused kernels' node records: k-epsilon K when a fused scenario uses it
    bool keps = false;
    double sigc = P[0].schmidt_number;
    for (int s = 0; s < ns; ++s)
      if (!ctx->batched || ctx->classes_fast_of_scen(s)) {
        keps |= P[s].diffusion_model >= 2;
        sigc = P[s].schmidt_number;
      }
    // 2D fields that change every step: the records are written by the time
    // interpolation itself (the nodes' values are in registers there)
    PackReq pack;
    pack.on = any_fast && !init_only && !ctx->frozen && ctx->dim == 2 && ctx->fields.nlayers == 1;
    pack.keps = keps;
    pack.fac = CMU / sigc;
    if (pack.on) {
      e = ensure_fast(ctx);
      if (e) return e;
    }
    if (!init_only) {
      e = enqueue_time_interp(ctx, t, launches, cap ? ctx->d_clock : nullptr, pack);
      if (e) return e;
    }
    // per-slot arrays (their buffers are swapped by the sort)
    a.init_tri = ctx->d_init_tri;
    a.del_step = ctx->d_del_step;
    a.depth = ctx->dim == 3 ? ctx->d_depth : nullptr;
    a.gpool = ctx->d_gpool;
    a.slot.pool = ctx->d_gpool;
    a.slot.rank = ctx->d_rank;
    a.slot.scen = ctx->d_scen;
    a.slot.diameter = ctx->d_diam;
    a.slot.volume = ctx->d_vol;
    a.f = ctx->fields;
    if (any_fast) {
      e = ensure_fast(ctx);
      if (e) return e;
      const FieldsDev& fd = ctx->fields;
      // frozen fields: the records packed at an earlier step are still the step's
      const bool fresh = pack.on || (ctx->frozen && ctx->pack_stamp == ctx->field_stamp &&
                                     ctx->packed_keps == (int)keps && ctx->packed_sigc == sigc);
      if (!fresh) {
        if (ctx->dim == 3) {
          int64_t nrec = (int64_t)fd.npoints * fd.nlayers;
          k_pack_nodes3<<<nblk(nrec), BLK, 0, ctx->stream>>>(
              fd.tmp[fd.fid_zs], fd.tmp[fd.fid_ux], fd.tmp[fd.fid_uy], fd.tmp[fd.fid_uz],
              keps ? fd.tmp[fd.fid_turbeng] : nullptr, keps ? fd.tmp[fd.fid_dissip] : nullptr, CMU / sigc,
              fd.npoints, fd.nlayers, ctx->d_node_old, ctx->d_nodes3);
        } else {
          k_pack_nodes<<<nblk(fd.npoints), BLK, 0, ctx->stream>>>(
              fd.tmp[fd.fid_ux], fd.tmp[fd.fid_uy], keps ? fd.tmp[fd.fid_turbeng] : nullptr,
              keps ? fd.tmp[fd.fid_dissip] : nullptr, CMU / sigc, fd.npoints, ctx->d_node_old, ctx->d_nodes);
        }
        ++launches;
        ctx->pack_stamp = ctx->field_stamp;
        ctx->packed_keps = (int)keps;
        ctx->packed_sigc = sigc;
      }
    }
    a.A = ctx->pool[ctx->cur];
    a.B = ctx->pool[1 - ctx->cur];
    a.step_id = ctx->step_counter;
    a.batch_step = ctx->batch_steps;
    a.bnd_stats = any_stats ? ctx->d_bnd + (size_t)step * nstat : nullptr;
    a.strand_xyz = strand_rec ? ctx->d_strand_xyz : nullptr;
    a.strand_key = strand_rec ? ctx->d_strand_key : nullptr;
    a.strand_count = strand_rec ? ctx->d_strand_count : nullptr;
    a.strand_cap = ctx->strand_cap;
    // active rank when rows were deleted or added since the last scan
    enqueue_rank(ctx, ctx->d_flag + 1, launches);
    bool uniform = can_uniform && !ctx->pending_init;
    {
      // scenario totals (after a rank scan), the scan request cleared, units
      // of the step and generator chunks: one block
      PrepArgs pa{};
      pa.scan = ctx->d_scan;
      pa.total = ctx->d_total;
      pa.scen_base = ctx->d_scen_base;
      pa.nscen = ns;
      pa.npool = ctx->n;
      pa.rank_flag = ctx->d_flag + 1;
      pa.nact = ctx->d_scen_active;
      pa.st = ctx->d_stream;
      pa.scen_end = a.scen_end;
      pa.batch_step = a.batch_step;
      pa.clock = a.clock;
      pa.scen_units = ctx->d_scen_units;
      pa.uniform = uniform ? 1 : 0;
      pa.pregen = ctx->d_pregen;
      pa.chunk_base = uniform && any_fast && ctx->inline_normals == 0 ? ctx->d_gen_base : nullptr;
      pa.gen_shift = ctx->dim == 3 ? GEN_SHIFT_3D : GEN_SHIFT_2D;
      if (uniform && method != 1) {
        pa.layout = 1;
        pa.ev = layout_args(ctx, method);
      }
      k_step_prep<<<1, BLK, 0, ctx->stream>>>(pa);
      ++launches;
    }
    if (step == 0 && !cap) HIPCHK(hipEventRecord(ctx->ev0, ctx->stream));
    if (uniform) {
      enqueue_events(ctx, method, one, false, launches);
      a.general = 0;
      a.guarded = 0;
      const int ti = n_timed;  // event slots of the timed (eager) steps
      if (!cap) {
        while (ctx->kev.size() < (size_t)3 * (ti + 1)) {
          hipEvent_t ev;
          HIPCHK(hipEventCreate(&ev));
          ctx->kev.push_back(ev);
        }
        HIPCHK(hipEventRecord(ctx->kev[3 * ti], ctx->stream));
      }
      // normals of the fused kernels: generated, or the inline samplers' draw starts
      if (any_fast && !samples_inline(ctx)) launch_gen(ctx, make_fast_args(ctx, a, 0, 0), method);
      if (any_fast && samples_inline(ctx))
        for (auto& c : cls)
          if (c.fast && c.hi > c.lo) launch_draw_start(ctx, make_fast_args(ctx, a, c.lo, c.hi - c.lo), method);
      if (!cap) HIPCHK(hipEventRecord(ctx->kev[3 * ti + 1], ctx->stream));
      for (auto& c : cls) {
        if (c.hi <= c.lo) continue;
        if (c.fast) {
          launch_fast_main(ctx, make_fast_args(ctx, a, c.lo, c.hi - c.lo), c.model, method, false, &a, true);
          ++nfast;
        } else {
          a.j0 = c.lo;
          a.n = c.hi - c.lo;
          launch_step(ctx, a, c.model, false);
          ++ngeneric;
        }
        ++launches;
      }
      if (!cap) {
        HIPCHK(hipEventRecord(ctx->kev[3 * ti + 2], ctx->stream));
        ++n_timed;
      }
    } else {
      HIPCHK(hipMemsetAsync(ctx->d_flag, 0xff, sizeof(int), ctx->stream));  // force the general path
    }
    // general mode: exact per-row unit counts, pool-order scan, rerun.  Not
    // enqueued when every row is known to draw the predicted units (the
    // draws_fixed proof); k_finish then reports any misprediction in flag[4].
    const bool skip_general = uniform && draws_fixed;
    int* g = ctx->d_flag;
    if (!skip_general) {
    k_rerun_reset<<<nblk(std::max(nstat, ns)), BLK, 0, ctx->stream>>>(
        ctx->d_stream_end, ctx->d_err_step, a.bnd_stats, nstat, ns, ctx->d_n_deleted, g, a.strand_count);
    const unsigned gg = std::min<unsigned>(nblk(ctx->n), GUARD_GRID);
    k_undelete<<<gg, BLK, 0, ctx->stream>>>(ctx->pool[0].inactive, ctx->d_del_step, ctx->n,
                                                      ctx->step_counter, g, a.clock);
    a.general = 0;
    a.guarded = 1;
    for (auto& c : cls) {
      if (c.hi <= c.lo) continue;
      a.j0 = c.lo;
      a.n = c.hi - c.lo;
      launch_step(ctx, a, c.model, true);
    }
    enqueue_scan(ctx, ctx->d_vals, ctx->d_scan, (int64_t)ctx->n, ctx->d_bsum, ctx->d_total, (const int64_t*)nullptr,
                 g, launches);
    k_gather_rank<<<gg, BLK, 0, ctx->stream>>>(ctx->d_scan, ctx->d_gpool, ctx->d_scen, ctx->d_scen_base,
                                                         ctx->n, nullptr, ctx->d_off, g);
    k_scen_totals<<<nblk(ns), BLK, 0, ctx->stream>>>(ctx->d_scan, ctx->d_total, ctx->d_scen_base, ns, ctx->n,
                                                     ctx->d_scen_units, g);
    launches += 4;
    enqueue_events(ctx, method, g, true, launches);
    a.general = 1;
    a.guarded = 1;
    for (auto& c : cls) {
      if (c.hi <= c.lo) continue;
      if (c.fast) {
        launch_fast_main(ctx, make_fast_args(ctx, a, c.lo, c.hi - c.lo), c.model, method, true, &a);
        if (!uniform) ++nfast;
      } else {
        a.j0 = c.lo;
        a.n = c.hi - c.lo;
        launch_step(ctx, a, c.model, false);
        if (!uniform) ++ngeneric;
      }
      ++launches;
    }
    }  // !skip_general
    if (ctx->nsecseg > 0 && !init_only) {
      const PoolDev& PA = ctx->pool[ctx->cur];
      const PoolDev& PB = ctx->pool[1 - ctx->cur];
      k_control_sections<<<nblk(ctx->n), BLK, 0, ctx->stream>>>(
          PA.x, PA.y, PB.x, PB.y, ctx->pool[0].inactive, ctx->d_del_step, ctx->step_counter, ctx->n, ctx->d_sec_seg,
          ctx->d_sec_of, ctx->nsecseg, ctx->d_sec_counts + (size_t)step * ctx->nsec);
      ++launches;
    }
    k_finish<<<nblk(ns), BLK, 0, ctx->stream>>>(ctx->d_stream, ctx->d_jt, ctx->d_scen_units, method, ns, ctx->d_flag,
                                                ctx->d_n_deleted, ctx->d_n_deleted_total, ctx->step_counter,
                                                ctx->d_step_done, ctx->d_err_step, ctx->d_err,
                                                a.clock ? ctx->d_clock : nullptr, a.strand_count,
                                                skip_general ? ctx->d_flag + 4 : nullptr);
    ++launches;
    ctx->cur = 1 - ctx->cur;
    ctx->have_last = !init_only;
    if (lsm1_fused_2d) ctx->us_from_up = true;
    ctx->pending_init = false;
    ++ctx->step_counter;
    ++ctx->steps_since_sort;
    if (!init_only) {
      t += dt;  // SimuTime.increment, simutime.pyx:81
      ++ctx->batch_steps;
    }
    return CYL_OK;
  };
  int step = 0;
  ctx->last_graph_steps = 0;
  // Graph replay of the loop: a period of 2 * sort_interval steps returns the
  // state copies, the spare buffers, the sort phase and the interpolated
  // fields' buffer parity to where they were, so one captured period is
  // launched repeatedly; the step counters live on the device, and with
  // time-varying fields each step's record and weight come from a table the
  // host computes for the call.  Needs the uniform draw assignment and no
  // per-step statistics.
  const int period = 2 * std::max(1, ctx->sort_interval);
  if (ctx->use_graph && !init_only && can_uniform && !per_step_records && nsteps >= 1 + 2 * period) {
    // one eager step first: arrivals initialized, generator flags and packed records in place
    e = one_step(step++, false);
    if (e) return e;
    e = ensure_sort_bufs(ctx);
    if (e) return e;
    e = ensure_fast(ctx);  // no allocation or packing inside the capture
    if (e) return e;
    const int nrep = (nsteps - step) / period;
    // time-varying fields: the record and weight of every replayed step, from
    // the host's own interpolation state (a step past the last record leaves
    // the steps to the eager loop, which reports it where it happens)
    bool tab_ok = true;
    if (!ctx->frozen && !ctx->pending_init) {
      const int64_t ntab = (int64_t)nrep * period;
      std::vector<InterpStep> tab((size_t)ntab);
      int rec = ctx->record;
      double tt = t;
      for (int64_t q = 0; q < ntab && tab_ok; ++q) {
        rec = lower_index(ctx->times, tt, rec);
        double dtt = rec + 1 < ctx->ntimes ? ctx->times[rec] - ctx->times[rec + 1] : 0.;
        tab_ok = rec + 1 < ctx->ntimes && std::fabs(dtt) >= EPS;
        if (tab_ok) tab[(size_t)q] = InterpStep{(tt - ctx->times[rec + 1]) / dtt, (int64_t)rec};
        tt += dt;
      }
      if (tab_ok) {
        if (ntab > ctx->interp_tab_cap) {
          if (ctx->d_interp_tab) HIPCHK(hipFree(ctx->d_interp_tab));
          ctx->d_interp_tab = nullptr;
          HIPCHK(hipMalloc(&ctx->d_interp_tab, (size_t)ntab * sizeof(InterpStep)));
          ctx->interp_tab_cap = ntab;
        }
        HIPCHK(hipMemcpy(ctx->d_interp_tab, tab.data(), (size_t)ntab * sizeof(InterpStep), hipMemcpyHostToDevice));
        ctx->interp_step0 = ctx->step_counter;
      }
    }
    if (!ctx->pending_init && tab_ok) {
      HIPCHK(hipStreamSynchronize(ctx->stream));
      int clk[2] = {ctx->step_counter, ctx->batch_steps};
      HIPCHK(hipMemcpy(ctx->d_clock, clk, sizeof(clk), hipMemcpyHostToDevice));
      a.clock = ctx->d_clock;
      hipGraph_t graph = nullptr;
      hipGraphExec_t exec = nullptr;
      HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      int ec = CYL_OK;
      for (int q = 0; q < period && !ec; ++q) ec = one_step(step + q, true);
      hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
      a.clock = nullptr;
      if (ec) {
        if (graph) (void)hipGraphDestroy(graph);
        return ec;
      }
      HIPCHK(ce);
      hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
      }
      // the capture advanced the host-side state by one period; each further
      // replay advances the counters and the clock by another
      for (int r = 0; r < nrep; ++r) {
        hipError_t le = hipGraphLaunch(exec, ctx->stream);
        if (le != hipSuccess) {
          (void)hipGraphExecDestroy(exec);
          (void)hipGraphDestroy(graph);
          HIPCHK(le);
        }
        if (r > 0) {
          ctx->step_counter += period;
          ctx->batch_steps += period;
          for (int q = 0; q < period; ++q) {
            if (!ctx->frozen) {
              double tf = 0.;
              int ie = time_interp_state(ctx, t, tf);  // the host's record of the replayed interpolation
              if (ie) return ie;
            }
            t += dt;
          }
        }
      }
      step += nrep * period;
      ctx->last_graph_steps = nrep * period;
      HIPCHK(hipStreamSynchronize(ctx->stream));
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
    }
  }
  for (; step < nsteps; ++step) {
    e = one_step(step, false);
    if (e) return e;
  }
  if (nsteps > 0) HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  if (nsteps > 0) HIPCHK(hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1));
  ctx->last_kernel_ms = kms;
  ctx->last_launches = launches;
  ctx->last_fast_launches = nfast;
  ctx->last_generic_launches = ngeneric;
  double sk = 0., sp = 0.;
  for (int q = 0; q < n_timed; ++q) {
    float ms = 0.f, mp = 0.f;
    HIPCHK(hipEventElapsedTime(&mp, ctx->kev[3 * q], ctx->kev[3 * q + 1]));
    HIPCHK(hipEventElapsedTime(&ms, ctx->kev[3 * q + 1], ctx->kev[3 * q + 2]));
    sk += ms;
    sp += mp;
  }
  ctx->last_step_kernel_ms = sk;
  ctx->last_prep_ms = sp;
  ctx->last_step_launches = n_timed;
  if (bnd_out && any_stats && ctx->nopen > 0) {
    std::vector<int> h(need_bnd);
    HIPCHK(hipMemcpy(h.data(), ctx->d_bnd, need_bnd * sizeof(int), hipMemcpyDeviceToHost));
    for (int s = 0; s < nsteps; ++s)
      for (int sc = 0; sc < ctx->nscen; ++sc)
        for (int q = 0; q < ctx->nopen; ++q)
          bnd_out[((size_t)s * ctx->nscen + sc) * ctx->nopen + q] =
              h[(size_t)s * nstat + (size_t)sc * std::max(ctx->nopen, 1) + q];
  }
  if (strand_rec) {
    int cnt[2];
    HIPCHK(hipMemcpy(cnt, ctx->d_strand_count, sizeof(cnt), hipMemcpyDeviceToHost));
    int nr = std::min(cnt[0], ctx->strand_cap);
    std::vector<double> xyz((size_t)nr * 3);
    std::vector<int2> key(nr);
    if (nr > 0) {
      HIPCHK(hipMemcpy(xyz.data(), ctx->d_strand_xyz, xyz.size() * sizeof(double), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(key.data(), ctx->d_strand_key, key.size() * sizeof(int2), hipMemcpyDeviceToHost));
    }
    // serial order of the reference's loop: steps in turn, rows in pool order
    std::vector<int> ord(nr);
    for (int q = 0; q < nr; ++q) ord[q] = q;
    std::sort(ord.begin(), ord.end(), [&](int u, int v) {
      const int pu = key[u].y & ~REC_DEPOSIT, pv = key[v].y & ~REC_DEPOSIT;
      return key[u].x != key[v].x ? key[u].x < key[v].x : pu < pv;
    });
    for (int q = 0; q < nr; ++q) {
      const int r = ord[q];
      if (key[r].y & REC_DEPOSIT) {
        for (int c = 0; c < 3; ++c) ctx->last_deposit.push_back(xyz[3 * r + c]);
        ctx->last_deposit_step.push_back(key[r].x - step0);
      } else {
        for (int c = 0; c < 3; ++c) ctx->last_strand.push_back(xyz[3 * r + c]);
      }
    }
  }
  if (ctx->nsecseg > 0 && !init_only) {
    ctx->last_sec_counts.assign((size_t)nsteps * ctx->nsec, 0);
    if (nsteps > 0)
      HIPCHK(hipMemcpy(ctx->last_sec_counts.data(), ctx->d_sec_counts,
                       (size_t)nsteps * ctx->nsec * sizeof(int), hipMemcpyDeviceToHost));
  }
  int flags[5];
  HIPCHK(hipMemcpy(flags, ctx->d_flag, 5 * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<unsigned long long> errs(ctx->nscen);
  HIPCHK(hipMemcpy(errs.data(), ctx->d_err, ctx->nscen * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  *time_inout = t;
  ctx->last_total_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
  if (flags[2]) return ctx->set_err(CYL_E_STEP, -1, "random-stream event table overflow");
  if (flags[4]) return ctx->set_err(CYL_E_STEP, -1, "draw count differs from the proven uniform assignment");
  for (int s = 0; s < ctx->nscen; ++s) {
    if (errs[s] != ~0ull) {
      int code = (int)(errs[s] & 0xff), pidx = (int)(errs[s] >> 8);
      static const char* msgs[] = {"",
                                   "Failed interpolation, point not in triangulation",
                                   "Failed interpolation, point is below bottom",
                                   "Failed interpolation, point is above surface",
                                   "3D interpolation failed: degenerate vertical layer detected",
                                   "LSM2 requires finite a1 >= 0, dt >= 0 and a1*dt",
                                   "Wrong time scheme",
                                   "random-stream window too small",
                                   "draw-count misprediction",
                                   "neighbourhood search needs triangle_adjacency",
                                   "LSM3 is not available in the device step",
                                   "random-stream event table inconsistent",
                                   "LSM3 requires finite a1 >= 0, A2, TL > 0 and dt >= 0",
                                   "LSM3 covariance has a nonzero singular row or a significantly negative pivot"};
      std::string m = code > 0 && code <= 13 ? msgs[code] : "device error";
      ctx->errcode = code;
      return ctx->set_err(CYL_E_STEP, pidx, m);
    }
  }
  return CYL_OK;
}

extern "C" int cyl_step(cyl_ctx* ctx, const cyl_params* p, double* time_inout, double dt, int nsteps, int* bnd) {
  if (!ctx || !p || !time_inout || nsteps < 0) return CYL_E_ARG;
  ctx->fields.frozen = ctx->frozen;
  return run_steps(ctx, p, time_inout, dt, nsteps, bnd, false);
}

extern "C" int cyl_batch_step(cyl_ctx* ctx, double* time_inout, double dt, int nsteps, int* bnd) {
  if (!ctx || !time_inout || nsteps < 0) return CYL_E_ARG;
  if (!ctx->batched) return ctx->set_err(CYL_E_STATE, -1, "no scenarios set");
  ctx->fields.frozen = ctx->frozen;
  return run_steps(ctx, &ctx->scen_par[0], time_inout, dt, nsteps, bnd, false);
}

extern "C" int cyl_init_particles(cyl_ctx* ctx, const cyl_params* p, double time, double dt) {
  if (!ctx || !p) return CYL_E_ARG;
  double t = time;
  ctx->fields.frozen = ctx->frozen;
  int e = run_steps(ctx, p, &t, dt, 1, nullptr, true);
  ctx->have_last = false;
  return e;
}

extern "C" int cyl_set_tuning(cyl_ctx* ctx, int sort_interval, int blocks_per_cu, int inline_normals,
                              int use_graph) {
  if (!ctx || sort_interval < 0 || blocks_per_cu < 0 || inline_normals < 0 || inline_normals > 2) return CYL_E_ARG;
  ctx->use_graph = use_graph ? 1 : 0;
  ctx->sort_interval = sort_interval;
  ctx->blocks_per_cu = blocks_per_cu;
  if (inline_normals != ctx->inline_normals) ctx->pregen_h.clear();  // re-upload the generator flags
  ctx->inline_normals = inline_normals;
  return CYL_OK;
}

extern "C" int cyl_last_timing(cyl_ctx* ctx, double* kms, double* tms, int* launches) {
  if (!ctx) return CYL_E_ARG;
  if (kms) *kms = ctx->last_step_kernel_ms;
  if (tms) *tms = ctx->last_total_ms;
  if (launches) *launches = ctx->last_step_launches;
  return CYL_OK;
}

extern "C" int cyl_control_sections_set(cyl_ctx* ctx, int nsec, const int* nseg_of, const double* seg) {
  if (!ctx || nsec < 0 || (nsec > 0 && (!nseg_of || !seg))) return CYL_E_ARG;
  HIPCHK(hipSetDevice(ctx->device));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  std::vector<double4> sg;
  std::vector<int> of;
  int q = 0;
  for (int s = 0; s < nsec; ++s) {
    if (nseg_of[s] < 0) return CYL_E_ARG;
    for (int k = 0; k < nseg_of[s]; ++k, ++q) {
      sg.push_back(make_double4(seg[4 * q], seg[4 * q + 1], seg[4 * q + 2], seg[4 * q + 3]));
      of.push_back(s);
    }
  }
  if (ctx->d_sec_seg) HIPCHK(hipFree(ctx->d_sec_seg));
  if (ctx->d_sec_of) HIPCHK(hipFree(ctx->d_sec_of));
  ctx->d_sec_seg = nullptr;
  ctx->d_sec_of = nullptr;
  ctx->nsec = nsec;
  ctx->nsecseg = (int)sg.size();
  ctx->last_sec_counts.clear();
  if (ctx->nsecseg > 0) {
    HIPCHK(hipMalloc(&ctx->d_sec_seg, sg.size() * sizeof(double4)));
    HIPCHK(hipMalloc(&ctx->d_sec_of, of.size() * sizeof(int)));
    HIPCHK(hipMemcpy(ctx->d_sec_seg, sg.data(), sg.size() * sizeof(double4), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_sec_of, of.data(), of.size() * sizeof(int), hipMemcpyHostToDevice));
  }
  return CYL_OK;
}

extern "C" int cyl_control_counts(cyl_ctx* ctx, int* out, int cap, int* nout) {
  if (!ctx || !nout) return CYL_E_ARG;
  *nout = (int)ctx->last_sec_counts.size();
  for (int i = 0; i < *nout && i < cap && out; ++i) out[i] = ctx->last_sec_counts[i];
  return CYL_OK;
}

extern "C" int cyl_stranded_get(cyl_ctx* ctx, double* xyz, int cap, int* n) {
  if (!ctx || !n) return CYL_E_ARG;
  *n = (int)(ctx->last_strand.size() / 3);
  for (int q = 0; q < *n && q < cap && xyz; ++q)
    for (int c = 0; c < 3; ++c) xyz[3 * q + c] = ctx->last_strand[3 * q + c];
  return CYL_OK;
}

extern "C" int cyl_deposited_get(cyl_ctx* ctx, double* xyz, int* step, int cap, int* n) {
  if (!ctx || !n) return CYL_E_ARG;
  *n = (int)ctx->last_deposit_step.size();
  for (int q = 0; q < *n && q < cap; ++q) {
    if (xyz)
      for (int c = 0; c < 3; ++c) xyz[3 * q + c] = ctx->last_deposit[3 * q + c];
    if (step) step[q] = ctx->last_deposit_step[q];
  }
  return CYL_OK;
}

extern "C" int cyl_last_path(cyl_ctx* ctx, int* fused_launches, int* generic_launches, int* graph_steps) {
  if (!ctx) return CYL_E_ARG;
  if (fused_launches) *fused_launches = ctx->last_fast_launches;
  if (generic_launches) *generic_launches = ctx->last_generic_launches;
  if (graph_steps) *graph_steps = ctx->last_graph_steps;
  return CYL_OK;
}

extern "C" int cyl_last_prep_timing(cyl_ctx* ctx, double* prep_ms) {
  if (!ctx || !prep_ms) return CYL_E_ARG;
  *prep_ms = ctx->last_prep_ms;
  return CYL_OK;
}

// Counting sort of the slots by element, then one gather pass that moves the
// current state copy and every per-slot array into the spare buffers, whose
// pointers are then swapped in.  The fluid velocity seen is carried only when
// the next step reads it (the fused 2D LSM-1 kernel rewrites it).  Active ranks
// are pool-order quantities and move with their slots.
// counters, cursors and spare per-slot buffers of the sort (allocated once per pool)
static int ensure_sort_bufs(cyl_ctx* ctx) {
  int n = ctx->n, nt = ctx->mesh.nt;
  Bufs& bb = ctx->part_bufs;
  int ncls = ctx->batched ? std::max<int>(1, (int)ctx->classes.size()) : 1;
  int nbin = ((nt >> SORT_SHIFT) + 3) * ncls;
  if (!ctx->d_scnt || ctx->sort_nt != nt || ctx->sort_ncls != ncls) {
    HIPCHK(dalloc(bb, &ctx->d_scnt, nbin));
    HIPCHK(dalloc(bb, &ctx->d_cursor, nbin));
    HIPCHK(dalloc(bb, &ctx->d_sort_tiles, nblk(nbin, SCAN_TILE) + 1));
    ctx->sort_nt = nt;
    ctx->sort_ncls = ncls;
  }
  if (ctx->d_scen && !ctx->sp_scen) HIPCHK(dalloc(bb, &ctx->sp_scen, n));
  if (!ctx->sp_gpool) {
    HIPCHK(dalloc(bb, &ctx->sp_gpool, n));
    HIPCHK(dalloc(bb, &ctx->sp_init_tri, n));
    HIPCHK(dalloc(bb, &ctx->sp_del_step, n));
    HIPCHK(dalloc(bb, &ctx->sp_inactive, n));
    HIPCHK(dalloc(bb, &ctx->sp_rank, n));
    HIPCHK(dalloc(bb, &ctx->sp_diam, n));
    HIPCHK(dalloc(bb, &ctx->sp_vol, n));
    HIPCHK(dalloc(bb, &ctx->sp_depth, ctx->dim == 3 ? n : 1));
  }
  return CYL_OK;
}

static int enqueue_sort(cyl_ctx* ctx, bool need_us, int& launches) {
  int e = ensure_sort_bufs(ctx);
  if (e) return e;
  int n = ctx->n, nt = ctx->mesh.nt;
  hipStream_t s = ctx->stream;
  int ncls = ctx->batched ? std::max<int>(1, (int)ctx->classes.size()) : 1;
  int nbin = ((nt >> SORT_SHIFT) + 3) * ncls;
  PoolDev& A = ctx->pool[ctx->cur];
  PoolDev& B = ctx->pool[1 - ctx->cur];
  k_fill_int<<<nblk(nbin), BLK, 0, s>>>(ctx->d_scnt, nbin, 0);  // a kernel, not a memset: graph-capturable
  k_sort_hist<<<nblk(n, BLK * SORT_ITEMS), BLK, 0, s>>>(A.tri, A.inactive, ctx->d_elem_rank, ctx->batched ? ctx->d_scen : nullptr,
                                      ctx->d_class_of_scen, n, nt, ctx->d_scnt, ctx->d_skey);
  enqueue_scan(ctx, ctx->d_scnt, ctx->d_cursor, (int64_t)nbin, ctx->d_sort_tiles, ctx->d_total,
               (const int64_t*)nullptr, (const int*)nullptr, launches);
  k_sort_place<<<nblk(n, BLK * SORT_ITEMS), BLK, 0, s>>>(ctx->d_skey, n, ctx->d_cursor, ctx->d_perm);
  PermArgs pa{};
  pa.perm = ctx->d_perm;
  pa.n = n;
  auto pd = [&](const double* src, double* dst) {
    pa.dsrc[pa.nd] = src;
    pa.ddst[pa.nd++] = dst;
  };
  auto pi = [&](const int* src, int* dst) {
    pa.isrc[pa.ni] = src;
    pa.idst[pa.ni++] = dst;
  };
  pd(A.x, B.x);
  pd(A.y, B.y);
  if (ctx->dim == 3) pd(A.z, B.z);
  for (int c = 0; c < ctx->dim; ++c) {
    pd(A.up[c], B.up[c]);
    if (need_us) pd(A.us[c], B.us[c]);
  }
  pd(ctx->d_diam, ctx->sp_diam);
  pd(ctx->d_vol, ctx->sp_vol);
  if (ctx->dim == 3) pd(ctx->d_depth, ctx->sp_depth);
  pi(A.tri, B.tri);
  if (ctx->dim == 3) pi(A.ll, B.ll);
  pi(ctx->d_gpool, ctx->sp_gpool);
  pi(ctx->d_init_tri, ctx->sp_init_tri);
  pi(ctx->d_del_step, ctx->sp_del_step);
  pi(ctx->pool[0].inactive, ctx->sp_inactive);
  pi(ctx->d_rank, ctx->sp_rank);  // a slot's active rank moves with it
  if (ctx->d_scen) pi(ctx->d_scen, ctx->sp_scen);
  k_permute_all<<<fast_grid(nblk(n)), BLK, 0, s>>>(pa);
  launches += 3;
  HIPCHK(hipGetLastError());
  ctx->cur = 1 - ctx->cur;
  std::swap(ctx->d_gpool, ctx->sp_gpool);
  std::swap(ctx->d_init_tri, ctx->sp_init_tri);
  std::swap(ctx->d_del_step, ctx->sp_del_step);
  std::swap(ctx->d_diam, ctx->sp_diam);
  std::swap(ctx->d_vol, ctx->sp_vol);
  std::swap(ctx->d_depth, ctx->sp_depth);
  std::swap(ctx->pool[0].inactive, ctx->sp_inactive);
  std::swap(ctx->d_rank, ctx->sp_rank);
  if (ctx->d_scen) std::swap(ctx->d_scen, ctx->sp_scen);
  ctx->pool[1].inactive = ctx->pool[0].inactive;
  ctx->have_last = false;
  return CYL_OK;
}
