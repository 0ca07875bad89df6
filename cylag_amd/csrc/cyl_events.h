// Random-stream event tables of one step.
//
// The reference draws every normal of a step from one serial PCG32 stream in
// pool order (random_utils.pyx:17, solver.pyx:459).  A particle's normals are
// therefore fixed by its first draw unit (a scan over the pool) and by the
// stream positions those units map to.  Rectangular normals take one position
// each.  A Ziggurat normal takes one position unless its first draw fails the
// fast test (random_utils.pyx:162-164), and a polar pair call takes two
// positions per attempt plus two per rejected attempt.  The kernels below find,
// per scenario window of the stream:
//   rng_method 2: every position whose draw fails the fast test (a slow
//     candidate), the value and length of the normal that would start there,
//     which candidates lie on the chain of normal starts, and for those the
//     normal index, the cumulative extra positions and the value;
//   rng_method 3: every rejected attempt, keyed r_j - j.
// Draws.events_before / zig_position / polar_position then map a unit to its
// position, and the normal generator replays runs of the stream from them.
//
// Pipeline (rng_method 2): chunk layout of the step's windows (ev_layout, in
// the step's preparation) -> one pass over the windows (each block finds the
// candidates of its span of chunks, compacts them in LDS and samples them into
// the span's region, and reduces each span's largest reach key) -> scan of
// the span counts and maxima -> gather into dense, position-ordered tables
// with the exclusive max-scan of the reach keys -> resolve with an (extra,
// count) scan within each span -> scan of the span totals -> compact ->
// buckets.
#pragma once
#include "cyl_kernels.h"

namespace cyl {

constexpr int ZCHUNK = 64;           // stream positions per scan thread
constexpr int REACH_SHIFT = 40;      // reach key: (scenario << 40) | last position
constexpr int64_t EC_MASK = 0xffffffffLL;

struct EventArgs {
  StreamDev* stream;          // [nscen]
  const JumpTable* jt;        // [nscen]
  const JumpTable8* jt8;      // [nscen]
  const ZigTables* zt;
  const int64_t* chunk_base;  // [nscen + 1] first chunk of each scenario's window this step (ev_layout)
  int nscen;
  int64_t nchunks;
  int* chunk_loc;             // per chunk: its first candidate's offset in its span
  int* span_cnt;              // per span: candidates (at most span_cap)
  int64_t* span_off;          // per span: first dense index (scan of span_cnt)
  int64_t* span_max;          // per span (rng_method 2): largest reach key, -1 if none
  int64_t* span_carry;        // per span: exclusive max-scan of span_max
  int span_cap;               // candidate records per span region
  int64_t* reg_pos;           // span regions (span_cap records per span of the chunk capacity): position, and for rng_method 2
  int64_t* reg_len;           //   the length, value and reach key of the normal starting there
  double* reg_val;
  int64_t* reg_key;
  int64_t* nev;               // total candidates
  int64_t* ev_begin;          // [nscen + 1] first candidate of each scenario
  int64_t* ev_pos;            // dense tables: candidate position (relative to the scenario's window start)
  int64_t* ev_len;            // length of the normal starting at the candidate (zig)
  double* ev_val;             // its value (zig)
  int64_t* reach;             // exclusive max-scan of the reach keys (zig)
  const int64_t* chunk_live;  // chunks of the step's windows (ev_layout)
  int64_t ev_cap;
  int* overflow;
  const int* cond;            // run only when *cond != 0
};

// candidate records per span region: a span is BLK chunks of ZCHUNK positions;
// Ziggurat slow starts come at about 1.2 % of positions (about 200 per span),
// rejected polar attempts at 1 - pi/4 of BLK * ZCHUNK / 2 attempts (about 1760)
constexpr int EV_SPAN_CAP_ZIG = 1024;
constexpr int EV_SPAN_CAP_POLAR = 4096;
static_assert(ZCHUNK == 64, "a chunk's candidates are one 64-bit mask");

__device__ __forceinline__ int scen_of(const int64_t* base, int nscen, int64_t i) {
  int lo = 0, hi = nscen - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// window needed for u units: zig normals take ~1.03 positions on average
__device__ __forceinline__ int64_t window_for(int method, int64_t units) {
  if (units <= 0) return 0;
  if (method == 2) return units + units / 8 + 4096;
  return 2 * (units + units / 2 + 4096);  // polar: attempts of 2 positions, ~21% rejected
}

// Layout of the step's event tables (one block).  Chunks: scenario q's window
// (window_for its units, stored in its stream) takes ceil(W / ZCHUNK)
// chunks from base[q], an exclusive scan over the scenarios; base[nscen] and
// *live receive the total, and a layout past the table's nchunks is reported
// (overflow) and *live clamped to the table.  Bucket index: the scenario's
// lookups ask for units up to its unit count u, so its live entries are the
// first (u >> bshift) + 2 of its capacity (beyond them a lookup searches the
// keys); lb_base is their exclusive scan.
struct LayoutArgs {
  StreamDev* st;
  const int64_t* units;
  int nscen, method, bshift;
  int64_t nchunks;
  const int64_t* bucket_base;  // [nscen + 1] bucket capacity of each scenario
  int64_t* base;               // [nscen + 1]
  int64_t* live;
  int64_t* lb_base;            // [nscen + 1]
  int* overflow;
};

__device__ void ev_layout(const LayoutArgs& a, int64_t* lds) {
  int64_t carry = 0, bcarry = 0;
  for (int t0 = 0; t0 < a.nscen; t0 += BLK) {
    const int q = t0 + threadIdx.x;
    int64_t c = 0, nb = 0;
    if (q < a.nscen) {
      const int64_t u = a.units[q], W = window_for(a.method, u);
      a.st[q].window = W;
      c = (W + ZCHUNK - 1) / ZCHUNK;
      const int64_t cap = a.bucket_base[q + 1] - a.bucket_base[q];
      nb = (u >> a.bshift) + 2;
      nb = nb < cap ? nb : cap;
    }
    int64_t tot, btot;
    const int64_t pre = block_excl_scan<OpSum, BLK>(c, tot, lds);
    const int64_t bpre = block_excl_scan<OpSum, BLK>(nb, btot, lds);
    if (q < a.nscen) {
      a.base[q] = carry + pre;
      a.lb_base[q] = bcarry + bpre;
    }
    carry += tot;
    bcarry += btot;
  }
  if (threadIdx.x == 0) {
    a.base[a.nscen] = carry;
    a.lb_base[a.nscen] = bcarry;
    if (carry > a.nchunks) atomicOr(a.overflow, 1);
    *a.live = carry < a.nchunks ? carry : a.nchunks;
  }
}

__global__ void __launch_bounds__(BLK) k_ev_layout(LayoutArgs a, const int* cond) {
  if (cond && *cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  ev_layout(a, lds);
}

__device__ __forceinline__ bool zig_fast(uint32_t u, const uint32_t* kn) {
  int32_t hz = (int32_t)u;
  uint32_t iz = (uint32_t)hz & 127u;
  uint32_t ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;  // libc abs(INT_MIN) wraps
  return ahz < kn[iz];
}

// One pass over the stream window, one chunk of ZCHUNK positions per thread,
// blocks striding over spans of BLK chunks.  Each thread draws its chunk and
// marks its candidates in a 64-bit mask (rng_method 2: draws failing the fast
// test; 3: rejected polar attempts); a block scan of the mask counts gives
// each its offset in the span, and the span's candidates go, in position
// order, to the span's region.  Under rng_method 2 the block then samples
// them, one candidate per thread: the generator state there is one
// byte-window jump from its chunk's start state (kept in LDS), and the value
// and length of the Ziggurat normal starting there and its reach key
// (scenario << REACH_SHIFT) | (pos + len - 1) are stored with it, and the
// block stores the largest key of each of its spans at the end.  A
// candidate whose normal reaches the base strip's tail loop (two correctly
// rounded logarithms per attempt; a few percent of candidates) is queued in
// LDS and finished by the block at the end, so that the waves sampling the
// other candidates do not wait on it.  The chunks are those of the step's
// layout (ev_layout).
constexpr int TAIL_CAP = 128;  // a block queues about 10-50 (a few % of its candidates); a full queue finishes in place
struct TailRec {
  uint64_t state;
  int64_t i, pos;
  int32_t hz, ncalls, scen, pad;
};

__device__ __forceinline__ void reg_finish(const EventArgs& e, int64_t r, double v, int64_t ncalls, int64_t pos,
                                           int s) {
  e.reg_val[r] = v;
  e.reg_len[r] = ncalls;
  e.reg_key[r] = ((int64_t)s << REACH_SHIFT) | (pos + ncalls - 1);
}

template <int METHOD>
__global__ void __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(5))) k_ev_write(EventArgs e) {
  if (e.cond && *e.cond == 0) return;
  __shared__ JumpTable8 t8;
  __shared__ uint32_t kn[128];
  __shared__ int64_t lds[BLK / 64];
  __shared__ uint64_t cmask[BLK], cstate[BLK];
  __shared__ int64_t cpos[BLK];
  __shared__ int cscen[BLK], cbase[BLK];
  __shared__ TailRec tails[METHOD == 2 ? TAIL_CAP : 1];
  __shared__ int ntail;
  if (threadIdx.x == 0) ntail = 0;
  const int s0 = scen_of(e.chunk_base, e.nscen, (int64_t)blockIdx.x * BLK);
  const uint64_t inc0 = e.stream[s0].inc;
  // the byte-window table of the block's first scenario; it depends on the
  // stream increment only, so it serves every scenario of that increment (one
  // stream advanced per scenario), the others jump with their own bit table
  {
    const uint64_t* src = reinterpret_cast<const uint64_t*>(&e.jt8[s0]);
    uint64_t* dst = reinterpret_cast<uint64_t*>(&t8);
    for (int q = threadIdx.x; q < (int)(sizeof(JumpTable8) / 8); q += BLK) dst[q] = src[q];
    if (METHOD == 2)
      for (int q = threadIdx.x; q < 128; q += BLK) kn[q] = e.zt->kn[q];
  }
  __syncthreads();  // the tables above are staged
  const int64_t nch = *e.chunk_live;
  const int span_cap = e.span_cap;
  for (int64_t span = (int64_t)blockIdx.x * BLK; span < nch; span += (int64_t)gridDim.x * BLK) {
    const int64_t ch = span + threadIdx.x;
    uint64_t m = 0, st0 = 0;
    int64_t p0 = 0;
    int s = 0;
    if (ch < nch) {
      s = scen_of(e.chunk_base, e.nscen, ch);
      const int64_t local = ch - e.chunk_base[s];
      const StreamDev& st = e.stream[s];
      p0 = local * ZCHUNK;
      if (p0 < st.window) {
        Pcg g;
        g.inc = st.inc;
        // polar: attempt a uses positions 2a, 2a+1; p0 is even
        g.state = st.inc == inc0 ? pcg_jump8(st.base, (uint64_t)p0, t8, e.jt[s]) : pcg_jump(st.base, (uint64_t)p0, e.jt[s]);
        st0 = g.state;
        if (METHOD == 2) {
          for (int q = 0; q < ZCHUNK; ++q)
            if (!zig_fast(pcg_next(g), kn)) m |= 1ull << q;
        } else {
          for (int q = 0; q < ZCHUNK / 2; ++q) {
            uint32_t u1 = pcg_next(g), u2 = pcg_next(g);
            if (!polar_accept(u1, u2)) m |= 1ull << q;
          }
        }
      }
    }
    cmask[threadIdx.x] = m;
    cstate[threadIdx.x] = st0;
    cpos[threadIdx.x] = p0;
    cscen[threadIdx.x] = s;
    int64_t total;
    const int64_t excl = block_excl_scan<OpSum, BLK>((int64_t)__popcll(m), total, lds);
    cbase[threadIdx.x] = (int)excl;
    if (ch < nch) e.chunk_loc[ch] = (int)excl;
    const int64_t sp = span / BLK;
    const int ncand = total < span_cap ? (int)total : span_cap;
    if (threadIdx.x == 0) {
      e.span_cnt[sp] = ncand;
      if (total > span_cap) atomicOr(e.overflow, 2);  // reported, never stored past the region
    }
    __syncthreads();
    const int64_t region = sp * span_cap;
    for (int ci = threadIdx.x; ci < ncand; ci += BLK) {
      // owner: the last thread whose first candidate is at or before ci
      int lo = 0, hi = BLK - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cbase[mid] <= ci) lo = mid;
        else hi = mid - 1;
      }
      uint64_t mm = cmask[lo];
      for (int k = ci - cbase[lo]; k > 0; --k) mm &= mm - 1;  // drop the lower candidates
      const int q = __ffsll((unsigned long long)mm) - 1;
      const int64_t r = region + ci;
      if (METHOD == 2) {
        const int so = cscen[lo];
        const int64_t pos = cpos[lo] + q;
        e.reg_pos[r] = pos;
        const uint64_t inc = e.stream[so].inc;
        Draws d;
        d.zt = e.zt;
        d.pre_pos = nullptr;
        d.g.inc = inc;
        d.g.state = inc == inc0 ? t8.mult[0][q] * cstate[lo] + t8.plus[0][q] : pcg_jump(cstate[lo], (uint64_t)q, e.jt[so]);
        d.ncalls = 0;
        double v;
        int32_t hz;
        if (d.zig_head(v, hz)) {
          reg_finish(e, r, v, d.ncalls, pos, so);
        } else {
          int k = atomicAdd(&ntail, 1);
          if (k < TAIL_CAP) tails[k] = TailRec{d.g.state, r, pos, hz, (int32_t)d.ncalls, so, 0};
          else reg_finish(e, r, d.zig_tail(hz), d.ncalls, pos, so);  // queue full: finish here
        }
      } else {
        e.reg_pos[r] = cpos[lo] / 2 + q;
      }
    }
    __syncthreads();  // the span's LDS entries are refilled by the next span
  }
  if (METHOD == 2) {
    const int nt = ntail < TAIL_CAP ? ntail : TAIL_CAP;
    for (int k = threadIdx.x; k < nt; k += BLK) {
      const TailRec t = tails[k];
      Draws d;
      d.zt = e.zt;
      d.pre_pos = nullptr;
      d.g.inc = e.stream[t.scen].inc;
      d.g.state = t.state;
      d.ncalls = t.ncalls;
      double v = d.zig_tail(t.hz);
      reg_finish(e, t.i, v, d.ncalls, t.pos, t.scen);
    }
    __syncthreads();  // every key of the block's spans is stored
    for (int64_t span = (int64_t)blockIdx.x * BLK; span < nch; span += (int64_t)gridDim.x * BLK) {
      const int64_t sp = span / BLK;
      const int n = e.span_cnt[sp];
      int64_t m = OpMax::ident;
      for (int i = threadIdx.x; i < n; i += BLK) {
        const int64_t k = e.reg_key[sp * span_cap + i];
        m = k > m ? k : m;
      }
      int64_t mx;
      block_excl_scan<OpMax, BLK>(m, mx, lds);
      if (threadIdx.x == 0) e.span_max[sp] = mx;
    }
  }
}

// Exclusive scans of the span counts and (rng_method 2) of the span maxima
// (one block, tiles of BLK * 8 spans), the total number of candidates, and the
// first candidate of each scenario: the dense index of the first candidate of
// its first chunk.
template <int METHOD>
__global__ void __launch_bounds__(BLK) k_ev_span_scan(EventArgs e) {
  if (e.cond && *e.cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  const int64_t nch = *e.chunk_live;
  const int64_t nsp = (nch + BLK - 1) / BLK;
  int64_t carry = 0, mcarry = OpMax::ident;
  constexpr int IT = 8;
  for (int64_t t0 = 0; t0 < nsp; t0 += (int64_t)BLK * IT) {
    int64_t v[IT], mv[IT], run = 0, mrun = OpMax::ident;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = t0 + (int64_t)threadIdx.x * IT + k;
      v[k] = i < nsp ? e.span_cnt[i] : 0;
      run += v[k];
      if (METHOD == 2) {
        mv[k] = i < nsp ? e.span_max[i] : OpMax::ident;
        mrun = mv[k] > mrun ? mv[k] : mrun;
      }
    }
    int64_t total;
    int64_t pre = block_excl_scan<OpSum, BLK>(run, total, lds);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = t0 + (int64_t)threadIdx.x * IT + k;
      if (i < nsp) e.span_off[i] = carry + pre;
      pre += v[k];
    }
    carry += total;
    if (METHOD == 2) {
      int64_t mtot;
      int64_t mpre = block_excl_scan<OpMax, BLK>(mrun, mtot, lds);
      mpre = mpre > mcarry ? mpre : mcarry;
#pragma unroll
      for (int k = 0; k < IT; ++k) {
        const int64_t i = t0 + (int64_t)threadIdx.x * IT + k;
        if (i < nsp) e.span_carry[i] = mpre;
        mpre = mv[k] > mpre ? mv[k] : mpre;
      }
      mcarry = mtot > mcarry ? mtot : mcarry;
    }
  }
  __syncthreads();  // span_off is complete (written by this block)
  if (threadIdx.x == 0) *e.nev = carry;
  for (int q = threadIdx.x; q <= e.nscen; q += BLK) {
    int64_t b = carry;
    if (q < e.nscen && e.chunk_base[q] < nch) {
      const int64_t ch = e.chunk_base[q];
      b = e.span_off[ch / BLK] + e.chunk_loc[ch];
    }
    e.ev_begin[q] = b < e.ev_cap ? b : e.ev_cap;  // an overflow is reported, never indexed past
  }
  if (threadIdx.x == 0 && carry > e.ev_cap) atomicOr(e.overflow, 2);
}

// The span regions gathered into the dense, position-ordered tables: block b
// copies the spans b, b + grid, ...  Under rng_method 2 it also writes the
// exclusive max-scan of the reach keys: the span's carry, then a block scan of
// the span's keys (EV_SPAN_CAP_ZIG / BLK consecutive keys per thread).
template <int METHOD>
__global__ void __launch_bounds__(BLK) k_ev_gather(EventArgs e) {
  if (e.cond && *e.cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  const int64_t nsp = (*e.chunk_live + BLK - 1) / BLK;
  const int span_cap = e.span_cap;
  for (int64_t sp = blockIdx.x; sp < nsp; sp += gridDim.x) {
    const int n = e.span_cnt[sp];
    const int64_t o = e.span_off[sp], r0 = sp * span_cap;
    for (int i = threadIdx.x; i < n; i += BLK) {
      const int64_t d = o + i;
      if (d >= e.ev_cap) break;
      e.ev_pos[d] = e.reg_pos[r0 + i];
      if (METHOD == 2) {
        e.ev_len[d] = e.reg_len[r0 + i];
        e.ev_val[d] = e.reg_val[r0 + i];
      }
    }
    if (METHOD == 2) {
      constexpr int PT = EV_SPAN_CAP_ZIG / BLK;
      int64_t k[PT], run = OpMax::ident;
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const int i = threadIdx.x * PT + q;
        k[q] = i < n ? e.reg_key[r0 + i] : OpMax::ident;
        run = k[q] > run ? k[q] : run;
      }
      int64_t tot;
      int64_t m = block_excl_scan<OpMax, BLK>(run, tot, lds);
      const int64_t c = e.span_carry[sp];
      m = m > c ? m : c;
#pragma unroll
      for (int q = 0; q < PT; ++q) {
        const int i = threadIdx.x * PT + q;
        if (i < n && o + i < e.ev_cap) e.reach[o + i] = m;
        m = k[q] > m ? k[q] : m;
      }
    }
  }
}

// Ziggurat chain resolution, one span of the dense tables per block.
// Candidate p is off the chain iff an on-chain slow start q < p has
// q + len(q) - 1 >= p.  A candidate that no earlier candidate reaches is
// surely on-chain; from there the chain is followed forward.  reach (exclusive
// max-scan of the keys) gives the furthest position reached by earlier
// candidates of the same scenario.  The packed value (len - 1) << 32 | 1 of an
// on-chain candidate (0 otherwise) is scanned within the span: pk receives
// the exclusive prefix within the span (ON_CHAIN set for an on-chain
// candidate) and span_pk the span's total.
constexpr int64_t ON_CHAIN = (int64_t)1 << 62;  // flag bit of the per-candidate prefix

struct ChainArgs {
  const int64_t* pos;
  const int64_t* len;
  const double* val;
  const int64_t* reach;
  const int64_t* ev_begin;  // [nscen + 1], the last entry clamped to the capacity
  const int* span_cnt;
  const int64_t* span_off;
  const int64_t* chunk_live;
  int nscen;
  int64_t ev_cap;
  int64_t* pk;              // per candidate: exclusive prefix within its span
  int64_t* span_pk;         // per span: total
  int64_t* span_pkoff;      // per span: exclusive scan of span_pk
  int64_t* scen_c;          // [nscen + 1]: the exclusive prefix at each scenario's first candidate
  int64_t* scen_c0;         // [nscen + 1]: its count half (first compacted event of the scenario)
  StreamDev* st;
  int64_t* key;             // compacted events: normal index
  int64_t* xtra;            //   cumulative extra positions (inclusive)
  double* cval;             //   value
  const int* bucket;
  const int64_t* bucket_base;
  const int64_t* lb_base;   // [nscen + 1] live bucket entries (ev_layout)
  int bshift;
  const int* cond;
};

__device__ __forceinline__ int64_t span_items(const ChainArgs& a, int64_t sp, int64_t& o) {
  const int64_t n = a.span_cnt[sp];
  o = a.span_off[sp];
  const int64_t ne = a.ev_begin[a.nscen];
  return o >= ne ? 0 : (o + n <= ne ? n : ne - o);
}

__global__ void __launch_bounds__(BLK) k_ev_resolve(ChainArgs a) {
  if (a.cond && *a.cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  const int64_t nsp = (*a.chunk_live + BLK - 1) / BLK;
  for (int64_t sp = blockIdx.x; sp < nsp; sp += gridDim.x) {
    int64_t o;
    const int64_t n = span_items(a, sp, o);
    // rounds of BLK candidates, one per thread (a span holds about 200), so
    // that a thread's chain walks do not queue behind each other
    int64_t carry = 0;
    for (int64_t r0 = 0; r0 < n; r0 += BLK) {
      const int64_t li = r0 + threadIdx.x;
      int64_t v = 0;
      if (li < n) {
        const int64_t i = o + li;
        const int s = scen_of(a.ev_begin, a.nscen, i);
        const int64_t b = a.ev_begin[s];
        auto reach_of = [&](int64_t j) -> int64_t {
          const int64_t m = a.reach[j];
          return (m >= 0 && (m >> REACH_SHIFT) == s) ? (m & ((1LL << REACH_SHIFT) - 1)) : -1;
        };
        int64_t j0 = i;
        while (j0 > b && reach_of(j0) >= a.pos[j0]) --j0;
        int64_t cur_end = a.pos[j0] + a.len[j0] - 1;
        int on = 1;
        for (int64_t l = j0 + 1; l <= i; ++l) {
          if (a.pos[l] > cur_end) {
            cur_end = a.pos[l] + a.len[l] - 1;
            on = 1;
          } else {
            on = 0;
          }
        }
        v = on ? ((a.len[i] - 1) << 32) | 1 : 0;
      }
      int64_t tot;
      const int64_t pre = block_excl_scan<OpSum, BLK>(v, tot, lds);
      if (li < n) a.pk[o + li] = (carry + pre) | (v & 1 ? ON_CHAIN : 0);
      carry += tot;
    }
    if (threadIdx.x == 0) a.span_pk[sp] = carry;
  }
}

// Exclusive scan of the span totals (one block), and each scenario's prefix at
// its first candidate (the total past the last one).
__global__ void __launch_bounds__(BLK) k_ev_chain_scan(ChainArgs a) {
  if (a.cond && *a.cond == 0) return;
  __shared__ int64_t lds[BLK / 64];
  const int64_t nsp = (*a.chunk_live + BLK - 1) / BLK;
  int64_t carry = 0;
  constexpr int IT = 8;
  for (int64_t t0 = 0; t0 < nsp; t0 += (int64_t)BLK * IT) {
    int64_t v[IT], run = 0;
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = t0 + (int64_t)threadIdx.x * IT + k;
      v[k] = i < nsp ? a.span_pk[i] : 0;
      run += v[k];
    }
    int64_t total;
    int64_t pre = block_excl_scan<OpSum, BLK>(run, total, lds);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t i = t0 + (int64_t)threadIdx.x * IT + k;
      if (i < nsp) a.span_pkoff[i] = carry + pre;
      pre += v[k];
    }
    carry += total;
  }
  __syncthreads();  // span_pkoff is complete (written by this block)
  const int64_t ne = a.ev_begin[a.nscen];
  for (int q = threadIdx.x; q <= a.nscen; q += BLK) {
    const int64_t d = a.ev_begin[q];
    int64_t c = carry;
    if (d < ne) {
      // the span holding candidate d: the last span whose first index is at or before it
      int64_t lo = 0, hi = nsp - 1;
      while (lo < hi) {
        const int64_t mid = (lo + hi + 1) >> 1;
        if (a.span_off[mid] <= d) lo = mid;
        else hi = mid - 1;
      }
      c = a.span_pkoff[lo] + (a.pk[d] & ~ON_CHAIN);
    }
    a.scen_c[q] = c;
    a.scen_c0[q] = c & EC_MASK;
  }
}

// Point each scenario's stream at its slice of the event tables (block 0,
// threads below nscen) and compact the on-chain candidates, one span per
// block: normal index, cumulative extra positions (inclusive) and value.
__global__ void __launch_bounds__(BLK) k_ev_compact(ChainArgs a) {
  if (a.cond && *a.cond == 0) return;
  if (blockIdx.x == 0)
    for (int q = threadIdx.x; q < a.nscen; q += BLK) {
      const int64_t c0 = a.scen_c0[q], c1 = a.scen_c0[q + 1];
      a.st[q].ev_key = a.key + c0;
      a.st[q].ev_extra = a.xtra + c0;
      a.st[q].ev_val = a.cval + c0;
      a.st[q].n_ev = c1 - c0;
      a.st[q].ev_bucket = a.bucket + a.bucket_base[q];
      a.st[q].n_bucket = a.lb_base[q + 1] - a.lb_base[q];
      a.st[q].ev_bshift = a.bshift;
    }
  const int64_t nsp = (*a.chunk_live + BLK - 1) / BLK;
  for (int64_t sp = blockIdx.x; sp < nsp; sp += gridDim.x) {
    int64_t o;
    const int64_t n = span_items(a, sp, o);
    const int64_t base = a.span_pkoff[sp];
    for (int64_t li = threadIdx.x; li < n; li += BLK) {
      const int64_t i = o + li;
      const int64_t p = a.pk[i];
      if (!(p & ON_CHAIN)) continue;
      const int64_t C = base + (p & ~ON_CHAIN);
      const int s = scen_of(a.ev_begin, a.nscen, i);
      const int64_t e_excl = (C >> 32) - (a.scen_c[s] >> 32);
      const int64_t c = C & EC_MASK;
      a.key[c] = a.pos[i] - e_excl;
      a.xtra[c] = e_excl + a.len[i] - 1;
      a.cval[c] = a.val[i];
    }
  }
}

// polar: key_j = r_j - j (j counted within the scenario), and the stream slices
__global__ void __launch_bounds__(BLK) k_ev_polar(StreamDev* st, int nscen, const int64_t* ev_begin,
                                                  const int64_t* nev, const int64_t* pos, int64_t* key,
                                                  const int* bucket, const int64_t* bucket_base,
                                                  const int64_t* lb_base, int bshift, const int* cond) {
  if (cond && *cond == 0) return;
  const int64_t ne = ev_begin[nscen];  // candidates, clamped to the capacity
  const int64_t end = ne > nscen ? ne : nscen;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < end; i += (int64_t)gridDim.x * BLK) {
    if (i < nscen) {
      st[i].ev_key = key + ev_begin[i];
      st[i].ev_extra = nullptr;
      st[i].ev_val = nullptr;
      st[i].n_ev = ev_begin[i + 1] - ev_begin[i];
      st[i].ev_bucket = bucket + bucket_base[i];
      st[i].n_bucket = lb_base[i + 1] - lb_base[i];
      st[i].ev_bshift = bshift;
    }
    if (i >= ne) continue;
    int s = scen_of(ev_begin, nscen, i);
    key[i] = pos[i] - (i - ev_begin[s]);
  }
}

// bucket index of every scenario's event keys: entry b = lower_bound(b << shift),
// the first event whose key is at least b << shift, for the scenario's live
// entries b < n_bucket (lb_base: their layout).  Event c of a scenario writes
// the entries b with key[c - 1] < b << shift <= key[c] (keys are non-negative
// and increasing); the entries past the last key hold n_ev.  cbase[s]: the
// scenario's first event in key (cbase[nscen]: the total).
__global__ void __launch_bounds__(BLK) k_ev_buckets(const StreamDev* st, int nscen, const int64_t* lb_base,
                                                    const int64_t* cbase, const int64_t* key, int* bucket,
                                                    const int64_t* bucket_base, const int* cond) {
  if (cond && *cond == 0) return;
  const int64_t stride = (int64_t)gridDim.x * BLK;
  const int64_t nlive = lb_base[nscen];
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < nlive; i += stride) {
    const int s = scen_of(lb_base, nscen, i);
    const int64_t b = i - lb_base[s], n = st[s].n_ev;
    if (n == 0 || b > (st[s].ev_key[n - 1] >> st[s].ev_bshift)) bucket[bucket_base[s] + b] = (int)n;
  }
  const int64_t nev = cbase[nscen];
  for (int64_t c = (int64_t)blockIdx.x * BLK + threadIdx.x; c < nev; c += stride) {
    const int s = scen_of(cbase, nscen, c);
    const int64_t l = c - cbase[s], nb = st[s].n_bucket;
    const int sh = st[s].ev_bshift;
    const int64_t lo = l == 0 ? 0 : (key[c - 1] >> sh) + 1;
    int64_t hi = key[c] >> sh;
    hi = hi < nb - 1 ? hi : nb - 1;
    int* bk = bucket + bucket_base[s];
    for (int64_t b = lo; b <= hi; ++b) bk[b] = (int)l;
  }
}

}  // namespace cyl
