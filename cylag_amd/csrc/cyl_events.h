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
// Pipeline (rng_method 2): count -> scan -> write and sample -> reach
// max-scan -> resolve -> (extra, count) scan -> compact -> buckets.
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
  const int64_t* scen_units;  // units needed per scenario this step
  const int64_t* chunk_base;  // first chunk of each scenario (host-sized capacity)
  int nscen;
  int64_t nchunks;
  int* chunk_count;           // candidates per chunk
  int64_t* chunk_off;         // scanned
  int64_t* nev;               // total candidates (scan total)
  int64_t* ev_begin;          // [nscen + 1] first candidate of each scenario
  int64_t* ev_pos;            // candidate position (relative to the scenario's window start)
  uint64_t* ev_state;         // generator state at the candidate (zig)
  int64_t* ev_len;            // length of the normal starting at the candidate (zig)
  double* ev_val;             // its value (zig)
  int64_t* ev_key;            // its reach key (zig)
  int64_t* chunk_live;        // end of the live chunks (set by the count pass)
  int64_t ev_cap;
  int* overflow;
  const int* cond;            // run only when *cond != 0
};

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

__device__ __forceinline__ bool zig_fast(uint32_t u, const uint32_t* kn) {
  int32_t hz = (int32_t)u;
  uint32_t iz = (uint32_t)hz & 127u;
  uint32_t ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;  // libc abs(INT_MIN) wraps
  return ahz < kn[iz];
}

// Scan of the stream window, one chunk of ZCHUNK positions per thread, blocks
// striding over spans of BLK chunks.  WRITE = false: count candidates per
// chunk, set each scenario's window, and (block 0) the end of the live chunks
// (no chunk at or past it holds a position of any window).  WRITE = true:
// write the candidates of the live chunks at the scanned offsets; under
// rng_method 2 the block then samples the candidates of its span (one
// contiguous range of the table, one candidate per thread): the value and
// length of the Ziggurat normal that starts there and its reach key
// (scenario << REACH_SHIFT) | (pos + len - 1).
// A candidate whose normal reaches the base strip's tail loop (two correctly
// rounded logarithms per attempt; a few percent of candidates) is queued in
// LDS and finished by the block at the end, so that the waves sampling the
// other candidates do not wait on it.
constexpr int TAIL_CAP = 256;
struct TailRec {
  uint64_t state;
  int64_t i;
  int32_t hz, ncalls;
};

__device__ __forceinline__ void ev_finish(const EventArgs& e, int64_t i, double v, int64_t ncalls) {
  e.ev_val[i] = v;
  e.ev_len[i] = ncalls;
  e.ev_key[i] = e.ev_key[i] | (e.ev_pos[i] + ncalls - 1);
}

template <int METHOD, bool WRITE>
__global__ void __launch_bounds__(BLK) k_ev_scan(EventArgs e) {
  if (e.cond && *e.cond == 0) return;
  __shared__ JumpTable8 t8;
  __shared__ uint32_t kn[128];
  __shared__ int64_t lds[BLK / 64];
  __shared__ TailRec tails[WRITE && METHOD == 2 ? TAIL_CAP : 1];
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
  __syncthreads();
  if (!WRITE && blockIdx.x == 0) {
    int64_t live = 0;
    for (int q = threadIdx.x; q < e.nscen; q += BLK) {
      int64_t W = window_for(METHOD, e.scen_units[q]);
      e.stream[q].window = W;
      if (W > 0) {
        int64_t end = e.chunk_base[q] + (W + ZCHUNK - 1) / ZCHUNK;
        live = end > live ? end : live;
      }
    }
    int64_t tot;
    block_excl_scan<OpMax, BLK>(live, tot, lds);
    // a window past its scenario's chunk capacity is reported (overflow) and
    // never scanned past the table
    if (threadIdx.x == 0) *e.chunk_live = tot < 0 ? 0 : (tot < e.nchunks ? tot : e.nchunks);
  }
  // the count pass visits every chunk; the write pass only the live ones
  // (the chunk scan is exact up to index chunk_live, which holds the total)
  const int64_t nch = WRITE ? *e.chunk_live : e.nchunks;
  if (WRITE && blockIdx.x == 0)
    for (int q = threadIdx.x; q <= e.nscen; q += BLK) {
      int64_t v = q < e.nscen && e.chunk_base[q] < nch ? e.chunk_off[e.chunk_base[q]] : *e.nev;
      e.ev_begin[q] = v < e.ev_cap ? v : e.ev_cap;  // an overflow is reported, never indexed past
    }
  for (int64_t span = (int64_t)blockIdx.x * BLK; span < nch; span += (int64_t)gridDim.x * BLK) {
    const int64_t ch = span + threadIdx.x;
    if (ch < nch) {
      int s = scen_of(e.chunk_base, e.nscen, ch);
      int64_t local = ch - e.chunk_base[s];
      int64_t W = window_for(METHOD, e.scen_units[s]);
      int64_t cap_chunks = (s + 1 < e.nscen ? e.chunk_base[s + 1] : e.nchunks) - e.chunk_base[s];
      int64_t p0 = local * ZCHUNK;
      int cnt = 0;
      int64_t w = WRITE ? e.chunk_off[ch] : 0;
      const int64_t skey = (int64_t)s << REACH_SHIFT;
      if (p0 < W) {
        if (!WRITE && (local + 1) * ZCHUNK < W && local + 1 == cap_chunks) atomicOr(e.overflow, 1);
        const StreamDev& st = e.stream[s];
        Pcg g;
        g.inc = st.inc;
        if (METHOD == 2) {
          g.state = st.inc == inc0 ? pcg_jump8(st.base, (uint64_t)p0, t8, e.jt[s]) : pcg_jump(st.base, (uint64_t)p0, e.jt[s]);
          for (int q = 0; q < ZCHUNK; ++q) {
            uint64_t at = g.state;
            if (zig_fast(pcg_next(g), kn)) continue;
            if (WRITE) {
              if (w < e.ev_cap) {
                e.ev_pos[w] = p0 + q;
                e.ev_state[w] = at;
                e.ev_key[w] = skey;
              } else {
                atomicOr(e.overflow, 2);
              }
              ++w;
            } else {
              ++cnt;
            }
          }
        } else {
          // polar: attempt a uses positions 2a, 2a+1
          int64_t a0 = p0 / 2;
          uint64_t d = (uint64_t)(2 * a0);
          g.state = st.inc == inc0 ? pcg_jump8(st.base, d, t8, e.jt[s]) : pcg_jump(st.base, d, e.jt[s]);
          for (int q = 0; q < ZCHUNK / 2; ++q) {
            uint32_t u1 = pcg_next(g), u2 = pcg_next(g);
            if (polar_accept(u1, u2)) continue;
            if (WRITE) {
              if (w < e.ev_cap) e.ev_pos[w] = a0 + q;
              else atomicOr(e.overflow, 2);
              ++w;
            } else {
              ++cnt;
            }
          }
        }
      }
      if (!WRITE) e.chunk_count[ch] = cnt;
    }
    if (WRITE && METHOD == 2) {
      // the span's candidates, written above by this block, in one range
      __syncthreads();
      int64_t c0 = e.chunk_off[span];
      int64_t c1 = span + BLK < nch ? e.chunk_off[span + BLK] : *e.nev;
      c0 = c0 < e.ev_cap ? c0 : e.ev_cap;
      c1 = c1 < e.ev_cap ? c1 : e.ev_cap;
      for (int64_t i = c0 + threadIdx.x; i < c1; i += BLK) {
        Draws d;
        d.zt = e.zt;
        d.pre_pos = nullptr;
        d.g.inc = e.stream[e.ev_key[i] >> REACH_SHIFT].inc;
        d.g.state = e.ev_state[i];
        d.ncalls = 0;
        double v;
        int32_t hz;
        if (d.zig_head(v, hz)) {
          ev_finish(e, i, v, d.ncalls);
        } else {
          int k = atomicAdd(&ntail, 1);
          if (k < TAIL_CAP) tails[k] = TailRec{d.g.state, i, hz, (int32_t)d.ncalls};
          else ev_finish(e, i, d.zig_tail(hz), d.ncalls);  // queue full: finish here
        }
      }
    }
  }
  if (WRITE && METHOD == 2) {
    __syncthreads();
    const int nt = ntail < TAIL_CAP ? ntail : TAIL_CAP;
    for (int k = threadIdx.x; k < nt; k += BLK) {
      const TailRec r = tails[k];
      Draws d;
      d.zt = e.zt;
      d.pre_pos = nullptr;
      d.g.inc = e.stream[e.ev_key[r.i] >> REACH_SHIFT].inc;
      d.g.state = r.state;
      d.ncalls = r.ncalls;
      double v = d.zig_tail(r.hz);
      ev_finish(e, r.i, v, d.ncalls);
    }
  }
}

// Ziggurat chain resolution.  Candidate p is off the chain iff an on-chain
// slow start q < p has q + len(q) - 1 >= p.  A candidate that no earlier
// candidate reaches is surely on-chain; from there the chain is followed
// forward.  reach (exclusive max-scan of the keys, tile prefix in tmax) gives
// the furthest position reached by earlier candidates of the same scenario.
// Writes (len - 1) << 32 | 1 for on-chain candidates, 0 otherwise.
__global__ void __launch_bounds__(BLK) k_ev_resolve(const int64_t* pos, const int64_t* len, const int64_t* ev_begin,
                                                    int nscen, const int64_t* nev, const int64_t* reach,
                                                    const int64_t* tmax, int64_t* pk, const int* cond) {
  if (cond && *cond == 0) return;
  const int64_t ne = ev_begin[nscen];  // candidates, clamped to the capacity
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < ne; i += (int64_t)gridDim.x * BLK) {
    int s = scen_of(ev_begin, nscen, i);
    int64_t b = ev_begin[s];
    auto reach_of = [&](int64_t j) -> int64_t {
      int64_t m = reach[j], t = tmax[j / SCAN_TILE];
      m = t > m ? t : m;
      return (m >= 0 && (m >> REACH_SHIFT) == s) ? (m & ((1LL << REACH_SHIFT) - 1)) : -1;
    };
    int64_t j0 = i;
    while (j0 > b && reach_of(j0) >= pos[j0]) --j0;
    int64_t cur_end = pos[j0] + len[j0] - 1;
    int on = 1;
    for (int64_t l = j0 + 1; l <= i; ++l) {
      if (pos[l] > cur_end) {
        cur_end = pos[l] + len[l] - 1;
        on = 1;
      } else {
        on = 0;
      }
    }
    pk[i] = on ? ((len[i] - 1) << 32) | 1 : 0;
  }
}

// Point each scenario's stream at its slice of the event tables (threads
// below nscen) and compact the on-chain candidates: normal index, cumulative
// extra positions (inclusive) and value.  ec = exclusive scan of pk.
__global__ void __launch_bounds__(BLK) k_ev_compact(StreamDev* st, int nscen, const int64_t* ev_begin,
                                                    const int64_t* nev, const int64_t* pos, const int64_t* len,
                                                    const double* val, const int64_t* pk, const int64_t* ec,
                                                    int64_t* key, int64_t* xtra, double* cval, const int* bucket,
                                                    const int64_t* bucket_base, int bshift, const int* cond) {
  if (cond && *cond == 0) return;
  const int64_t ne = ev_begin[nscen];  // candidates, clamped to the capacity
  const int64_t end = ne > nscen ? ne : nscen;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < end; i += (int64_t)gridDim.x * BLK) {
    if (i < nscen) {
      int64_t c0 = ec[ev_begin[i]] & EC_MASK, c1 = ec[ev_begin[i + 1]] & EC_MASK;
      st[i].ev_key = key + c0;
      st[i].ev_extra = xtra + c0;
      st[i].ev_val = cval + c0;
      st[i].n_ev = c1 - c0;
      st[i].ev_bucket = bucket + bucket_base[i];
      st[i].n_bucket = bucket_base[i + 1] - bucket_base[i];
      st[i].ev_bshift = bshift;
    }
    if (i >= ne || !(pk[i] & 1)) continue;
    int s = scen_of(ev_begin, nscen, i);
    int64_t eb = ec[ev_begin[s]] >> 32;
    int64_t e_excl = (ec[i] >> 32) - eb;
    int64_t c = ec[i] & EC_MASK;
    key[c] = pos[i] - e_excl;
    xtra[c] = e_excl + len[i] - 1;
    cval[c] = val[i];
  }
}

// polar: key_j = r_j - j (j counted within the scenario), and the stream slices
__global__ void __launch_bounds__(BLK) k_ev_polar(StreamDev* st, int nscen, const int64_t* ev_begin,
                                                  const int64_t* nev, const int64_t* pos, int64_t* key,
                                                  const int* bucket, const int64_t* bucket_base, int bshift,
                                                  const int* cond) {
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
      st[i].n_bucket = bucket_base[i + 1] - bucket_base[i];
      st[i].ev_bshift = bshift;
    }
    if (i >= ne) continue;
    int s = scen_of(ev_begin, nscen, i);
    key[i] = pos[i] - (i - ev_begin[s]);
  }
}

// bucket index of every scenario's event keys: entry b = lower_bound(b << shift);
// buckets past the last key (the step's units end there) hold n_ev directly
__global__ void __launch_bounds__(BLK) k_ev_buckets(const StreamDev* st, int nscen, const int64_t* bucket_base,
                                                    int64_t ntot, int* bucket, const int* cond) {
  if (cond && *cond == 0) return;
  for (int64_t i = (int64_t)blockIdx.x * BLK + threadIdx.x; i < ntot; i += (int64_t)gridDim.x * BLK) {
    int s = scen_of(bucket_base, nscen, i);
    int64_t b = i - bucket_base[s], u = b << st[s].ev_bshift;
    const int64_t* k = st[s].ev_key;
    int64_t lo = 0, hi = st[s].n_ev;
    if (hi > 0 && k[hi - 1] < u) lo = hi;
    while (lo < hi) {
      int64_t mid = (lo + hi) >> 1;
      if (k[mid] < u) lo = mid + 1;
      else hi = mid;
    }
    bucket[i] = (int)lo;
  }
}

}  // namespace cyl
