// PCG32 stream and normal samplers shared by host and device code.
//
// The generator is the reference's global stream (cylag/lsm/random_utils.pyx:12-53):
// XSH-RR 64/32 with multiplier 6364136223846793005.  On the device every
// particle jumps from the step's base state to its own offset in the serial
// stream, so the draws equal those of the serial pool-order loop
// (cylag/core/solver.pyx:459).  Jumps use a per-stream table of the affine
// maps for 2^k steps, which is the same map pcg32_advance builds bit by bit.
#pragma once
#include <stdint.h>

#ifndef CYL_HD
#define CYL_HD __host__ __device__
#endif

namespace cyl {

constexpr uint64_t PCG_MULT = 6364136223846793005ULL;
constexpr double TWO_SQRT3 = 3.4641016151377544;  // constants.pyx:24
constexpr double ZIG_R = 3.442619855899;          // random_utils.pyx:113

struct Pcg {
  uint64_t state, inc;
};

CYL_HD inline uint32_t pcg_next(Pcg& g) {
  // random_utils.pyx:22-28
  uint64_t old = g.state;
  g.state = old * PCG_MULT + g.inc;
  uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
  uint32_t rot = (uint32_t)(old >> 59);
  return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

CYL_HD inline void pcg_seed(Pcg& g, uint64_t seed, uint64_t stream) {
  // random_utils.pyx:30-37
  g.state = 0;
  g.inc = (stream << 1) | 1u;
  pcg_next(g);
  g.state += seed;
  pcg_next(g);
}

// affine map of `delta` steps: state -> mult*state + plus   (random_utils.pyx:39-53)
CYL_HD inline void pcg_advance_map(uint64_t inc, uint64_t delta, uint64_t& mult, uint64_t& plus) {
  uint64_t am = 1, ap = 0, cm = PCG_MULT, cp = inc;
  while (delta > 0) {
    if (delta & 1) {
      am *= cm;
      ap = ap * cm + cp;
    }
    cp = (cm + 1) * cp;
    cm *= cm;
    delta >>= 1;
  }
  mult = am;
  plus = ap;
}

CYL_HD inline void pcg_advance(Pcg& g, uint64_t delta) {
  uint64_t m, p;
  pcg_advance_map(g.inc, delta, m, p);
  g.state = m * g.state + p;
}

// Jump table for one stream: entry k advances by 2^k.
struct JumpTable {
  uint64_t mult[64];
  uint64_t plus[64];
};

inline void make_jump_table(uint64_t inc, JumpTable& jt) {
  uint64_t cm = PCG_MULT, cp = inc;
  for (int k = 0; k < 64; ++k) {
    jt.mult[k] = cm;
    jt.plus[k] = cp;
    cp = (cm + 1) * cp;
    cm *= cm;
  }
}

// state after `delta` draws from `base` (maps for distinct powers of two commute)
CYL_HD inline uint64_t pcg_jump(uint64_t base, uint64_t delta, const JumpTable& jt) {
  uint64_t s = base;
  int k = 0;
  while (delta) {
    if (delta & 1) s = jt.mult[k] * s + jt.plus[k];
    delta >>= 1;
    ++k;
  }
  return s;
}

// Byte-window jump table: entry [l][v] advances by v << (8 l), l = 0..3, so a
// jump below 2^32 is four affine steps with no branch (entry [l][0] is the identity).
struct JumpTable8 {
  uint64_t mult[4][256];
  uint64_t plus[4][256];
};

inline void make_jump_table8(uint64_t inc, JumpTable8& t) {
  for (int l = 0; l < 4; ++l)
    for (int v = 0; v < 256; ++v) pcg_advance_map(inc, (uint64_t)v << (8 * l), t.mult[l][v], t.plus[l][v]);
}

CYL_HD inline uint64_t pcg_jump8(uint64_t base, uint64_t delta, const JumpTable8& t8, const JumpTable& jt) {
  uint64_t s = base;
  for (int l = 0; l < 4; ++l) {
    uint32_t v = (uint32_t)(delta >> (8 * l)) & 255u;
    s = t8.mult[l][v] * s + t8.plus[l][v];
  }
  uint64_t hi = delta >> 32;
  int k = 32;
  while (hi) {
    if (hi & 1) s = jt.mult[k] * s + jt.plus[k];
    hi >>= 1;
    ++k;
  }
  return s;
}

CYL_HD inline double u32_to_uniform(uint32_t u) { return u * (1.0 / 4294967296.0); }

// rectangular normal approximation, random_utils.pyx:217-219
CYL_HD inline double rect_normal(Pcg& g) { return (u32_to_uniform(pcg_next(g)) - 0.5) * TWO_SQRT3; }

// Ziggurat tables (computed on the host with the host libm, then uploaded)
struct ZigTables {
  uint32_t kn[128];
  double wn[128];
  double fn[128];
};

// random_utils.pyx:154-194; uses exp/log of the compiling target
template <class ExpF, class LogF>
CYL_HD inline double zig_normal_t(Pcg& g, const ZigTables& z, ExpF ex, LogF lg) {
  for (;;) {
    int32_t hz = (int32_t)pcg_next(g);
    uint32_t iz = (uint32_t)hz & 127u;
    uint32_t ahz = hz < 0 ? (uint32_t)(-(int64_t)hz) : (uint32_t)hz;  // libc abs(INT_MIN) wraps
    if (ahz < z.kn[iz]) return hz * z.wn[iz];
    if (iz == 0) {
      for (;;) {
        double x = -lg(u32_to_uniform(pcg_next(g))) * 0.2904764;
        double y = -lg(u32_to_uniform(pcg_next(g)));
        if (y + y >= x * x) return hz > 0 ? ZIG_R + x : -ZIG_R - x;
      }
    }
    double x = hz * z.wn[iz];
    if (z.fn[iz] + u32_to_uniform(pcg_next(g)) * (z.fn[iz - 1] - z.fn[iz]) < ex(-0.5 * x * x)) return x;
  }
}

// Marsaglia polar attempt loop, random_utils.pyx:90-108 (pair) / 72-88 (single)
template <class LogF, class SqrtF>
CYL_HD inline void polar_pair_t(Pcg& g, double& a, double& b, LogF lg, SqrtF sq) {
  double x1 = 0., x2 = 0., w = 2.0;
  while (w >= 1.0 || w == 0.0) {
    x1 = 2.0 * u32_to_uniform(pcg_next(g)) - 1.0;
    x2 = 2.0 * u32_to_uniform(pcg_next(g)) - 1.0;
    w = x1 * x1 + x2 * x2;
  }
  w = sq((-2.0 * lg(w)) / w);
  a = x1 * w;
  b = x2 * w;
}

// Is a polar attempt whose two uniforms are (u1, u2) accepted?
CYL_HD inline bool polar_accept(uint32_t u1, uint32_t u2) {
  double x1 = 2.0 * u32_to_uniform(u1) - 1.0;
  double x2 = 2.0 * u32_to_uniform(u2) - 1.0;
  double w = x1 * x1 + x2 * x2;
  return !(w >= 1.0 || w == 0.0);
}

}  // namespace cyl
