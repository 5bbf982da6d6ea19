// Host draws of the memetic generation in the reference order, bit-identical to NumPy's
// Generator(PCG64) calls that the reference strategies make each generation
// (tournament_selection.py:113, pmx_crossover.py:155 via route_helpers.py:74-108,
// inversion_mutation.py:84-92):
//
//   cand  = integers(0, P, (P, k), int32)          tournament candidates
//   first = integers(0, n, P // 2, int32)          PMX cut pairs: first point,
//   sh    = integers(0, n - 1, P // 2, int32)      second = sh + (sh >= first)
//   probs = random(P)                              inversion: which rows mutate,
//   first = integers(0, n, P, int32)               and their position pairs
//   sh    = integers(0, n - 1, P, int32)
//
// Published algorithms: PCG64 is O'Neill's PCG XSL-RR 128/64 (128-bit LCG, multiplier
// 0x2360ED051FC65DA44385DF649FCCF645; the output is taken after the step); a 32-bit draw is the
// low half of a 64-bit output with the high half kept for the next 32-bit draw (the generator
// state's has_uint32 / uinteger); a bounded 32-bit integer uses Lemire's multiply-and-reject
// method (2019) with the rejection threshold 2^32 mod range; a double is (u64 >> 11) * 2^-53.
// NumPy makes the same calls one at a time, each with a fixed Python/argument cost of several
// microseconds; here a whole run of generations is one call.
#include <stdint.h>
#include <string.h>

#include "../../include/tspmi.h"

namespace {

typedef unsigned __int128 u128;

struct Pcg64 {
  u128 state, inc;
  int has32;
  uint32_t buf32;

  uint64_t next64() {
    const u128 mult = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;
    state = state * mult + inc;
    const uint64_t x = (uint64_t)(state >> 64) ^ (uint64_t)state;
    const unsigned rot = (unsigned)(state >> 122);
    return (x >> rot) | (x << ((64 - rot) & 63));
  }
  uint32_t next32() {
    if (has32) {
      has32 = 0;
      return buf32;
    }
    const uint64_t v = next64();
    has32 = 1;
    buf32 = (uint32_t)(v >> 32);
    return (uint32_t)v;
  }
  double next_double() { return (double)(next64() >> 11) * (1.0 / 9007199254740992.0); }
  // integers(0, bound) for bound >= 1 (a bound of 1 draws nothing).
  uint32_t below(uint32_t bound) {
    if (bound <= 1) return 0;
    uint64_t m = (uint64_t)next32() * bound;
    uint32_t left = (uint32_t)m;
    if (left < bound) {
      const uint32_t threshold = (uint32_t)(0u - bound) % bound;
      while (left < threshold) {
        m = (uint64_t)next32() * bound;
        left = (uint32_t)m;
      }
    }
    return (uint32_t)(m >> 32);
  }
};

}  // namespace

extern "C" int tspmi_draw_generations(uint64_t* state, int G, int P, int n, int k, int32_t* cand, int32_t* cut_lo,
                                      int32_t* cut_hi, double* probs, int32_t* pos1, int32_t* pos2,
                                      uint64_t* states_out) {
  if (state == nullptr || G < 0 || P < 1 || n < 2 || k < 1) return TSPMI_EINVAL;
  if (G > 0 && (cand == nullptr || probs == nullptr || pos1 == nullptr || pos2 == nullptr ||
                cut_lo == nullptr || cut_hi == nullptr))
    return TSPMI_EINVAL;
  Pcg64 g;
  g.state = ((u128)state[0] << 64) | state[1];
  g.inc = ((u128)state[2] << 64) | state[3];
  g.has32 = state[4] != 0;
  g.buf32 = (uint32_t)state[5];
  const int half = P / 2;
  const int hw = half > 0 ? half : 1;
  for (int t = 0; t < G; ++t) {
    int32_t* c = cand + (size_t)t * P * k;
    for (long x = 0; x < (long)P * k; ++x) c[x] = (int32_t)g.below((uint32_t)P);
    int32_t* lo = cut_lo + (size_t)t * hw;
    int32_t* hi = cut_hi + (size_t)t * hw;
    if (half > 0) {
      // first points, then the shifted second points: two separate calls in the reference
      for (int q = 0; q < half; ++q) lo[q] = (int32_t)g.below((uint32_t)n);
      for (int q = 0; q < half; ++q) {
        const int32_t first = lo[q];
        int32_t second = (int32_t)g.below((uint32_t)(n - 1));
        second += (second >= first);
        lo[q] = first < second ? first : second;
        hi[q] = first < second ? second : first;
      }
    } else {
      lo[0] = hi[0] = 0;
    }
    double* pr = probs + (size_t)t * P;
    for (int r = 0; r < P; ++r) pr[r] = g.next_double();
    int32_t* a = pos1 + (size_t)t * P;
    int32_t* b = pos2 + (size_t)t * P;
    for (int r = 0; r < P; ++r) a[r] = (int32_t)g.below((uint32_t)n);
    for (int r = 0; r < P; ++r) {
      int32_t s = (int32_t)g.below((uint32_t)(n - 1));
      b[r] = s + (s >= a[r]);
    }
    if (states_out != nullptr) {
      uint64_t* o = states_out + (size_t)t * 6;
      o[0] = (uint64_t)(g.state >> 64);
      o[1] = (uint64_t)g.state;
      o[2] = (uint64_t)(g.inc >> 64);
      o[3] = (uint64_t)g.inc;
      o[4] = (uint64_t)g.has32;
      o[5] = g.buf32;
    }
  }
  state[0] = (uint64_t)(g.state >> 64);
  state[1] = (uint64_t)g.state;
  state[4] = (uint64_t)g.has32;
  state[5] = g.buf32;
  return TSPMI_OK;
}
