// lj_gen.hpp -- seeded, counter-based synthetic columns (host and device).
//
// Value j of stream `stream` under seed `seed` is a pure function of
// (seed, stream, j), so a column can be produced in any order, in parallel, on
// the host or on the device, and both give the same bits.
//
//  * SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
//    generators", OOPSLA 2014): output j of the generator whose state starts at
//    `key` is mix64(key + (j+1)*gamma).
//  * Bounded integers: Lemire, "Fast random integer generation in an interval"
//    (ACM TOMACS 2019), multiply-shift with the exact rejection step; a
//    rejected draw is replaced by the next SplitMix64 output of a private chain.
//  * Calendar: days since 1970-01-01 -> (y, m, d) by H. Hinnant's
//    civil_from_days ("chrono-Compatible Low-Level Date Algorithms").
#pragma once

#include <stdint.h>

#include "lipjoin/block/hash.hpp"  // LJ_HD, mulhi64

namespace lj {

#define LJ_GAMMA 0x9E3779B97F4A7C15ull

LJ_HD uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

LJ_HD uint64_t stream_key(uint64_t seed, uint64_t stream) {
  return mix64(seed ^ mix64(stream * 0xD1B54A32D192ED03ull + 0x632BE59BD9B4E019ull));
}

LJ_HD uint64_t rand_at(uint64_t key, uint64_t j) { return mix64(key + (j + 1) * LJ_GAMMA); }

// Uniform integer in [0, range), 1 <= range <= 2^32.
LJ_HD uint64_t bounded_at(uint64_t key, uint64_t j, uint64_t range) {
  uint64_t x = rand_at(key, j);
  uint64_t lo = x * range;
  uint64_t hi = mulhi64(x, range);
  if (lo < range) {
    uint64_t t = (0ull - range) % range;
    while (lo < t) {
      x = mix64(x + LJ_GAMMA);
      lo = x * range;
      hi = mulhi64(x, range);
    }
  }
  return hi;
}

// 1992-01-01 as days since 1970-01-01.
#define LJ_DAY0 8035

LJ_HD void civil_from_days(int64_t z, int64_t& y, int64_t& m, int64_t& d) {
  z += 719468;
  const int64_t era = (z >= 0 ? z : z - 146096) / 146097;
  const int64_t doe = z - era * 146097;
  const int64_t yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
  const int64_t doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
  const int64_t mp = (5 * doy + 2) / 153;
  d = doy - (153 * mp + 2) / 5 + 1;
  m = mp < 10 ? mp + 3 : mp - 9;
  y = yoe + era * 400 + (m <= 2);
}

LJ_HD int32_t datekey_of_day(int64_t day) {  // day 0 = 1992-01-01
  int64_t y, m, d;
  civil_from_days(day + LJ_DAY0, y, m, d);
  return (int32_t)(y * 10000 + m * 100 + d);
}

LJ_HD int32_t year_of_day(int64_t day) {
  int64_t y, m, d;
  civil_from_days(day + LJ_DAY0, y, m, d);
  return (int32_t)y;
}

enum GenKind : int32_t {
  GEN_UNIFORM = 0,      // lo + U[0, hi-lo]
  GEN_DATEKEY = 1,      // datekey(lo + U[0, hi))
  GEN_SEQUENCE = 2,     // lo + j
  GEN_YEAR_OF_DAY = 3,  // year(lo + j)
  GEN_DATEKEY_SEQ = 4   // datekey(lo + j)
};

LJ_HD int32_t gen_value(int32_t kind, uint64_t key, int64_t lo, int64_t hi, int64_t j) {
  switch (kind) {
    case GEN_UNIFORM: return (int32_t)(lo + (int64_t)bounded_at(key, (uint64_t)j, (uint64_t)(hi - lo + 1)));
    case GEN_DATEKEY: return datekey_of_day(lo + (int64_t)bounded_at(key, (uint64_t)j, (uint64_t)hi));
    case GEN_SEQUENCE: return (int32_t)(lo + j);
    case GEN_YEAR_OF_DAY: return year_of_day(lo + j);
    case GEN_DATEKEY_SEQ: return datekey_of_day(lo + j);
  }
  return 0;
}

}  // namespace lj
