// IEEE binary16 for the host emulation of the kernels (round to nearest even).
#pragma once

#include <stdint.h>
#include <string.h>

struct __half {
  uint16_t bits;
};

inline __half __ushort_as_half(unsigned short u) { return __half{u}; }
inline unsigned short __half_as_ushort(__half h) { return h.bits; }

inline float __half2float(__half h) {
  const uint32_t s = (uint32_t)(h.bits >> 15) << 31;
  const uint32_t e = (h.bits >> 10) & 0x1f;
  uint32_t m = h.bits & 0x3ff;
  uint32_t u;
  if (e == 0) {
    if (m == 0) {
      u = s;
    } else {  // subnormal: normalise
      int k = 0;
      while (!(m & 0x400)) {
        m <<= 1;
        ++k;
      }
      u = s | ((uint32_t)(127 - 15 - k + 1) << 23) | ((m & 0x3ff) << 13);
    }
  } else if (e == 31) {
    u = s | 0x7f800000u | (m << 13);
  } else {
    u = s | ((e + 112) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &u, 4);
  return f;
}

inline __half __float2half_rn(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  const uint16_t s = (uint16_t)((u >> 16) & 0x8000);
  const uint32_t a = u & 0x7fffffffu;
  if (a >= 0x7f800000u) return __half{(uint16_t)(s | (a > 0x7f800000u ? 0x7e00 : 0x7c00))};
  if (a >= 0x477ff000u) return __half{(uint16_t)(s | 0x7c00)};  // rounds to infinity
  if (a < 0x33000001u) return __half{s};                         // rounds to zero
  const int e = (int)(a >> 23) - 127;
  uint32_t m = (a & 0x7fffff) | 0x800000;
  int shift;
  uint32_t he;
  if (e < -14) {  // subnormal half
    shift = 13 + (-14 - e);
    he = 0;
  } else {
    shift = 13;
    he = (uint32_t)(e + 15);
    m &= 0x7fffff;
  }
  uint32_t q = m >> shift;
  const uint32_t rem = m & ((1u << shift) - 1);
  const uint32_t half = 1u << (shift - 1);
  if (rem > half || (rem == half && (q & 1))) ++q;
  uint32_t out = (he << 10) + q;  // a mantissa carry moves into the exponent
  return __half{(uint16_t)(s | out)};
}
