// GSL gsl_rng_ranlxd1 (M. Luescher's second-generation RANLUX, 48-bit
// variant, luxury p = 202) as used by the reference for the seed table and
// the per-pencil white noise (Aarambam/LPT/Basis/main.c:317-319, 392-401).
//
// State: twelve 48-bit integers and a borrow.  The recursion
//   x_n = x_{n-5} - x_{n-12} - c_{n-1}  (mod 2^48)
// is exact in 64-bit integers (GSL holds the same values as doubles k*2^-48).
// After every twelve outputs the generator runs 202 steps and then returns
// the last twelve values in generation order.  The block below always starts
// an update at register slot 0 and renames the registers afterwards, so every
// register index is a compile-time constant.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define AAR_HD __host__ __device__ __forceinline__
#else
#define AAR_HD inline
#endif

struct Ranlxd {
  uint64_t s[12];
  uint64_t carry;
};

static constexpr uint64_t kMask48 = (1ull << 48) - 1;
static constexpr double kTwoM48 = 1.0 / 281474976710656.0;

// gsl_rng_set for ranlxd (GSL rng/ranlxd.c ranlxd_set_lux): 31 seed bits feed
// the shift register b_{n+31} = b_n ^ b_{n+18}; each state word takes the
// complements of 48 consecutive register bits, most significant first.
AAR_HD void ranlxd_seed(Ranlxd& r, uint32_t seed) {
  uint32_t w = seed ? seed : 1u;  // seed 0 maps to 1
  w &= 0x7fffffffu;               // the reference's seeds are < 2^31
#pragma unroll
  for (int k = 0; k < 12; ++k) {
    uint64_t x = 0;
#pragma unroll 4
    for (int l = 0; l < 48; ++l) {
      uint32_t b = w & 1u;
      uint32_t nb = (w ^ (w >> 18)) & 1u;
      w = (w >> 1) | (nb << 30);
      x = 2 * x + (1u - b);
    }
    r.s[k] = x;
  }
  r.carry = 0;
}

AAR_HD void ranlxd_step(uint64_t& dst, uint64_t lag5, uint64_t& c) {
  uint64_t y = lag5 - dst - c;
  c = y >> 63;
  dst = y & kMask48;
}

// One luxury update (202 steps), then rename so outputs are s[0..11].
AAR_HD void ranlxd_update(Ranlxd& r) {
  uint64_t c = r.carry;
  uint64_t* s = r.s;
#pragma unroll 1
  for (int b = 0; b < 16; ++b) {
#pragma unroll
    for (int q = 0; q < 12; ++q) ranlxd_step(s[q], s[(q + 7) % 12], c);
  }
#pragma unroll
  for (int q = 0; q < 10; ++q) ranlxd_step(s[q], s[(q + 7) % 12], c);
  uint64_t t[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) t[q] = s[(q + 10) % 12];
#pragma unroll
  for (int q = 0; q < 12; ++q) s[q] = t[q];
  r.carry = c;
}

#if defined(__HIPCC__)
// The same update on the device, the 48-bit words held as 32-bit halves: a step is two
// subtractions with borrow (the borrow of the high half is the next step's c, since the words
// are below 2^48) and a mask, where the 64-bit form takes five instructions.
__device__ __forceinline__ void ranlxd_update_dev(Ranlxd& r) {
  uint32_t lo[12], hi[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    lo[q] = (uint32_t)r.s[q];
    hi[q] = (uint32_t)(r.s[q] >> 32);
  }
  uint32_t c = (uint32_t)r.carry;
  auto step = [&](int q) {
    const int l = (q + 7) % 12;
    uint32_t b0, b1;
    const uint32_t yl = __builtin_subc(lo[l], lo[q], c, &b0);
    const uint32_t yh = __builtin_subc(hi[l], hi[q], b0, &b1);
    c = b1;
    lo[q] = yl;
    hi[q] = yh & 0xffffu;
  };
#pragma unroll 1
  for (int b = 0; b < 16; ++b) {
#pragma unroll
    for (int q = 0; q < 12; ++q) step(q);
  }
#pragma unroll
  for (int q = 0; q < 10; ++q) step(q);
#pragma unroll
  for (int q = 0; q < 12; ++q) {
    const int o = (q + 10) % 12;
    r.s[q] = ((uint64_t)hi[o] << 32) | lo[o];
  }
  r.carry = c;
}
#endif

// Sequential generator with the GSL calling semantics (host seed table and
// the rare redraw path of the Gaussian kernel).
struct RanlxdStream {
  Ranlxd r;
  int pos;  // next output index in s[], 12 = exhausted
  AAR_HD void seed(uint32_t s) {
    ranlxd_seed(r, s);
    pos = 12;
  }
  AAR_HD double uniform() {
    if (pos == 12) {
      ranlxd_update(r);
      pos = 0;
    }
    return (double)r.s[pos++] * kTwoM48;
  }
};
