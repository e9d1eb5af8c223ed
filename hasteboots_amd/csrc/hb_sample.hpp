// Counter-based samplers for the evaluation keys, shared by the host key generation and
// the GPU key-generation kernels so that both produce the same words.
//
// Every coefficient of every RLWE sample of the keys draws from its own ChaCha20 block
// (RFC 8439 block function): key from the seed, nonce (stream kind, row index, tag),
// block counter = position of the coefficient in that row's samples. Of the block's 16
// words, 0-3 give the uniform a-coefficient and 4-15 up to three Box-Muller attempts for
// the Gaussian error, so no coefficient depends on another's draws (the reference's
// FieldUniformSampler / FieldDiscreteGaussianSampler draw sequentially from one AES-CTR
// stream, algebra/src/random.rs:11-35, 156-170; its randomness is not reproduced here).
//
// uniform: floor(r * q / 2^128) for a 128-bit r, statistical distance below 2^-64 from
// uniform on [0, q). gaussian: N(0, sigma) rounded to the nearest integer, cut at 6 sigma
// (the reference's cut), negative values as q - |v|; if all three attempts fall outside
// the cut (probability about 1e-26) the error is 0.
//
// The floating-point steps use only IEEE-754 add, multiply and fused multiply-add, with
// contraction disabled, so host and device round identically:
//   det_log: x = m 2^e with m in (sqrt(1/2), sqrt(2)], ln m = 2 atanh(s), s = (m-1)/(m+1)
//     (the quotient by Newton-Raphson reciprocal steps), atanh by its Taylor series to
//     s^23 (|s| <= 0.172, truncation below 1e-18 relative);
//   det_cos2pi: reduction by the symmetries of cos(2 pi u) to 2 pi t, t in [0, 1/8] (exact
//     subtractions), then the Taylor series of cos or sin to degree 20 / 21;
//   det_sqrt: inverse square root by Newton-Raphson from a bit-level first guess, one
//     residual correction.
#pragma once
#include <cstdint>

#include "hb_field.hpp"

namespace hb {

constexpr uint32_t kKeyStreamTag = 0x68623032u;  // "hb02": counter-based key streams

HB_HD uint32_t rotl32_hd(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

// ChaCha20 block: 20 rounds over (constants, key, counter, nonce), plus the input.
HB_HD void chacha20_block(const uint32_t key[8], uint32_t counter, uint32_t n0, uint32_t n1,
                          uint32_t n2, uint32_t out[16]) {
  uint32_t x[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1],
                    key[2],      key[3],      key[4],      key[5],      key[6], key[7],
                    counter,     n0,          n1,          n2};
  uint32_t in[16];
  for (int i = 0; i < 16; i++) in[i] = x[i];
#define HB_QR(a, b, c, d)                    \
  x[a] += x[b]; x[d] = rotl32_hd(x[d] ^ x[a], 16); \
  x[c] += x[d]; x[b] = rotl32_hd(x[b] ^ x[c], 12); \
  x[a] += x[b]; x[d] = rotl32_hd(x[d] ^ x[a], 8);  \
  x[c] += x[d]; x[b] = rotl32_hd(x[b] ^ x[c], 7);
  for (int i = 0; i < 10; i++) {
    HB_QR(0, 4, 8, 12) HB_QR(1, 5, 9, 13) HB_QR(2, 6, 10, 14) HB_QR(3, 7, 11, 15)
    HB_QR(0, 5, 10, 15) HB_QR(1, 6, 11, 12) HB_QR(2, 7, 8, 13) HB_QR(3, 4, 9, 14)
  }
#undef HB_QR
  for (int i = 0; i < 16; i++) out[i] = x[i] + in[i];
}

// splitmix64 expansion of a 64-bit seed into the 256-bit ChaCha20 key.
HB_HD void derive_key(uint64_t seed, uint32_t key[8]) {
  uint64_t s = seed;
  for (int i = 0; i < 4; i++) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    key[2 * i] = (uint32_t)z;
    key[2 * i + 1] = (uint32_t)(z >> 32);
  }
}

HB_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  Gl::mul_wide(a, b, lo, hi);
  return hi;
}

// floor(r * q / 2^128), r = w[0] | w[1] << 32 | w[2] << 64 | w[3] << 96.
HB_HD uint64_t uniform_mod(const uint32_t w[4], uint64_t q) {
  const uint64_t lo = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
  const uint64_t hi = (uint64_t)w[2] | ((uint64_t)w[3] << 32);
  const uint64_t t = mulhi64(q, lo);
  uint64_t ulo, uhi;
  Gl::mul_wide(q, hi, ulo, uhi);
  const uint64_t s = ulo + t;
  return uhi + (s < ulo ? 1u : 0u);
}

HB_HD double bits_to_double(uint64_t b) {
  union { uint64_t u; double d; } v;
  v.u = b;
  return v.d;
}
HB_HD uint64_t double_to_bits(double d) {
  union { uint64_t u; double d; } v;
  v.d = d;
  return v.u;
}

// Uniform double in (0, 1] from 64 bits: ((bits >> 11) + 1) / 2^53.
HB_HD double unit_from(uint32_t lo, uint32_t hi) {
  const uint64_t b = (uint64_t)lo | ((uint64_t)hi << 32);
  return (double)((b >> 11) + 1) * 0x1.0p-53;
}

// Natural logarithm of a normal x in (0, 1].
HB_HD double det_log(double x) {
#pragma clang fp contract(off)
  const uint64_t b = double_to_bits(x);
  int e = (int)((b >> 52) & 0x7FF) - 1023;
  double m = bits_to_double((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);  // [1, 2)
  if (m > 0x1.6a09e667f3bcdp+0) {  // sqrt(2)
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;  // exact
  // s = f / (2 + f): reciprocal of d/2 in [0.79, 1.21] by Newton-Raphson, then one
  // residual step on the quotient
  const double d = 2.0 + f;
  const double h = d * 0.5;  // exact
  double r = 2.0 - h;
  for (int i = 0; i < 5; i++) r = __builtin_fma(r, __builtin_fma(-h, r, 1.0), r);
  const double rh = r * 0.5;  // ~ 1 / d
  double s = f * rh;
  s = __builtin_fma(__builtin_fma(-s, d, f), rh, s);
  const double t = s * s;
  // 2 atanh(s) = 2 s (1 + t/3 + t^2/5 + ... + t^11/23)
  double p = 0x1.642c8590b2164p-5;
  p = __builtin_fma(p, t, 0x1.8618618618618p-5);
  p = __builtin_fma(p, t, 0x1.af286bca1af28p-5);
  p = __builtin_fma(p, t, 0x1.e1e1e1e1e1e1ep-5);
  p = __builtin_fma(p, t, 0x1.1111111111111p-4);
  p = __builtin_fma(p, t, 0x1.3b13b13b13b14p-4);
  p = __builtin_fma(p, t, 0x1.745d1745d1746p-4);
  p = __builtin_fma(p, t, 0x1.c71c71c71c71cp-4);
  p = __builtin_fma(p, t, 0x1.2492492492492p-3);
  p = __builtin_fma(p, t, 0x1.999999999999ap-3);
  p = __builtin_fma(p, t, 0x1.5555555555555p-2);
  const double ln_m = 2.0 * __builtin_fma(s * t, p, s);
  const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
  const double de = (double)e;
  return __builtin_fma(de, ln2_hi, __builtin_fma(de, ln2_lo, ln_m));
}

// cos(2 pi u) for u in [0, 1].
HB_HD double det_cos2pi(double u) {
#pragma clang fp contract(off)
  double x = u > 0.5 ? 1.0 - u : u;  // exact; cos(2 pi u) = cos(2 pi (1 - u))
  double sign = 1.0;
  if (x > 0.25) {  // cos(pi - a) = -cos(a)
    x = 0.5 - x;
    sign = -1.0;
  }
  const bool use_sin = x > 0.125;  // cos(pi/2 - a) = sin(a)
  if (use_sin) x = 0.25 - x;
  const double two_pi_hi = 0x1.921fb54442d18p+2, two_pi_lo = 0x1.1a62633145c07p-52;
  const double a = __builtin_fma(x, two_pi_hi, x * two_pi_lo);  // in [0, pi/4]
  const double t = a * a;
  double r;
  if (use_sin) {
    double p = 0x1.71b8ef6dcf572p-66;
    p = __builtin_fma(p, t, -0x1.2f49b46814157p-57);
    p = __builtin_fma(p, t, 0x1.952c77030ad4ap-49);
    p = __builtin_fma(p, t, -0x1.ae7f3e733b81fp-41);
    p = __builtin_fma(p, t, 0x1.6124613a86d09p-33);
    p = __builtin_fma(p, t, -0x1.ae64567f544e4p-26);
    p = __builtin_fma(p, t, 0x1.71de3a556c734p-19);
    p = __builtin_fma(p, t, -0x1.a01a01a01a01ap-13);
    p = __builtin_fma(p, t, 0x1.1111111111111p-7);
    p = __builtin_fma(p, t, -0x1.5555555555555p-3);
    r = __builtin_fma(a * t, p, a);
  } else {
    double p = 0x1.e542ba4020225p-62;
    p = __builtin_fma(p, t, -0x1.6827863b97d97p-53);
    p = __builtin_fma(p, t, 0x1.ae7f3e733b81fp-45);
    p = __builtin_fma(p, t, -0x1.93974a8c07c9dp-37);
    p = __builtin_fma(p, t, 0x1.1eed8eff8d898p-29);
    p = __builtin_fma(p, t, -0x1.27e4fb7789f5cp-22);
    p = __builtin_fma(p, t, 0x1.a01a01a01a01ap-16);
    p = __builtin_fma(p, t, -0x1.6c16c16c16c17p-10);
    p = __builtin_fma(p, t, 0x1.5555555555555p-5);
    p = __builtin_fma(p, t, -0x1.0000000000000p-1);
    r = __builtin_fma(t, p, 1.0);
  }
  return sign * r;
}

// Square root of v >= 0 (normal or zero).
HB_HD double det_sqrt(double v) {
#pragma clang fp contract(off)
  if (v == 0.0) return 0.0;
  double r = bits_to_double(0x5FE6EB50C7B537A9ull - (double_to_bits(v) >> 1));  // ~ v^-1/2
  const double hv = v * 0.5;
  for (int i = 0; i < 5; i++) r = r * __builtin_fma(-hv * r, r, 1.5);
  double y = v * r;
  y = __builtin_fma(__builtin_fma(-y, y, v), r * 0.5, y);
  return y;
}

// Rounded N(0, sigma) from w[0..11] (three attempts of two 64-bit uniforms), cut at
// 6 sigma, as a canonical word mod q.
HB_HD uint64_t gaussian_mod(const uint32_t w[12], double sigma, uint64_t q) {
#pragma clang fp contract(off)
  const double cut = 6.0 * sigma;
  for (int k = 0; k < 3; k++) {
    const double u1 = unit_from(w[4 * k], w[4 * k + 1]);
    const double u2 = unit_from(w[4 * k + 2], w[4 * k + 3]);
    const double z = det_sqrt(-2.0 * det_log(u1)) * det_cos2pi(u2) * sigma;
    if (z < cut && z > -cut) {
      const double rz = __builtin_round(z);
      return rz < 0.0 ? q - (uint64_t)(-rz) : (uint64_t)rz;
    }
  }
  return 0;
}

// One coefficient of a key RLWE sample: stream (kind, row), counter = position.
HB_HD void key_coefficient(const uint32_t key[8], uint32_t kind, uint32_t row, uint32_t counter,
                           uint64_t q, double sigma, uint64_t* a, uint64_t* e) {
  uint32_t w[16];
  chacha20_block(key, counter, kind, row, kKeyStreamTag, w);
  *a = uniform_mod(w, q);
  *e = gaussian_mod(w + 4, sigma, q);
}

}  // namespace hb
