// Field arithmetic shared by the HIP kernels and the host-side key generation.
//
// Goldilocks (P = 2^64 - 2^32 + 1) keeps the reference's raw-word semantics where
// they are observable: words live in [0, 2^64) and are canonicalised only when a
// value is read (algebra/src/goldilocks/mod.rs:26-59). add/sub follow
// algebra/src/modulus/goldilocks/ops.rs:27-55 exactly, and neg_raw reproduces the
// wrapping `P - x` of ops.rs:69-71, which for a raw word x > P yields -x + 2^32 - 1.
//
// BabyBear (P = 2^31 - 2^27 + 1) is carried in plain canonical form; products use
// Montgomery reduction against constants that are stored pre-multiplied by
// R = 2^32, so mont_mul(x, y*R) = x*y mod P and data never leaves plain form.
// The reduction is the one of algebra/src/modulus/baby_bear/mod.rs:108-122
// (MU = P^-1 mod 2^32).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HB_HD __host__ __device__ __forceinline__
#else
#define HB_HD inline
#endif

namespace hb {

using u32 = uint32_t;
using u64 = uint64_t;

// ------------------------------- Goldilocks -------------------------------
struct Gl {
  using word = u64;
  static constexpr u64 P = 0xFFFFFFFF00000001ull;
  static constexpr u64 EPS = 0xFFFFFFFFull;
  static constexpr u32 MODULUS_BITS = 64;
  static constexpr u64 GENERATOR = 7;
  static constexpr bool IS_GOLDILOCKS = true;

  HB_HD static u64 canon(u64 x) { return x >= P ? x - P : x; }

  // ops.rs:27-34, bit-exact including the raw output word.
  HB_HD static u64 add(u64 x, u64 y) {
    u64 s = x + y;
    u64 adj = (s < x) ? EPS : 0;
    u64 s2 = s + adj;
    if (s2 < s) s2 += EPS;
    return s2;
  }
  // add for a canonical y: x + y < 2^64 + P, so the EPS of a carry cannot carry again and the
  // second correction of add never applies (the same word, one compare fewer).
  HB_HD static u64 add_cy(u64 x, u64 y) {
    const u64 s = x + y;
    return s + (s < x ? EPS : 0);
  }
  // ops.rs:48-55, bit-exact.
  HB_HD static u64 sub(u64 x, u64 y) {
    u64 d = x - y;
    u64 adj = (x < y) ? EPS : 0;
    u64 d2 = d - adj;
    if (d < adj) d2 -= EPS;
    return d2;
  }
  // ops.rs:69-71 with wrapping arithmetic.
  HB_HD static u64 neg_raw(u64 x) { return x == 0 ? 0 : P - x; }

  // Canonical a - b for canonical inputs.
  HB_HD static u64 sub_canon(u64 a, u64 b) {
    u64 d = a - b;
    return a < b ? d + P : d;
  }

  // 64x64 -> 128 product as (lo, hi).
  HB_HD static void mul_wide(u64 a, u64 b, u64& lo, u64& hi) {
#if defined(__HIP_DEVICE_COMPILE__)
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 p01 = (u64)a0 * b1 + (p00 >> 32);
    u64 p10 = (u64)a1 * b0 + (u32)p01;
    u64 p11 = (u64)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (u32)p00;
    hi = p11;
#else
    unsigned __int128 p = (unsigned __int128)a * b;
    lo = (u64)p;
    hi = (u64)(p >> 64);
#endif
  }
  // reduce128 (algebra/src/modulus/goldilocks/mod.rs:139-150): result in [0, 2^64),
  // congruent mod P.
  HB_HD static u64 reduce128(u64 lo, u64 hi) {
    u64 hh = hi >> 32;
    u64 hl = hi & EPS;
    u64 t0 = lo - hh;
    if (lo < hh) t0 -= EPS;
    u64 t1 = (hl << 32) - hl;
    u64 r = t0 + t1;
    if (r < t0) r += EPS;
    return r;
  }
  HB_HD static u64 mul(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return reduce128(lo, hi);
  }
  // Montgomery reduction with R = 2^64: for x = lo + hi * 2^64 with hi < P returns
  // x * 2^-64 mod P in [0, P). m = lo * P^-1 mod 2^64 = lo * (1 + 2^32); b = m*P >> 64
  // is formed from m without a multiply; r = hi - b, plus P on borrow.
  HB_HD static u64 mont_reduce(u64 lo, u64 hi) {
    u64 a = lo + (lo << 32);
    u64 e = a < lo ? 1 : 0;
    u64 b = a - (a >> 32) - e;
    u64 r = hi - b;
    return hi < b ? r - EPS : r;
  }
  // a * c * 2^-64 mod P: device constants (twiddles, keys) are stored as c * 2^64 mod P,
  // so mont_mul(x, c * 2^64) = x * c with x in plain form. Canonical output.
  HB_HD static u64 mont_mul(u64 a, u64 c) {
    u64 lo, hi;
    mul_wide(a, c, lo, hi);
    return mont_reduce(lo, hi);
  }
  // Host: plain canonical value -> Montgomery constant form (times 2^64 = EPS mod P).
  HB_HD static u64 to_mont(u64 v) { return canon(mul(v, EPS)); }
};

// -------------------------------- BabyBear --------------------------------
struct Bb {
  using word = u32;
  static constexpr u32 P = 0x78000001u;
  static constexpr u32 MU = 0x88000001u;  // P^-1 mod 2^32
  static constexpr u32 MU_NEG = 0x77FFFFFFu;  // -P^-1 mod 2^32
  static constexpr u32 R_MOD = 0x0FFFFFFEu;   // 2^32 mod P = 2^28 - 2
  static constexpr u32 MODULUS_BITS = 31;
  static constexpr u64 GENERATOR = 31;
  static constexpr bool IS_GOLDILOCKS = false;

  HB_HD static u32 canon(u32 x) { return x; }
  // Canonical inputs (< P < 2^31): the wrapped value of the wrong branch is >= 2^32 - P,
  // above every canonical word, so an unsigned min picks the result (no compare and
  // select, so no carry-flag dependence on the VALU).
  HB_HD static u32 umin(u32 a, u32 b) { return a < b ? a : b; }
  HB_HD static u32 add(u32 a, u32 b) {
    const u32 s = a + b;
    return umin(s, s - P);
  }
  HB_HD static u32 add_cy(u32 a, u32 b) { return add(a, b); }
  HB_HD static u32 sub(u32 a, u32 b) {
    const u32 d = a - b;
    return umin(d, d + P);
  }
  HB_HD static u32 neg_raw(u32 x) { return x == 0 ? 0 : P - x; }
  HB_HD static u32 sub_canon(u32 a, u32 b) { return sub(a, b); }
  // x * y * R^-1 mod P for x*y < P * 2^32, canonical. t * P matches x in the low 32 bits,
  // so (x - t P) / 2^32 = hi(x) - hi(t P) exactly, in (-P, P).
  HB_HD static u32 mont_reduce(u64 x) {
    const u32 t = (u32)x * MU;
    const u32 r = (u32)(x >> 32) - (u32)(((u64)t * P) >> 32);
    return umin(r, r + P);
  }
  HB_HD static u32 mont_mul(u32 a, u32 b) { return mont_reduce((u64)a * b); }
  // Host helpers (plain modular arithmetic).
  HB_HD static u32 mul_plain(u32 a, u32 b) { return (u32)(((u64)a * b) % P); }
  HB_HD static u32 to_mont(u32 a) { return (u32)(((u64)a << 32) % P); }
};

}  // namespace hb
