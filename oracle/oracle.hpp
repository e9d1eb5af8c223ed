// CPU oracle for TFHE NAND gate bootstrapping — TEST INFRASTRUCTURE ONLY.
//
// This is a source-faithful C++ restatement of the reference Rust `tfhe` gate
// bootstrapping path. It exists to check the HIP product bit for bit and to act
// as the timed CPU baseline. Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load it; the product never does.
//
// Every function cites the reference file:line it follows. Raw-word semantics of
// the Goldilocks field (values in [0, 2^64), canonicalised only by value()) are
// reproduced exactly, including the wrapping negation of a raw word above P.
//
// NTT convention (the reference delegates to the un-vendored `concrete-ntt`
// crate, whose root choice and slot order are unpinned): forward output slot i
// holds X(psi^(2*brv(i)+1)), i.e. evaluation order bit-reversed, with
// psi = g^((P-1)/(2N)), g = 7 (Goldilocks) or g = 31 (BabyBear). Forward and
// inverse outputs are canonical, in [0, P) — the range the reference documents
// for its NTT wrapper (algebra/src/transformation/mod.rs:77-95).
//
// Parity pins: field ops against u128 closed forms; NTT against the O(N^2)
// evaluation formula and schoolbook negacyclic products; decomposition by
// recomposition; external product by schoolbook; blind rotation + key switch
// semantically (decrypts to NAND with noise < Q/16) and against an independent
// brute-force path that uses schoolbook products in place of NTTs.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

using u32 = uint32_t;
using u64 = uint64_t;
using u128 = unsigned __int128;

// ---------------------------------------------------------------------------
// Goldilocks, P = 2^64 - 2^32 + 1, raw u64 words.
// algebra/src/modulus/goldilocks/mod.rs:9-12 (EPSILON, P),
// algebra/src/modulus/goldilocks/ops.rs:23-95 (add/sub/neg/mul),
// algebra/src/goldilocks/mod.rs:51-59 (new/value), :114-127 (decomposition).
// ---------------------------------------------------------------------------
struct Gl {
  using word = u64;
  static constexpr u64 P = 0xFFFFFFFF00000001ull;
  static constexpr u64 EPS = 0xFFFFFFFFull;
  static constexpr u32 MODULUS_BITS = 64;  // 64 - P.leading_zeros()
  static constexpr u64 GENERATOR = 7;

  // The conditional corrections are written as masks (0 - flag) & EPS so that the
  // compiler emits no data-dependent branches (a carry is a coin flip for random words).
  static inline u64 eps_if(bool c) { return (0 - (u64)c) & EPS; }
  // ops.rs:27-34
  static inline u64 add(u64 x, u64 y) {
    u64 sum = x + y;
    u64 sum2 = sum + eps_if(sum < x);
    return sum2 + eps_if(sum2 < sum);
  }
  // ops.rs:48-55
  static inline u64 sub(u64 x, u64 y) {
    u64 diff = x - y;
    u64 adj = eps_if(x < y);
    u64 diff2 = diff - adj;
    return diff2 - eps_if(diff < adj);
  }
  // ops.rs:69-71 — wraps in release builds for a raw x > P. Such events are counted
  // (per thread) so tests can prove a workload exercises them.
  static inline u64& raw_neg_events() {
    static thread_local u64 count = 0;
    return count;
  }
  static inline u64 neg(u64 x) {
    if (x > P) raw_neg_events()++;
    return x == 0 ? 0 : P - x;
  }
  // mod.rs:139-150 (reduce128) with add_no_canonicalize_trashing_input :91-125
  static inline u64 reduce128(u128 x) {
    u64 x_lo = (u64)x, x_hi = (u64)(x >> 64);
    u64 x_hi_hi = x_hi >> 32;
    u64 x_hi_lo = x_hi & EPS;
    u64 t0 = x_lo - x_hi_hi - eps_if(x_lo < x_hi_hi);
    u64 t1 = x_hi_lo * EPS;
    u64 res = t0 + t1;
    return res + eps_if(res < t0);
  }
  // ops.rs:85-87
  static inline u64 mul(u64 x, u64 y) { return reduce128((u128)x * y); }
  // mod.rs:26-33 (to_canonical_u64)
  static inline u64 value(u64 x) { return x >= P ? x - P : x; }
  // goldilocks/mod.rs:52-54 — Field::new stores the word as given.
  static inline u64 from_value(u64 v) { return v; }
  static inline bool is_zero_raw(u64 x) { return x == 0; }  // goldilocks/mod.rs:313-316
  static inline u64 canon(u64 x) { return value(x); }
};

// ---------------------------------------------------------------------------
// BabyBear, P = 2^31 - 2^27 + 1, Montgomery u32 words (R = 2^32, MU = P^-1).
// algebra/src/modulus/baby_bear/mod.rs:11-122, ops.rs:9-76,
// algebra/src/baby_bear/mod.rs:28-48, :114-127.
// ---------------------------------------------------------------------------
struct Bb {
  using word = u32;
  static constexpr u32 P = 0x78000001u;
  static constexpr u32 MU = 0x88000001u;
  static constexpr u32 MODULUS_BITS = 31;  // 32 - P.leading_zeros()
  static constexpr u64 GENERATOR = 31;

  static inline u32 to_monty(u32 x) { return (u32)(((u64)x << 32) % P); }  // mod.rs:70-72
  // mod.rs:108-122
  static inline u32 monty_reduce(u64 x) {
    u64 t = (x * (u64)MU) & 0xFFFFFFFFull;
    u64 u = t * (u64)P;
    u64 x_sub_u = x - u;
    bool over = x < u;
    u32 hi = (u32)(x_sub_u >> 32);
    u32 corr = over ? P : 0;
    return hi + corr;
  }
  static inline u32 from_monty(u32 x) { return monty_reduce((u64)x); }
  // ops.rs:9-16
  static inline u32 add(u32 x, u32 y) {
    u32 sum = x + y;
    u32 corr = sum - P;
    bool over = sum < P;
    return over ? sum : corr;
  }
  // ops.rs:30-35
  static inline u32 sub(u32 x, u32 y) {
    u32 diff = x - y;
    bool over = x < y;
    return diff + (over ? P : 0);
  }
  static inline u32 neg(u32 x) { return sub(0, x); }  // ops.rs:49-51
  static inline u32 mul(u32 x, u32 y) { return monty_reduce((u64)x * y); }  // ops.rs:65-68
  static inline u64 value(u32 x) { return from_monty(x); }
  static inline u32 from_value(u64 v) { return to_monty((u32)v); }
  static inline bool is_zero_raw(u32 x) { return x == 0; }
  static inline u32 canon(u32 x) { return x; }  // Montgomery words are always in [0, P)
};

// ---------------------------------------------------------------------------
// Parameter sets: tfhe/src/bfhe/parameters.rs:45-99, vfhe/src/bfhe/parameters.rs:26-84.
// ---------------------------------------------------------------------------
enum SetId { BABYBEAR = 0, GOLDILOCKS = 1, ZAMA = 2 };

struct Params {
  int set;
  int is_goldilocks;
  u32 n;        // lwe_dimension
  u32 N;        // ring_dimension
  u32 br_bits;  // blind_rotation_basis_bits
  u32 ks_bits;  // key_switching_basis_bits
  u32 br_len;   // decompose_len for blind rotation
  u32 ks_len;   // decompose_len for key switching
  u64 q;        // modulus value
  u32 t;        // lwe_plain_modulus
};

bool get_params(int set, Params* out);

}  // namespace orc
