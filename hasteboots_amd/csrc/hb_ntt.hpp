// Wave-level negacyclic NTT for gfx950 (wave64), shared by the bootstrap and
// microbench kernels.
//
// One wavefront owns one polynomial of N = 64*E words: lane l holds E words in
// registers. The transform is the Cooley-Tukey / Gentleman-Sande pair whose
// forward output slot i is X(psi^(2*bitrev(i)+1)) (bit-reversed evaluation order),
// split into three register passes joined by two exchanges through a padded LDS
// buffer of N + N/16 words:
//   pass 1: log2(E) stages; lane l holds positions l + 64e (twiddles are uniform
//           across the wave and compile-time constants);
//   pass 2: 4 stages inside each 64-word block; lane holds positions
//           64b + j0 + 4k of block b = u>>2, j0 = u&3 for u = l + 64v;
//   pass 3: 2 stages inside each 4-word block; lane holds positions 4(l+64q)+r,
//           i.e. 4 consecutive words, so global loads/stores of the NTT-domain
//           layout are 32-byte (Goldilocks) or 16-byte (BabyBear) vectors per lane.
// The inverse runs the passes in reverse and takes/gives the same layouts, so the
// NTT-domain data never needs a permutation between kernels and keys.
//
// The reference delegates these transforms to concrete-ntt
// (algebra/src/transformation/concrete.rs:51-61, 138-148).
#pragma once
#include <utility>

#include "hb_field.hpp"

namespace hb {

// The wavefront's index in its one-dimensional workgroup as a wave-uniform (scalar) value.
// threadIdx.x >> 6 is the same in the 64 lanes of a wave, but the compiler takes it for a
// per-lane value, and every address derived from it (a gate's scratch, accumulator and
// outputs) would then sit in VGPR pairs.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Intra-wave LDS ordering: all lanes of one wavefront execute LDS operations in
// order; the fences stop the compiler from moving accesses across the exchange.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class F>
struct DF;

template <>
struct DF<Gl> {
  using W = u64;
  __device__ __forceinline__ static W add(W a, W b) { return Gl::add(a, b); }
  __device__ __forceinline__ static W sub(W a, W b) { return Gl::sub(a, b); }
  // Constants (twiddles, keys) are stored times 2^64 (Montgomery constant form).
  __device__ __forceinline__ static W mul(W a, W c) { return Gl::mont_mul(a, c); }
  __device__ __forceinline__ static W canon(W a) { return Gl::canon(a); }
};

template <>
struct DF<Bb> {
  using W = u32;
  __device__ __forceinline__ static W add(W a, W b) { return Bb::add(a, b); }
  __device__ __forceinline__ static W sub(W a, W b) { return Bb::sub(a, b); }
  // Constants (twiddles, keys) are stored times R = 2^32.
  // Montgomery reduction of x < 2^32 P, canonical x / 2^32 mod P: with t = -x / P mod 2^32,
  // x + t P (< 2^33 P < 2^64, one v_mad_u64_u32 with x as the addend) has a zero low word and
  // its high word is in [0, 2P). t's product is pinned to full 64 bits (the empty asm makes
  // it live) so it compiles to v_mad_u64_u32 rather than v_mul_lo_u32, which issues slower
  // on gfx950. Three multiplies and a conditional subtract: 5 instructions.
  __device__ __forceinline__ static W redc(u64 x) {
    u64 t = (u64)(u32)x * Bb::MU_NEG;
    asm("" : "+v"(t));
    const u32 r = (u32)(((u64)(u32)t * Bb::P + x) >> 32);
    return Bb::umin(r, r - Bb::P);
  }
  // The product is pinned too: otherwise the compiler recomputes its low word with a
  // separate v_mul_lo_u32 and adds the full product in last (four multiplies, not three).
  __device__ __forceinline__ static W mul(W a, W c) {
    u64 x = (u64)a * c;
    asm("" : "+v"(x));
    return redc(x);
  }
  // Lazy sums of products: s = hi 2^32 + lo = hi R_MOD + lo (mod P), below 2^60 for any s.
  __device__ __forceinline__ static u64 fold(u64 s) {
    return (u64)(u32)(s >> 32) * Bb::R_MOD + (u32)s;
  }
  // A canonical accumulator a entering a lazy sum in Montgomery scale (a 2^32 mod P), < 2^59.
  __device__ __forceinline__ static u64 lazy_in(W a) { return (u64)a * Bb::R_MOD; }
  __device__ __forceinline__ static W lazy_out(u64 s) { return redc(fold(s)); }
  // Words in [0, 2P) (forward-transform outputs) to canonical.
  __device__ __forceinline__ static W canon(W a) { return Bb::umin(a, a - Bb::P); }
};

template <class F>
__device__ __forceinline__ void ct_bfly(typename DF<F>::W& x, typename DF<F>::W& y,
                                        typename DF<F>::W w) {
  auto t = DF<F>::mul(y, w);
  y = DF<F>::sub(x, t);
  x = DF<F>::add(x, t);
}

template <class F>
__device__ __forceinline__ void gs_bfly(typename DF<F>::W& x, typename DF<F>::W& y,
                                        typename DF<F>::W w) {
  auto s = DF<F>::add(x, y);
  y = DF<F>::mul(DF<F>::sub(x, y), w);
  x = s;
}

// Pairs of independent butterflies. The generic form is two scalar butterflies;
// Goldilocks uses the hand-scheduled gfx950 sequences of hb_gl_asm.inc, which
// interleave the two carry chains (t = y*w is a Montgomery product, canonical; CT
// outputs lie in [0, 2^64); GS takes and returns canonical words). `_u` variants take
// wave-uniform twiddles (SGPRs), `_v` per-lane twiddles.
template <class F>
struct Bfly {
  using W = typename DF<F>::W;
  __device__ __forceinline__ static void ct2_u(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    ct_bfly<F>(xa, ya, wa);
    ct_bfly<F>(xb, yb, wb);
  }
  __device__ __forceinline__ static void ct2_v(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    ct_bfly<F>(xa, ya, wa);
    ct_bfly<F>(xb, yb, wb);
  }
  __device__ __forceinline__ static void gs2_u(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    gs_bfly<F>(xa, ya, wa);
    gs_bfly<F>(xb, yb, wb);
  }
  __device__ __forceinline__ static void gs2_v(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    gs_bfly<F>(xa, ya, wa);
    gs_bfly<F>(xb, yb, wb);
  }
  // Inverse-transform input normalisation (identity: inputs are canonical).
  __device__ __forceinline__ static W inv_in(W x) { return x; }
};

// BabyBear CT butterflies on lazy words in [0, 2P) (2P < 2^32): x is reduced on entry, the
// product is canonical, and the outputs x + t and x - t + P stay in [0, 2P), one instruction
// fewer than canonical outputs. A forward transform therefore returns words in [0, 2P)
// (DF<Bb>::canon reduces them); GS butterflies keep canonical words.
template <>
struct Bfly<Bb> {
  using W = u32;
  __device__ __forceinline__ static void ct(W& x, W& y, W w) {
    const W xc = Bb::umin(x, x - Bb::P);
    const W t = DF<Bb>::mul(y, w);
    y = xc + Bb::P - t;
    x = xc + t;
  }
  __device__ __forceinline__ static void ct2_u(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    ct(xa, ya, wa);
    ct(xb, yb, wb);
  }
  __device__ __forceinline__ static void ct2_v(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    ct(xa, ya, wa);
    ct(xb, yb, wb);
  }
  __device__ __forceinline__ static void gs2_u(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    gs_bfly<Bb>(xa, ya, wa);
    gs_bfly<Bb>(xb, yb, wb);
  }
  __device__ __forceinline__ static void gs2_v(W& xa, W& ya, W wa, W& xb, W& yb, W wb) {
    gs_bfly<Bb>(xa, ya, wa);
    gs_bfly<Bb>(xb, yb, wb);
  }
  __device__ __forceinline__ static W inv_in(W x) { return x; }
};

#if defined(__HIP_DEVICE_COMPILE__)
#include "hb_gl_asm.inc"

template <>
struct Bfly<Gl> {
  using W = u64;
  __device__ __forceinline__ static void ct2_u(u64& xa, u64& ya, u64 wa, u64& xb, u64& yb, u64 wb) {
    u64 Xa, Ya, Xb, Yb, dummy, s1a, s2a, s1b, s2b;
    HB_GL_CT2_S_ASM;
    xa = Xa, ya = Ya, xb = Xb, yb = Yb;
  }
  __device__ __forceinline__ static void ct2_v(u64& xa, u64& ya, u64 wa, u64& xb, u64& yb, u64 wb) {
    u64 Xa, Ya, Xb, Yb, dummy, s1a, s2a, s1b, s2b;
    HB_GL_CT2_V_ASM;
    xa = Xa, ya = Ya, xb = Xb, yb = Yb;
  }
  __device__ __forceinline__ static void gs2_u(u64& xa, u64& ya, u64 wa, u64& xb, u64& yb, u64 wb) {
    u32 Xalo, Xahi, Xblo, Xbhi;
    u64 Ya, Yb, dummy, s1a, s2a, s1b, s2b;
    HB_GL_GS2_S_ASM;
    xa = ((u64)Xahi << 32) | Xalo, ya = Ya, xb = ((u64)Xbhi << 32) | Xblo, yb = Yb;
  }
  __device__ __forceinline__ static void gs2_v(u64& xa, u64& ya, u64 wa, u64& xb, u64& yb, u64 wb) {
    u32 Xalo, Xahi, Xblo, Xbhi;
    u64 Ya, Yb, dummy, s1a, s2a, s1b, s2b;
    HB_GL_GS2_V_ASM;
    xa = ((u64)Xahi << 32) | Xalo, ya = Ya, xb = ((u64)Xbhi << 32) | Xblo, yb = Yb;
  }
  __device__ __forceinline__ static u64 inv_in(u64 x) { return Gl::canon(x); }
};

// (pa, pb) += (x * ka, x * kb) with Montgomery-form keys; pa, pb in [0, 2^64).
__device__ __forceinline__ void gl_mac2(u64& pa, u64& pb, u64 x, u64 ka, u64 kb) {
  u64 Pa, Pb, dummy, s1a, s2a, s1b, s2b;
  HB_GL_MAC2_ASM;
  pa = Pa, pb = Pb;
}

// (ra, rb) = (xa * wa, xb * wb) with Montgomery-form constants wa, wb (per lane); canonical.
__device__ __forceinline__ void gl_mul2(u64& ra, u64& rb, u64 xa, u64 wa, u64 xb, u64 wb) {
  u64 Ra, Rb, dummy, s1a, s2a, s1b, s2b;
  HB_GL_MUL2_ASM;
  ra = Ra, rb = Rb;
}

// Butterfly pairs whose twiddles are +-2^e (hb_gl_asm.inc power-of-two macros). TA/TB:
// multiplier type 0 (A), 1 (BL), 2 (BH) with shift immediates S*1, S*2; NA/NB: the
// twiddle is the negation of what the macro multiplies by. For CT a negated twiddle
// swaps the two outputs; for GS it swaps the inputs (the sum is symmetric, the
// difference changes sign).
#define HB_GL_P2_DISPATCH(KIND)                                              \
  if constexpr (TA == 0 && TB == 0) { HB_GL_##KIND##P2_A_A_ASM; }            \
  else if constexpr (TA == 0 && TB == 1) { HB_GL_##KIND##P2_A_BL_ASM; }      \
  else if constexpr (TA == 0 && TB == 2) { HB_GL_##KIND##P2_A_BH_ASM; }      \
  else if constexpr (TA == 1 && TB == 0) { HB_GL_##KIND##P2_BL_A_ASM; }      \
  else if constexpr (TA == 1 && TB == 1) { HB_GL_##KIND##P2_BL_BL_ASM; }     \
  else if constexpr (TA == 1 && TB == 2) { HB_GL_##KIND##P2_BL_BH_ASM; }     \
  else if constexpr (TA == 2 && TB == 0) { HB_GL_##KIND##P2_BH_A_ASM; }      \
  else if constexpr (TA == 2 && TB == 1) { HB_GL_##KIND##P2_BH_BL_ASM; }     \
  else { HB_GL_##KIND##P2_BH_BH_ASM; }

template <int TA, int SA1, int SA2, bool NA, int TB, int SB1, int SB2, bool NB>
__device__ __forceinline__ void gl_ct2_pow2(u64& xa_, u64& ya_, u64& xb_, u64& yb_) {
  const u64 xa = xa_, ya = ya_, xb = xb_, yb = yb_;
  u64 Xa, Ya, Xb, Yb, dummy, s1a, s2a, s1b, s2b;
  HB_GL_P2_DISPATCH(CT)
  xa_ = NA ? Ya : Xa, ya_ = NA ? Xa : Ya;
  xb_ = NB ? Yb : Xb, yb_ = NB ? Xb : Yb;
}

// CT pairs whose twiddles may also be 1: types 0-2 as above, 3 (TN: y any word) and 4 (TC: y
// canonical) for a unit twiddle; N*: the twiddle is the negation (outputs swapped).
#define HB_GL_CTX_CASE(A, B, NA_, NB_) \
  else if constexpr (TA == NA_ && TB == NB_) { HB_GL_CTP2_##A##_##B##_ASM; }
template <int TA, int SA1, int SA2, bool NA, int TB, int SB1, int SB2, bool NB>
__device__ __forceinline__ void gl_ct2_unit(u64& xa_, u64& ya_, u64& xb_, u64& yb_) {
  if constexpr (TA < 3 && TB < 3) {
    gl_ct2_pow2<TA, SA1, SA2, NA, TB, SB1, SB2, NB>(xa_, ya_, xb_, yb_);
  } else {
    const u64 xa = xa_, ya = ya_, xb = xb_, yb = yb_;
    u64 Xa, Ya, Xb, Yb, dummy, s1a, s2a, s1b, s2b;
    if constexpr (false) {}
    HB_GL_CTX_CASE(A, TN, 0, 3) HB_GL_CTX_CASE(A, TC, 0, 4)
    HB_GL_CTX_CASE(BL, TN, 1, 3) HB_GL_CTX_CASE(BL, TC, 1, 4)
    HB_GL_CTX_CASE(BH, TN, 2, 3) HB_GL_CTX_CASE(BH, TC, 2, 4)
    HB_GL_CTX_CASE(TN, A, 3, 0) HB_GL_CTX_CASE(TN, BL, 3, 1) HB_GL_CTX_CASE(TN, BH, 3, 2)
    HB_GL_CTX_CASE(TN, TN, 3, 3) HB_GL_CTX_CASE(TN, TC, 3, 4)
    HB_GL_CTX_CASE(TC, A, 4, 0) HB_GL_CTX_CASE(TC, BL, 4, 1) HB_GL_CTX_CASE(TC, BH, 4, 2)
    HB_GL_CTX_CASE(TC, TN, 4, 3) HB_GL_CTX_CASE(TC, TC, 4, 4)
    else static_assert(TA < 0, "unknown twiddle type pair");
    xa_ = NA ? Ya : Xa, ya_ = NA ? Xa : Ya;
    xb_ = NB ? Yb : Xb, yb_ = NB ? Xb : Yb;
  }
}
#undef HB_GL_CTX_CASE

template <int TA, int SA1, int SA2, bool NA, int TB, int SB1, int SB2, bool NB>
__device__ __forceinline__ void gl_gs2_pow2(u64& xa_, u64& ya_, u64& xb_, u64& yb_) {
  const u64 xa = NA ? ya_ : xa_, ya = NA ? xa_ : ya_;
  const u64 xb = NB ? yb_ : xb_, yb = NB ? xb_ : yb_;
  u32 Xalo, Xahi, Xblo, Xbhi;
  u64 Ya, Yb, dummy, s1a, s2a, s1b, s2b;
  HB_GL_P2_DISPATCH(GS)
  xa_ = ((u64)Xahi << 32) | Xalo, ya_ = Ya, xb_ = ((u64)Xbhi << 32) | Xblo, yb_ = Yb;
}
#undef HB_GL_P2_DISPATCH
#else
// Host compilation pass only parses device code; this body is never compiled for a GPU.
__device__ inline void gl_mac2(u64& pa, u64& pb, u64 x, u64 ka, u64 kb) {
  pa = Gl::add(pa, Gl::mont_mul(x, ka));
  pb = Gl::add(pb, Gl::mont_mul(x, kb));
}
__device__ inline void gl_mul2(u64& ra, u64& rb, u64 xa, u64 wa, u64 xb, u64 wb) {
  ra = Gl::mont_mul(xa, wa);
  rb = Gl::mont_mul(xb, wb);
}
#endif

// Pass-1 twiddles as compile-time constants. psi = g^((P-1)/2N) (g = 7 Goldilocks,
// 31 BabyBear) exactly as the host tables; entry k < E is psi^bitrev(k) (or psi^-bitrev(k))
// in the field's constant form (times 2^64 Goldilocks, 2^32 BabyBear). Pass 1 only
// touches indices below E, so these wave-uniform constants replace table loads there.
__host__ __device__ constexpr u64 ct_mulmod(u64 a, u64 b, u64 p) {
  return (u64)((unsigned __int128)a * b % p);
}
__host__ __device__ constexpr u64 ct_powmod(u64 b, u64 e, u64 p) {
  u64 r = 1;
  b %= p;
  while (e) {
    if (e & 1) r = ct_mulmod(r, b, p);
    b = ct_mulmod(b, b, p);
    e >>= 1;
  }
  return r;
}
__host__ __device__ constexpr u64 ct_bitrev(u64 x, int bits) {
  u64 r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
template <class F, int LOGN, bool INV>
struct Pass1Tw {
  using W = typename DF<F>::W;
  static constexpr int E = (1 << LOGN) / 64;
  W w[E];
  __host__ __device__ constexpr Pass1Tw() : w{} {
    const u64 p = F::P;
    const u64 psi = ct_powmod(F::GENERATOR, (p - 1) >> (LOGN + 1), p);
    const u64 base = INV ? ct_powmod(psi, p - 2, p) : psi;
    const u64 r = F::IS_GOLDILOCKS ? (u64)(((unsigned __int128)1 << 64) % p) : ((u64)1 << 32) % p;
    for (int k = 0; k < E; k++)
      w[k] = (W)ct_mulmod(ct_powmod(base, ct_bitrev(k, LOGN), p), r, p);
  }
};
template <int E>
struct Log2 {
  static constexpr int value = E == 16 ? 4 : E == 32 ? 5 : E == 8 ? 3 : -1;
};

// Index of the k-th butterfly's upper element in a stage whose partner distance is half.
__host__ __device__ constexpr int bf_lo(int k, int half) { return (k / half) * 2 * half + k % half; }

// Goldilocks pass-1 twiddles are powers of two: 2 has order 192 mod P, so the 2^LOGN-th
// roots psi^bitrev(i) for i < E are +-2^e. Classification for the shift-based
// multipliers of hb_gl_asm.inc (with 2^96 = -1):
//   e in [1, 32):  type A, y * 2^e by shift and fold;
//   e in [32, 96): y * 2^e = -(y * 2^(j - 64)), j = e - 32: types BL (j < 32) / BH (j > 32),
//                  the Montgomery reduction of y * 2^j.
// ok = 0 for anything else (the caller falls back to the general product).
struct Pow2Tw {
  int ok, neg, type, s1, s2;
};
template <int LOGN, bool INV>
__host__ __device__ constexpr Pow2Tw pow2_twiddle(int i) {
  const u64 p = Gl::P;
  const u64 psi = ct_powmod(Gl::GENERATOR, (p - 1) >> (LOGN + 1), p);
  const u64 base = INV ? ct_powmod(psi, p - 2, p) : psi;
  const u64 w = ct_powmod(base, ct_bitrev(i, LOGN), p);
  u64 v = 1;
  int e = -1;
  for (int k = 0; k < 192; k++) {
    if (v == w) {
      e = k;
      break;
    }
    v = ct_mulmod(v, 2, p);
  }
  if (e < 0) return {0, 0, 0, 0, 0};
  int neg = 0;
  if (e >= 96) neg = 1, e -= 96;
  if (e >= 1 && e < 32) return {1, neg, 0, e, 32 - e};
  if (e >= 32 && e < 96) {
    const int j = e - 32;
    if (j >= 1 && j < 32) return {1, !neg, 1, j, 32 - j};
    if (j >= 33 && j < 64) return {1, !neg, 2, j - 32, 64 - j};
  }
  return {0, 0, 0, 0, 0};
}

// Pass 1 at compile-time indices: butterfly pair P of stage S (twiddle indices known to
// the compiler), Goldilocks twiddles as shift-based multipliers where possible.
template <class F, int E, bool INV, int S, int P>
__device__ __forceinline__ void pass1_pair(typename DF<F>::W (&x)[E]) {
  constexpr int LOGE = Log2<E>::value;
  constexpr int m = 1 << S, half = (E / 2) >> S;
  constexpr int e0 = bf_lo(2 * P, half), e1 = bf_lo(2 * P + 1, half);
  constexpr int ia = m + e0 / (2 * half), ib = m + e1 / (2 * half);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (F::IS_GOLDILOCKS) {
    constexpr Pow2Tw a = pow2_twiddle<LOGE + 6, INV>(ia), b = pow2_twiddle<LOGE + 6, INV>(ib);
    if constexpr (a.ok && b.ok) {
      if constexpr (!INV)
        gl_ct2_pow2<a.type, a.s1, a.s2, (bool)a.neg, b.type, b.s1, b.s2, (bool)b.neg>(
            x[e0], x[e0 + half], x[e1], x[e1 + half]);
      else
        gl_gs2_pow2<a.type, a.s1, a.s2, (bool)a.neg, b.type, b.s1, b.s2, (bool)b.neg>(
            x[e0], x[e0 + half], x[e1], x[e1 + half]);
      return;
    }
  }
#endif
  constexpr Pass1Tw<F, LOGE + 6, INV> T{};
  if constexpr (!INV)
    Bfly<F>::ct2_u(x[e0], x[e0 + half], T.w[ia], x[e1], x[e1 + half], T.w[ib]);
  else
    Bfly<F>::gs2_u(x[e0], x[e0 + half], T.w[ia], x[e1], x[e1 + half], T.w[ib]);
}

template <class F, int E, bool INV, int S, int... Ps>
__device__ __forceinline__ void pass1_stage(typename DF<F>::W (&x)[E],
                                            std::integer_sequence<int, Ps...>) {
  (pass1_pair<F, E, INV, S, Ps>(x), ...);
}

// Forward: stages S0 .. LOGE-1; inverse: LOGE-1 .. 0.
template <class F, int E, bool INV, int S0, int... Is>
__device__ __forceinline__ void pass1_all(typename DF<F>::W (&x)[E],
                                          std::integer_sequence<int, Is...>) {
  constexpr int LOGE = Log2<E>::value;
  (pass1_stage<F, E, INV, INV ? LOGE - 1 - Is : S0 + Is>(
       x, std::make_integer_sequence<int, E / 4>{}),
   ...);
}

// Forward transform. In: x[e] = position lane + 64e. Out: x[4q+r] = slot 4(lane+64q)+r.
// tw = psi_rev (psi^bitrev(k), constants in the field's constant form); pass 1 uses the
// compile-time copy of its first E entries (Pass1Tw). twv is the same table for the
// per-lane reads of passes 2 and 3, normally a copy in LDS.
// LDS word offsets of the three register layouts (buffer of N + N/16 words, one pad of 4
// words per 64). With lane < 64 each padded index is a per-lane base plus a constant,
// so every exchange access is one base register and an immediate offset:
//   layout 1 (lane + 64e):               lane + 68e
//   layout 2 (64(u>>2) + (u&3) + 4k,      68(lane>>2) + (lane&3) + 1088v + 4k
//             u = lane + 64v)
//   layout 3 (4(lane + 64q) + r):         4lane + 4(lane>>4) + 272q + r
__device__ __forceinline__ int lds_base2(int lane) { return 68 * (lane >> 2) + (lane & 3); }
__device__ __forceinline__ int lds_base3(int lane) { return 4 * lane + 4 * (lane >> 4); }

// Keeps loads issued above it ahead of the arithmetic below it.
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }

// Forward transform. In: x[e] = position lane + 64e. Out: x[4q+r] = slot 4(lane+64q)+r,
// words in [0, 2^64) (Goldilocks) or [0, 2P) (BabyBear); DF<F>::canon reduces them.
// Pass 1 uses the compile-time twiddles (Pass1Tw); twv is psi_rev (psi^bitrev(k) in the
// field's constant form) for the per-lane reads of passes 2 and 3, normally a copy in
// LDS. Each stage's twiddles are requested one stage ahead of their use.
// With SMALL (Goldilocks only), every input word is below 2^16 (gadget digits): stage 0,
// whose twiddle is psi^(N/2) = 2^48, is then exact without reduction:
//   x + 2^48 y = (2^16 y) : x   and   x - 2^48 y + P = ~(2^16 y) : (x + 1)
// as (high 32 bits : low 32 bits); both are valid words in [0, 2^64).
// PF: request per-lane twiddles a stage ahead (for one wave per SIMD, where nothing else
// hides LDS latency); without it they are read just before use, which needs fewer
// registers (for two waves per SIMD, which hide each other's latency).
// HALF (E = 32): both exchanges run one 16-word group at a time through a buffer of
// 64*16 + 64 words (half of N + N/16): group v of layout 2 (positions [1024v, 1024v+1024))
// holds exactly layout 1's words e in [16v, 16v+16) and layout 3's q in [4v, 4v+4), so each
// group's exchange is self-contained. Less LDS per wave, more waves per CU.
template <class F, int E, bool SMALL = false, bool PF = true, bool HALF = false>
__device__ __forceinline__ void ntt_fwd_wave(typename DF<F>::W (&x)[E], typename DF<F>::W* lds,
                                             const typename DF<F>::W* twv, int lane) {
  using W = typename DF<F>::W;
  constexpr int LOGE = Log2<E>::value;
  constexpr int V = E / 16;
  static_assert(LOGE > 0, "E must be 8, 16 or 32");
  const int b1 = lane, b2 = lds_base2(lane), b3 = lds_base3(lane);
  constexpr bool SMALL0 = SMALL && F::IS_GOLDILOCKS;
  if constexpr (SMALL0) {
    constexpr Pass1Tw<F, LOGE + 6, false> T{};
    // Montgomery form of 2^48 is 2^112 = -2^16 mod P
    static_assert((u64)T.w[1] == Gl::P - (1ull << 16), "stage-0 twiddle is not 2^48");
#pragma unroll
    for (int e = 0; e < E / 2; e++) {
      const u32 a0 = (u32)x[e], t1 = (u32)x[e + E / 2] << 16;
      x[e] = ((u64)t1 << 32) | a0;
      x[e + E / 2] = ((u64)(~t1) << 32) | (a0 + 1u);
    }
  }
  // pass 1
  constexpr int S0 = SMALL0 ? 1 : 0;
  pass1_all<F, E, false, S0>(x, std::make_integer_sequence<int, LOGE - S0>{});
  static_assert(!HALF || V == 2, "half-buffer exchanges are for two 16-word groups");
  if constexpr (HALF) {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int e = 0; e < 16; e++) lds[b1 + 68 * e] = x[16 * v + e];
      wave_sync();
#pragma unroll
      for (int k = 0; k < 16; k++) x[16 * v + k] = lds[b2 + 4 * k];
      wave_sync();
    }
  } else {
#pragma unroll
    for (int e = 0; e < E; e++) lds[b1 + 68 * e] = x[e];
    wave_sync();
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) x[16 * v + k] = lds[b2 + 1088 * v + 4 * k];
    }
  }
  // pass 2: block b = (lane + 64v) >> 2, stage s uses twv[(E << s) + (b << s) + j], j < 2^s
  W t2[V][4][8];
  if constexpr (PF) t2[0][0][0] = twv[E + (lane >> 2)];
  wave_sync();
#pragma unroll
  for (int v = 0; v < V; v++) {
    const int b = (lane + 64 * v) >> 2;
#pragma unroll
    for (int s = 0; s < 4; s++) {
      if constexpr (!PF) {
#pragma unroll
        for (int j = 0; j < (1 << s); j++) t2[v][s][j] = twv[(E << s) + (b << s) + j];
      } else if (s < 3) {
#pragma unroll
        for (int j = 0; j < (2 << s); j++) t2[v][s + 1][j] = twv[(E << (s + 1)) + (b << (s + 1)) + j];
      } else if (v + 1 < V) {
        t2[v + 1][0][0] = twv[E + ((lane + 64 * (v + 1)) >> 2)];
      }
      if constexpr (PF) sched_fence();
      const int half = 8 >> s;
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const int k0 = bf_lo(2 * p, half), k1 = bf_lo(2 * p + 1, half);
        Bfly<F>::ct2_v(x[16 * v + k0], x[16 * v + k0 + half], t2[v][s][k0 / (2 * half)],
                       x[16 * v + k1], x[16 * v + k1 + half], t2[v][s][k1 / (2 * half)]);
      }
      if constexpr (PF) sched_fence();
    }
  }
  if constexpr (HALF) {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) lds[b2 + 4 * k] = x[16 * v + k];
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int r = 0; r < 4; r++) x[16 * v + 4 * q + r] = lds[b3 + 272 * q + r];
      }
      if (v + 1 < V) wave_sync();
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) lds[b2 + 1088 * v + 4 * k] = x[16 * v + k];
    }
    wave_sync();
#pragma unroll
    for (int q = 0; q < E / 4; q++) {
#pragma unroll
      for (int r = 0; r < 4; r++) x[4 * q + r] = lds[b3 + 272 * q + r];
    }
  }
  // pass 3 twiddles, requested behind the exchange reads (PF) or per group
  W t3[E / 4][3];
  auto load3 = [&](int q) {
    const int c = lane + 64 * q;
    t3[q][0] = twv[16 * E + c];
    t3[q][1] = twv[32 * E + 2 * c];
    t3[q][2] = twv[32 * E + 2 * c + 1];
  };
  if constexpr (PF) {
#pragma unroll
    for (int q = 0; q < E / 4; q++) load3(q);
  }
  wave_sync();
  // pass 3
#pragma unroll
  for (int q = 0; q < E / 4; q++) {
    if constexpr (!PF) load3(q);
    Bfly<F>::ct2_v(x[4 * q + 0], x[4 * q + 2], t3[q][0], x[4 * q + 1], x[4 * q + 3], t3[q][0]);
    Bfly<F>::ct2_v(x[4 * q + 0], x[4 * q + 1], t3[q][1], x[4 * q + 2], x[4 * q + 3], t3[q][2]);
  }
}

// ---------------------------------------------------------------------------------------
// Twisted forward transform of gadget digits (Goldilocks, N = 1024, every input word below
// 2^16): the same slot values as ntt_fwd_wave<Gl, 16, true> with fewer general products.
// Stages 0-3 (pass 1) split X^1024 + 1 into 16 blocks X^64 - Z_g, Z_g = psi^(2 bitrev10(16 +
// g)), block g holding positions [64g, 64g + 64). Substituting X = beta_g Y with beta_g =
// psi^bitrev5(16 + g) (so beta_g^32 = psi^bitrev10(16 + g)) turns block g into Y^64 - 1: its
// words are multiplied by beta_g^p (the twist, p = position in the block), and the six
// remaining Cooley-Tukey stages take the Y-domain twiddles psi^bitrev10(2^s + i) beta_g^-t
// (t = 2^(9-s)), which do not depend on g and are 64th roots of unity, i.e. +-2^e or 1.
// Pass 2 (stages 4-7) thus runs on shifts and additions at compile-time twiddles; pass 3
// (stages 8, 9) keeps general products with 48 per-lane twiddles. The leaves are the same
// evaluations in the same order. Table twx (Montgomery constant form, 1136 words): the
// twist factor beta_g^p of block g, position p at 68 g + p (blocks padded like the exchange
// buffer, so that a wave's reads spread over the LDS banks), [1088, 1104) the stage-8
// twiddle of group ip, [1104, 1136) the stage-9 twiddles of groups (ip, h) at
// 1104 + 2 ip + h, where ip = the group's index within its block (bits 5..2 of the
// position). (kTwxWords, hb_kernels.hpp.)

// Type of a Y-domain pass-2 twiddle for gl_ct2_unit: stage 4 + S, group IP within the block,
// W = psi^(bitrev10(2^(4+S) + IP) - 2^(5-S)); CANON: the butterfly's y is canonical.
struct UnitTw {
  int type, s1, s2, neg;
};
template <int S, int IP, bool CANON>
__host__ __device__ constexpr UnitTw twisted_pass2_twiddle() {
  const u64 p = Gl::P;
  const u64 psi = ct_powmod(Gl::GENERATOR, (p - 1) >> 11, p);
  const u64 x = (ct_bitrev((1u << (4 + S)) + IP, 10) + 2048 - (32u >> S)) % 2048;
  const u64 w = ct_powmod(psi, x, p);
  u64 v = 1;
  int e = -1;
  for (int k = 0; k < 192; k++) {
    if (v == w) {
      e = k;
      break;
    }
    v = ct_mulmod(v, 2, p);
  }
  int neg = 0;
  if (e >= 96) neg = 1, e -= 96;
  if (e == 0) return {CANON ? 4 : 3, 0, 0, neg};
  if (e >= 1 && e < 32) return {0, e, 32 - e, neg};
  if (e >= 32 && e < 96) {
    const int j = e - 32;
    if (j >= 1 && j < 32) return {1, j, 32 - j, !neg};
    if (j >= 33 && j < 64) return {2, j - 32, 64 - j, !neg};
  }
  return {-1, 0, 0, 0};  // no shift form (2^32, 2^64): not a pass-2 twiddle
}

template <int S, int P>
__device__ __forceinline__ void twisted_pass2_pair(u64 (&x)[16]) {
  constexpr int half = 8 >> S;
  constexpr int k0 = bf_lo(2 * P, half), k1 = bf_lo(2 * P + 1, half);
  // after the twist every word is canonical: stage 4 takes y canonical
  constexpr UnitTw a = twisted_pass2_twiddle<S, k0 / (2 * half), S == 0>();
  constexpr UnitTw b = twisted_pass2_twiddle<S, k1 / (2 * half), S == 0>();
  static_assert(a.type >= 0 && b.type >= 0, "pass-2 twiddle without a shift form");
#if defined(__HIP_DEVICE_COMPILE__)
  gl_ct2_unit<a.type, a.s1, a.s2, (bool)a.neg, b.type, b.s1, b.s2, (bool)b.neg>(
      x[k0], x[k0 + half], x[k1], x[k1 + half]);
#endif
}

template <int S, int... Ps>
__device__ __forceinline__ void twisted_pass2_stage(u64 (&x)[16], std::integer_sequence<int, Ps...>) {
  (twisted_pass2_pair<S, Ps>(x), ...);
}

// Stages 2 and 3 (pass 1, compile-time shifts), the exchange, the twist, pass 2 and pass 3
// of the twisted transform, from x after stages 0 and 1 (x[e] = position lane + 64e, any
// words). Out: x[4q + r] = slot 4(lane + 64q) + r, words in [0, 2^64). PF: issue the twist
// and pass-3 table reads ahead of the exchanges (one wave per SIMD) instead of just before
// use.
template <bool PF>
__device__ __forceinline__ void ntt_fwd_twisted_from2(u64 (&x)[16], u64* lds, const u64* twx,
                                                      int lane) {
  constexpr int E = 16;
  const int b1 = lane, b2 = lds_base2(lane), b3 = lds_base3(lane);
  pass1_all<Gl, E, false, 2>(x, std::make_integer_sequence<int, 2>{});
  // exchange to pass 2's layout: positions 64b + j0 + 4k, b = lane >> 2, j0 = lane & 3
#pragma unroll
  for (int e = 0; e < E; e++) lds[b1 + 68 * e] = x[e];
  const u64* tw1 = twx + 68 * (lane >> 2) + (lane & 3);
  u64 w[E];
  if constexpr (PF) {
#pragma unroll
    for (int k = 0; k < E; k++) w[k] = tw1[4 * k];
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < E; k++) x[k] = lds[b2 + 4 * k];
  // twist: x * beta^p, canonical
#pragma unroll
  for (int k = 0; k < E; k += 2) {
    if constexpr (!PF) w[k] = tw1[4 * k], w[k + 1] = tw1[4 * k + 4];
    gl_mul2(x[k], x[k + 1], x[k], w[k], x[k + 1], w[k + 1]);
  }
  twisted_pass2_stage<0>(x, std::make_integer_sequence<int, 4>{});
  twisted_pass2_stage<1>(x, std::make_integer_sequence<int, 4>{});
  twisted_pass2_stage<2>(x, std::make_integer_sequence<int, 4>{});
  twisted_pass2_stage<3>(x, std::make_integer_sequence<int, 4>{});
  // exchange to pass 3's layout: positions 4(lane + 64q) + r
#pragma unroll
  for (int k = 0; k < E; k++) lds[b2 + 4 * k] = x[k];
  const int ip = lane & 15;
  const u64 w8 = twx[1088 + ip], w9a = twx[1104 + 2 * ip], w9b = twx[1105 + 2 * ip];
  wave_sync();
#pragma unroll
  for (int q = 0; q < E / 4; q++) {
#pragma unroll
    for (int r = 0; r < 4; r++) x[4 * q + r] = lds[b3 + 272 * q + r];
  }
  wave_sync();
#pragma unroll
  for (int q = 0; q < E / 4; q++) {
    Bfly<Gl>::ct2_v(x[4 * q + 0], x[4 * q + 2], w8, x[4 * q + 1], x[4 * q + 3], w8);
    Bfly<Gl>::ct2_v(x[4 * q + 0], x[4 * q + 1], w9a, x[4 * q + 2], x[4 * q + 3], w9b);
  }
}

// In: x[e] = digit word of position lane + 64e (< 2^16). Out: x[4q + r] = slot 4(lane + 64q) + r,
// words in [0, 2^64). twx: the table above (normally an LDS copy).
template <bool PF>
__device__ __forceinline__ void ntt_fwd_digits(u64 (&x)[16], u64* lds, const u64* twx, int lane) {
  constexpr int E = 16;
  // stage 0 exactly (inputs below 2^16; twiddle 2^48), then stage 1 at compile-time shifts
#pragma unroll
  for (int e = 0; e < E / 2; e++) {
    const u32 a0 = (u32)x[e], t1 = (u32)x[e + E / 2] << 16;
    x[e] = ((u64)t1 << 32) | a0;
    x[e + E / 2] = ((u64)(~t1) << 32) | (a0 + 1u);
  }
  pass1_all<Gl, E, false, 1>(x, std::make_integer_sequence<int, 1>{});
  ntt_fwd_twisted_from2<PF>(x, lds, twx, lane);
}

#if defined(__HIP_DEVICE_COMPILE__)
// Stages 0 and 1 of the byte-digit groups G and G + 1 (HB_GL_R4BYTES2_ASM).
template <u32 G>
__device__ __forceinline__ void byte_stages01(u64 (&x)[16], const u32 (&w)[4]) {
  u64 Y0a, Y3a, Y0b, Y3b, dummy, s1a, s2a, s1b, s2b;
  u32 Y1alo, Y1ahi, Y2alo, Y2ahi, Y1blo, Y1bhi, Y2blo, Y2bhi;
  HB_GL_R4BYTES2_ASM;
  x[G] = Y0a, x[G + 4] = ((u64)Y1ahi << 32) | Y1alo;
  x[G + 8] = ((u64)Y2ahi << 32) | Y2alo, x[G + 12] = Y3a;
  x[G + 1] = Y0b, x[G + 5] = ((u64)Y1bhi << 32) | Y1blo;
  x[G + 9] = ((u64)Y2bhi << 32) | Y2blo, x[G + 13] = Y3b;
}

// ntt_fwd_digits for byte digits held packed: w[m] byte g = the digit of position
// lane + 64(4m + g) (store_digits8's layout). Stages 0 and 1 of each group of positions
// lane + 64(g + {0, 4, 8, 12}) are formed from byte-placed dwords (v_perm_b32 selector bytes
// 0-3 pick the second source's bytes, 0x0C a zero byte) inside HB_GL_R4BYTES2_ASM, with no
// per-digit extraction.
template <bool PF>
__device__ __forceinline__ void ntt_fwd_bytes(u64 (&x)[16], const u32 (&w)[4], u64* lds,
                                              const u64* twx, int lane) {
  byte_stages01<0>(x, w);
  byte_stages01<2>(x, w);
  ntt_fwd_twisted_from2<PF>(x, lds, twx, lane);
}
#else
// Host compilation pass only parses device code; this body is never compiled for a GPU.
template <bool PF>
__device__ inline void ntt_fwd_bytes(u64 (&x)[16], const u32 (&w)[4], u64* lds, const u64* twx,
                                     int lane) {
  for (int e = 0; e < 16; e++) x[e] = (w[e >> 2] >> (8 * (e & 3))) & 0xFFu;
  ntt_fwd_digits<PF>(x, lds, twx, lane);
}
#endif

// Inverse transform without the 1/N factor. In: slot layout of ntt_fwd_wave.
// Out: x[e] = position lane + 64e, canonical. itwv = psi^-bitrev(k) for passes 3 and 2
// (normally an LDS copy); pass 1 uses the compile-time twiddles.
template <class F, int E, bool PF = true, bool HALF = false>
__device__ __forceinline__ void ntt_inv_wave(typename DF<F>::W (&x)[E], typename DF<F>::W* lds,
                                             const typename DF<F>::W* itwv, int lane) {
  using W = typename DF<F>::W;
  constexpr int LOGE = Log2<E>::value;
  constexpr int V = E / 16;
  const int b1 = lane, b2 = lds_base2(lane), b3 = lds_base3(lane);
  W t3[E / 4][3];
  auto load3 = [&](int q) {
    const int c = lane + 64 * q;
    t3[q][0] = itwv[16 * E + c];
    t3[q][1] = itwv[32 * E + 2 * c];
    t3[q][2] = itwv[32 * E + 2 * c + 1];
  };
  if constexpr (PF) {
#pragma unroll
    for (int q = 0; q < E / 4; q++) load3(q);
    sched_fence();
  }
#pragma unroll
  for (int e = 0; e < E; e++) x[e] = Bfly<F>::inv_in(x[e]);
  // pass 3
#pragma unroll
  for (int q = 0; q < E / 4; q++) {
    if constexpr (!PF) load3(q);
    Bfly<F>::gs2_v(x[4 * q + 0], x[4 * q + 1], t3[q][1], x[4 * q + 2], x[4 * q + 3], t3[q][2]);
    Bfly<F>::gs2_v(x[4 * q + 0], x[4 * q + 2], t3[q][0], x[4 * q + 1], x[4 * q + 3], t3[q][0]);
  }
  static_assert(!HALF || V == 2, "half-buffer exchanges are for two 16-word groups");
  if constexpr (HALF) {  // group v: layout 3's q in [4v, 4v+4) -> layout 2's group v
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
#pragma unroll
        for (int r = 0; r < 4; r++) lds[b3 + 272 * q + r] = x[16 * v + 4 * q + r];
      }
      wave_sync();
#pragma unroll
      for (int k = 0; k < 16; k++) x[16 * v + k] = lds[b2 + 4 * k];
      if (v + 1 < V) wave_sync();
    }
  } else {
#pragma unroll
    for (int q = 0; q < E / 4; q++) {
#pragma unroll
      for (int r = 0; r < 4; r++) lds[b3 + 272 * q + r] = x[4 * q + r];
    }
    wave_sync();
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) x[16 * v + k] = lds[b2 + 1088 * v + 4 * k];
    }
  }
  // pass 2, stages 3..0; stage s uses itwv[(E << s) + (b << s) + j], j < 2^s
  W t2[V][4][8];
  if constexpr (PF) {
#pragma unroll
    for (int j = 0; j < 8; j++) t2[0][3][j] = itwv[(E << 3) + ((lane >> 2) << 3) + j];
  }
  wave_sync();
#pragma unroll
  for (int v = 0; v < V; v++) {
    const int b = (lane + 64 * v) >> 2;
#pragma unroll
    for (int s = 3; s >= 0; s--) {
      if constexpr (!PF) {
#pragma unroll
        for (int j = 0; j < (1 << s); j++) t2[v][s][j] = itwv[(E << s) + (b << s) + j];
      } else if (s > 0) {
#pragma unroll
        for (int j = 0; j < (1 << (s - 1)); j++)
          t2[v][s - 1][j] = itwv[(E << (s - 1)) + (b << (s - 1)) + j];
      } else if (v + 1 < V) {
        const int bn = (lane + 64 * (v + 1)) >> 2;
#pragma unroll
        for (int j = 0; j < 8; j++) t2[v + 1][3][j] = itwv[(E << 3) + (bn << 3) + j];
      }
      if constexpr (PF) sched_fence();
      const int half = 8 >> s;
#pragma unroll
      for (int p = 0; p < 4; p++) {
        const int k0 = bf_lo(2 * p, half), k1 = bf_lo(2 * p + 1, half);
        Bfly<F>::gs2_v(x[16 * v + k0], x[16 * v + k0 + half], t2[v][s][k0 / (2 * half)],
                       x[16 * v + k1], x[16 * v + k1 + half], t2[v][s][k1 / (2 * half)]);
      }
      if constexpr (PF) sched_fence();
    }
  }
  if constexpr (HALF) {  // group v -> layout 1's e in [16v, 16v+16)
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) lds[b2 + 4 * k] = x[16 * v + k];
      wave_sync();
#pragma unroll
      for (int e = 0; e < 16; e++) x[16 * v + e] = lds[b1 + 68 * e];
      wave_sync();
    }
  } else {
#pragma unroll
    for (int v = 0; v < V; v++) {
#pragma unroll
      for (int k = 0; k < 16; k++) lds[b2 + 1088 * v + 4 * k] = x[16 * v + k];
    }
    wave_sync();
#pragma unroll
    for (int e = 0; e < E; e++) x[e] = lds[b1 + 68 * e];
    wave_sync();
  }
  // pass 1
  pass1_all<F, E, true, 0>(x, std::make_integer_sequence<int, LOGE>{});
}

// Inverse transform (without 1/N) shared by four wavefronts, one quarter each, for the
// latency-mode kernels: the same Gentleman-Sande butterflies and twiddles as
// ntt_inv_wave (stage t pairs positions j, j + t with itwv[N/(2t) + j/(2t)]), with each
// lane holding 4 words instead of 16. Wave qw (0..3) takes slots [256 qw, 256 qw + 256)
// (x[k] = slot 4(lane + 64 qw) + k, as x[4 qw + k] of ntt_fwd_wave's layout) and runs
// stages t = 1..128 through its own LDS buffer `lds` (words [0, 316)):
//   t = 1, 2:     positions 4(lane + 64 qw) + k
//   t = 4, 8:     256 qw + 16(lane >> 2) + (lane & 3) + 4k
//   t = 16, 32:   256 qw + 64(lane >> 4) + (lane & 15) + 16k
//   t = 64, 128:  256 qw + lane + 64k
// leaving its quarter in `lds` (at coop_a3 of the quarter-local position). After a
// workgroup barrier ntt_inv_cross reads positions 64 qw + lane + 256k from the four
// quarters' buffers and runs t = 256, 512; out: x[k] = position 64 qw + lane + 256k,
// canonical. The layouts are padded so that the 64 lanes' words spread over the banks.
__device__ __forceinline__ int coop_a1(int p) { return p + (p >> 3); }
__device__ __forceinline__ int coop_a2(int p) { return p + 4 * (p >> 4); }
__device__ __forceinline__ int coop_a3(int p) { return p + 4 * (p >> 6); }

// Gentleman-Sande pair of the N = 1024 inverse transform at the compile-time twiddle indices
// IA, IB (psi^-bitrev(I)): Goldilocks shift forms where both have one (pow2_twiddle; every
// index below 32 has), else the general product with the table words itwv[IA], itwv[IB].
template <class F, int IA, int IB>
__device__ __forceinline__ void gs2_const(typename DF<F>::W& xa, typename DF<F>::W& ya,
                                          typename DF<F>::W& xb, typename DF<F>::W& yb,
                                          const typename DF<F>::W* itwv) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (F::IS_GOLDILOCKS) {
    constexpr Pow2Tw a = pow2_twiddle<10, true>(IA), b = pow2_twiddle<10, true>(IB);
    if constexpr (a.ok && b.ok) {
      gl_gs2_pow2<a.type, a.s1, a.s2, (bool)a.neg, b.type, b.s1, b.s2, (bool)b.neg>(xa, ya, xb,
                                                                                   yb);
      return;
    }
  }
#endif
  Bfly<F>::gs2_v(xa, ya, itwv[IA], xb, yb, itwv[IB]);
}

// Stages t = 64, 128 of quarter QW of the four-wave inverse transform.
template <class F, int QW>
__device__ __forceinline__ void inv_quarter_top(typename DF<F>::W (&x)[4],
                                                const typename DF<F>::W* itwv) {
  gs2_const<F, 8 + 2 * QW, 9 + 2 * QW>(x[0], x[1], x[2], x[3], itwv);
  gs2_const<F, 4 + QW, 4 + QW>(x[0], x[2], x[1], x[3], itwv);
}

template <class F>
__device__ __forceinline__ void ntt_inv_quarter(typename DF<F>::W (&x)[4], typename DF<F>::W* lds,
                                                const typename DF<F>::W* itwv, int lane, int qw) {
  using W = typename DF<F>::W;
#pragma unroll
  for (int k = 0; k < 4; k++) x[k] = Bfly<F>::inv_in(x[k]);
  {  // t = 1, 2
    const int c = lane + 64 * qw;
    const W wa = itwv[512 + 2 * c], wb = itwv[512 + 2 * c + 1], w2 = itwv[256 + c];
    Bfly<F>::gs2_v(x[0], x[1], wa, x[2], x[3], wb);
    Bfly<F>::gs2_v(x[0], x[2], w2, x[1], x[3], w2);
#pragma unroll
    for (int k = 0; k < 4; k++) lds[coop_a1(4 * lane + k)] = x[k];
  }
  wave_sync();
  {  // t = 4, 8
    const int g = 16 * qw + (lane >> 2), base = 16 * (lane >> 2) + (lane & 3);
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = lds[coop_a1(base + 4 * k)];
    const W wa = itwv[128 + 2 * g], wb = itwv[128 + 2 * g + 1], w8 = itwv[64 + g];
    Bfly<F>::gs2_v(x[0], x[1], wa, x[2], x[3], wb);
    Bfly<F>::gs2_v(x[0], x[2], w8, x[1], x[3], w8);
#pragma unroll
    for (int k = 0; k < 4; k++) lds[coop_a2(base + 4 * k)] = x[k];
  }
  wave_sync();
  {  // t = 16, 32
    const int b = 4 * qw + (lane >> 4), base = 64 * (lane >> 4) + (lane & 15);
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = lds[coop_a2(base + 16 * k)];
    const W wa = itwv[32 + 2 * b], wb = itwv[32 + 2 * b + 1], w32 = itwv[16 + b];
    Bfly<F>::gs2_v(x[0], x[1], wa, x[2], x[3], wb);
    Bfly<F>::gs2_v(x[0], x[2], w32, x[1], x[3], w32);
#pragma unroll
    for (int k = 0; k < 4; k++) lds[coop_a3(base + 16 * k)] = x[k];
  }
  wave_sync();
  {  // t = 64, 128
#pragma unroll
    for (int k = 0; k < 4; k++) x[k] = lds[coop_a3(lane + 64 * k)];
    if constexpr (F::IS_GOLDILOCKS) {
      // twiddle indices 8 + 2 qw, 9 + 2 qw and 4 + qw: shifts, per wave-uniform qw
      switch (qw) {
        case 0: inv_quarter_top<F, 0>(x, itwv); break;
        case 1: inv_quarter_top<F, 1>(x, itwv); break;
        case 2: inv_quarter_top<F, 2>(x, itwv); break;
        default: inv_quarter_top<F, 3>(x, itwv); break;
      }
    } else {
      const W wa = itwv[8 + 2 * qw], wb = itwv[8 + 2 * qw + 1], w128 = itwv[4 + qw];
      Bfly<F>::gs2_v(x[0], x[1], wa, x[2], x[3], wb);
      Bfly<F>::gs2_v(x[0], x[2], w128, x[1], x[3], w128);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) lds[coop_a3(lane + 64 * k)] = x[k];
  }
}

// Stages t = 256, 512 of the four-wave inverse transform; q0: the LDS buffer of the wave
// holding quarter 0, quarter k's buffer at q0 + k * stride (in words). Call after a
// barrier that follows every quarter's ntt_inv_quarter.
template <class F>
__device__ __forceinline__ void ntt_inv_cross(typename DF<F>::W (&x)[4],
                                              const typename DF<F>::W* q0, int stride,
                                              const typename DF<F>::W* itwv, int lane, int qw) {
  using W = typename DF<F>::W;
  const int a = coop_a3(64 * qw + lane);
#pragma unroll
  for (int k = 0; k < 4; k++) x[k] = q0[(size_t)k * stride + a];
  gs2_const<F, 2, 3>(x[0], x[1], x[2], x[3], itwv);
  gs2_const<F, 1, 1>(x[0], x[2], x[1], x[3], itwv);
}

}  // namespace hb
