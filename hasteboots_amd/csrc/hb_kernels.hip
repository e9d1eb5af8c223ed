// HIP kernels for gfx950: fused gate bootstrapping (blind rotation + key switching +
// sample extraction, one wavefront per gate) and the NTT / external-product
// microbench kernels.
//
// Device layouts (all in HBM, written once by key upload):
//   BSK  [i < n][c: c_neg_s_m, c_m][j < br_len][p: a, b][N slots]  — NTT domain,
//        every word pre-multiplied by 1/N (so the inverse transform skips its
//        normalisation) and, for BabyBear, by R = 2^32 (Montgomery constant form).
//   KSK  [j < ks_len][p][N slots], same scaling.
//   Slot order is the forward-NTT output order (bit-reversed evaluation order).
//   Per-gate scratch (kScratchWordsPerGate words): the accumulator RLWE, 2N words, raw
//   Goldilocks words exactly as the reference's add_reduce produces them, then up to 2N
//   words for the byte digits of the rotated accumulator (store_digits8, store_digits5).
#include "hb_kernels.hpp"
#include "hb_ntt.hpp"

namespace hb {

template <class F>
struct DevConst;
template <>
struct DevConst<Gl> {
  using W = u64;
};
template <>
struct DevConst<Bb> {
  using W = u32;
};

// ----------------------------------------------------------------------------
// Gate input combination (tfhe/src/bfhe/evaluate.rs:113-231) on canonical words,
// returning the canonical value read by lwe_modulus_switch.
// ----------------------------------------------------------------------------
template <class F>
__device__ __forceinline__ typename DF<F>::W combine(int gate, typename DF<F>::W a,
                                                     typename DF<F>::W b, typename DF<F>::W c) {
  using W = typename DF<F>::W;
  W v;
  switch (gate) {
    case GATE_XOR:
    case GATE_XNOR: {
      W d = F::sub(a, b);  // sub_component_wise_ref then double_inplace
      v = F::add(d, d);
      break;
    }
    case GATE_MAJORITY:
      v = F::add(F::add(a, b), c);
      break;
    case GATE_BOOTSTRAP:
      v = a;
      break;
    default:  // nand, and, or, nor: add_component_wise_ref (lattice/src/lwe.rs:72-81)
      v = F::add(a, b);
  }
  return F::canon(v);
}

// lwe_modulus_switch (fhe_core/src/modulus_switch.rs:12-38): IEEE f64
// round(v * 2N / Q), half away from zero, 2N -> 0.
__device__ __forceinline__ u32 mod_switch(u64 v, double qf, u32 two_n) {
  double r = ::round((double)v * (double)two_n / qf);
  u32 t = (u32)(u64)r;
  return t >= two_n ? t - two_n : t;
}

// LUT entry k (tfhe/src/bfhe/evaluate.rs:267-465, lut_step = 1), or the caller's LUT.
template <class F>
__device__ __forceinline__ typename DF<F>::W lut_at(const BootArgs& a, int k) {
  using W = typename DF<F>::W;
  if (a.lut) return (W)a.lut[k];
  return k < (int)a.lut_mid ? (W)a.lut_lo : (W)a.lut_hi;
}

// Load 4 consecutive words (32 B Goldilocks / 16 B BabyBear) as one vector.
__device__ __forceinline__ void load4(const u64* __restrict__ p, u64 (&v)[4]) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
  ulonglong2 x = q[0], y = q[1];
  v[0] = x.x;
  v[1] = x.y;
  v[2] = y.x;
  v[3] = y.y;
}
__device__ __forceinline__ void load4(const u32* __restrict__ p, u32 (&v)[4]) {
  uint4 x = *reinterpret_cast<const uint4*>(p);
  v[0] = x.x;
  v[1] = x.y;
  v[2] = x.z;
  v[3] = x.w;
}

// acc += NTT(digit) ⊙ key for one gadget row: key rows (a, b) of N slots each. Goldilocks
// accumulators stay in [0, 2^64) (the inverse transform canonicalises its input); BabyBear
// accumulators are lazy 64-bit sums of the raw products (one v_mad_u64_u32 each), which
// gadget_mac folds and reduces.
// ntt_add_mul_assign_fast (algebra/src/polynomial/univariate/ntt_polynomial.rs:620-629).
// Key words are software-pipelined one 4-slot group ahead: group 0 arrives in (ka0, kb0),
// loaded by the caller before the forward transform, and group q+1 is in flight while
// group q is multiplied; the scheduling barriers keep each load ahead of the arithmetic it
// overlaps.
template <class F, int E, class A>
__device__ __forceinline__ void mac_row(A (&pa)[E], A (&pb)[E],
                                        const typename DF<F>::W (&x)[E],
                                        const typename DF<F>::W* __restrict__ ka,
                                        const typename DF<F>::W* __restrict__ kb,
                                        const typename DF<F>::W (&ka0)[4],
                                        const typename DF<F>::W (&kb0)[4], int lane) {
  using W = typename DF<F>::W;
  W va[4], vb[4];
#pragma unroll
  for (int r = 0; r < 4; r++) va[r] = ka0[r], vb[r] = kb0[r];
#pragma unroll
  for (int q = 0; q < E / 4; q++) {
    W na[4], nb[4];
    if (q + 1 < E / 4) {
      const int slot = 4 * (lane + 64 * (q + 1));
      load4(ka + slot, na);
      load4(kb + slot, nb);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if constexpr (F::IS_GOLDILOCKS) {
        gl_mac2(pa[4 * q + r], pb[4 * q + r], x[4 * q + r], va[r], vb[r]);
      } else {
        static_assert(sizeof(A) == 8, "BabyBear sums are 64-bit");
        pa[4 * q + r] += (u64)x[4 * q + r] * va[r];
        pb[4 * q + r] += (u64)x[4 * q + r] * vb[r];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    if (q + 1 < E / 4) {
#pragma unroll
      for (int r = 0; r < 4; r++) va[r] = na[r], vb[r] = nb[r];
    }
  }
}

// Forward transform of gadget digits (every word below 2^16): the twisted transform for
// Goldilocks (twv: the kTwxWords table of ntt_fwd_digits), the plain one for BabyBear (twv:
// psi_rev). The slot values are the same either way.
template <class F, int E, bool PF>
__device__ __forceinline__ void digit_ntt(typename DF<F>::W (&x)[E], typename DF<F>::W* lds,
                                          const typename DF<F>::W* twv, int lane) {
  if constexpr (F::IS_GOLDILOCKS && E == 16)
    ntt_fwd_digits<PF>(x, lds, twv, lane);
  else
    ntt_fwd_wave<F, E, true, PF>(x, lds, twv, lane);
}

// Σ over one gadget of digit NTTs: for j < len, x = NTT(digit_j(src)), acc += x ⊙ key_j.
// Decomposition: decompose_lsb_bits_inplace (native_polynomial.rs:340-348) on canonical
// values, digit_j = (v >> j*bits) & mask. twv: digit_ntt's forward table (LDS copy). S: the source word type (a field word, or u32 for a window of at most 32 bits of
// digits). A: the accumulator type (field words for Goldilocks, lazy 64-bit sums for
// BabyBear, folded after every second row: the transform's words are below 2P, so
// 2 (2P-1)(P-1) + 2^60 < 2^64). swap: pa takes the b products and pb the a products.
template <class F, int E, bool PF, class S, class A>
__device__ __forceinline__ void gadget_rows(A (&pa)[E], A (&pb)[E], const S (&src)[E], u32 len,
                                            u32 bits, const typename DF<F>::W* __restrict__ key,
                                            typename DF<F>::W* lds,
                                            const typename DF<F>::W* twv, int lane,
                                            bool swap = false) {
  using W = typename DF<F>::W;
  constexpr int N = 64 * E;
  const W mask = (W)((1u << bits) - 1u);
  const int oa = swap ? N : 0, ob = swap ? 0 : N;
  for (u32 j = 0; j < len; j++) {
    const W* ka = key + (size_t)(2 * j) * N + oa;
    const W* kb = key + (size_t)(2 * j) * N + ob;
    W ka0[4], kb0[4];
    load4(ka + 4 * lane, ka0);
    load4(kb + 4 * lane, kb0);
    __builtin_amdgcn_sched_barrier(0);
    W x[E];
    const u32 sh = j * bits;
#pragma unroll
    for (int e = 0; e < E; e++) x[e] = (W)(src[e] >> sh) & mask;
    digit_ntt<F, E, PF>(x, lds, twv, lane);  // digits < 2^bits <= 2^16
    mac_row<F, E>(pa, pb, x, ka, kb, ka0, kb0, lane);
    if constexpr (!F::IS_GOLDILOCKS) {
      if ((j & 1) == 1 && j + 1 < len) {
#pragma unroll
        for (int e = 0; e < E; e++) pa[e] = DF<F>::fold(pa[e]), pb[e] = DF<F>::fold(pb[e]);
      }
    }
  }
}

// gadget_rows on canonical field-word accumulators (pa, pb in and out).
template <class F, int E, bool PF = true, class S = typename DF<F>::W>
__device__ __forceinline__ void gadget_mac(typename DF<F>::W (&pa)[E], typename DF<F>::W (&pb)[E],
                                           const S (&src)[E], u32 len, u32 bits,
                                           const typename DF<F>::W* __restrict__ key,
                                           typename DF<F>::W* lds,
                                           const typename DF<F>::W* twv, int lane,
                                           bool swap = false) {
  if constexpr (F::IS_GOLDILOCKS) {
    gadget_rows<F, E, PF>(pa, pb, src, len, bits, key, lds, twv, lane, swap);
  } else {
    u64 sa[E], sb[E];
#pragma unroll
    for (int e = 0; e < E; e++) sa[e] = DF<F>::lazy_in(pa[e]), sb[e] = DF<F>::lazy_in(pb[e]);
    gadget_rows<F, E, PF>(sa, sb, src, len, bits, key, lds, twv, lane, swap);
#pragma unroll
    for (int e = 0; e < E; e++) pa[e] = DF<F>::lazy_out(sa[e]), pb[e] = DF<F>::lazy_out(sb[e]);
  }
}

// Byte digits of one rotated accumulator component (8-bit gadget digits of 64-bit words,
// decompose_lsb_bits_inplace with bits = 8, native_polynomial.rs:340-348) in the layout the
// digit loop reads: digit j of word k = lane + 64 (4m + b) is byte b of dword m of the 16
// bytes at dig[64 j + lane]. Per group of four words and half (digits 0-3 from the low
// dwords, 4-7 from the high), one byte permute per word pair interleaves two digits
// (a0.bj, a1.bj, a0.bj+1, a1.bj+1) and two more gather each digit's four bytes.
// v_perm_b32 byte selectors 0-3 pick bytes of its second source, 4-7 of its first.
__device__ __forceinline__ void store_digits8(const u64 (&v)[16], uint4* dig, int lane) {
  u32 d[8][4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const u32 a0 = (u32)(v[4 * m + 0] >> (32 * h)), a1 = (u32)(v[4 * m + 1] >> (32 * h));
      const u32 a2 = (u32)(v[4 * m + 2] >> (32 * h)), a3 = (u32)(v[4 * m + 3] >> (32 * h));
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const u32 sel = q ? 0x07030602u : 0x05010400u;
        const u32 t01 = __builtin_amdgcn_perm(a1, a0, sel);
        const u32 t23 = __builtin_amdgcn_perm(a3, a2, sel);
        d[4 * h + 2 * q][m] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
        d[4 * h + 2 * q + 1][m] = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 8; j++) dig[64 * j + lane] = make_uint4(d[j][0], d[j][1], d[j][2], d[j][3]);
}

// Digit j of four words (a0..a3 = the words shifted right by j * bits, low 32 bits) as one
// dword of bytes (a0, a1, a2, a3) & mask: two byte permutes gather the low bytes pairwise
// (a0.b0, a1.b0, a0.b1, a1.b1), a third the four low bytes.
__device__ __forceinline__ u32 pack_low_bytes(u32 a0, u32 a1, u32 a2, u32 a3, u32 mask4) {
  const u32 t01 = __builtin_amdgcn_perm(a1, a0, 0x05010400u);
  const u32 t23 = __builtin_amdgcn_perm(a3, a2, 0x05010400u);
  return __builtin_amdgcn_perm(t23, t01, 0x05040100u) & mask4;
}

// store_digits8's layout for a gadget of 13 digits of 5 bits (decompose_lsb_bits_inplace
// with bits = 5, native_polynomial.rs:340-348; digit 12 holds bits 60-63): one byte per
// digit, so the digit loop runs the byte-digit transform.
__device__ __forceinline__ void store_digits5(const u64 (&v)[16], uint4* dig, int lane) {
#pragma unroll
  for (int j = 0; j < 13; j++) {
    u32 d[4];
#pragma unroll
    for (int m = 0; m < 4; m++)
      d[m] = pack_low_bytes((u32)(v[4 * m] >> (5 * j)), (u32)(v[4 * m + 1] >> (5 * j)),
                            (u32)(v[4 * m + 2] >> (5 * j)), (u32)(v[4 * m + 3] >> (5 * j)),
                            0x1F1F1F1Fu);
    dig[64 * j + lane] = make_uint4(d[0], d[1], d[2], d[3]);
  }
}

// gadget_mac for a Goldilocks blind-rotation gadget of byte-sized digits, the digits read
// from their byte-transposed copy (store_digits8, store_digits5): one 16-byte load per
// digit, issued a digit ahead, in place of a 16-word source held in registers through the
// digit loop.
template <int E, bool PF>
__device__ __forceinline__ void gadget_mac_dig8(u64 (&pa)[E], u64 (&pb)[E],
                                                const uint4* __restrict__ dig, u32 len,
                                                const u64* __restrict__ key, u64* lds,
                                                const u64* twv, int lane) {
  static_assert(E == 16, "one 16-byte load holds a lane's 16 digits");
  constexpr int N = 64 * E;
  uint4 d = dig[lane];
  for (u32 j = 0; j < len; j++) {
    const u64* ka = key + (size_t)(2 * j) * N;
    u64 ka0[4], kb0[4];
    load4(ka + 4 * lane, ka0);
    load4(ka + N + 4 * lane, kb0);
    const uint4 dn = dig[64 * (j + 1 < len ? j + 1 : j) + lane];
    __builtin_amdgcn_sched_barrier(0);
    const u32 w[4] = {d.x, d.y, d.z, d.w};
    u64 x[E];
    ntt_fwd_bytes<PF>(x, w, lds, twv, lane);
    mac_row<Gl, E>(pa, pb, x, ka, ka + N, ka0, kb0, lane);
    d = dn;
  }
}

// Latency mode's share of the Goldilocks 8-bit gadget: digits j0 and j0 + 1 (nd = 2; pair
// mode: j0 alone, nd = 1) of the rotated component (`dst`, canonical words), keys from row
// j0 (`key`). The digits are packed from registers into ntt_fwd_bytes' layout (dword m byte
// g = the digit of word 4m + g) with four byte permutes per group of four words, as
// store_digits8 does for all eight. swap: pa takes the b products and pb the a products.
template <bool PF>
__device__ __forceinline__ void gadget_mac_bytes2(u64 (&pa)[16], u64 (&pb)[16],
                                                  const u64 (&dst)[16], u32 j0,
                                                  const u64* __restrict__ key, u64* lds,
                                                  const u64* twv, int lane, bool swap,
                                                  u32 nd = 2) {
  constexpr int N = 1024;
  const int oa = swap ? N : 0, ob = swap ? 0 : N;
  u32 d[4], dn[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    const u32 a0 = (u32)(dst[4 * m + 0] >> (8 * j0)), a1 = (u32)(dst[4 * m + 1] >> (8 * j0));
    const u32 a2 = (u32)(dst[4 * m + 2] >> (8 * j0)), a3 = (u32)(dst[4 * m + 3] >> (8 * j0));
    const u32 t01 = __builtin_amdgcn_perm(a1, a0, 0x05010400u);
    const u32 t23 = __builtin_amdgcn_perm(a3, a2, 0x05010400u);
    d[m] = __builtin_amdgcn_perm(t23, t01, 0x05040100u);
    dn[m] = __builtin_amdgcn_perm(t23, t01, 0x07060302u);
  }
#pragma unroll 1
  for (u32 j = 0; j < nd; j++) {
    const u64* ka = key + (size_t)(2 * j) * N + oa;
    const u64* kb = key + (size_t)(2 * j) * N + ob;
    u64 ka0[4], kb0[4];
    load4(ka + 4 * lane, ka0);
    load4(kb + 4 * lane, kb0);
    __builtin_amdgcn_sched_barrier(0);
    u64 x[16];
    ntt_fwd_bytes<PF>(x, d, lds, twv, lane);
    mac_row<Gl, 16>(pa, pb, x, ka, kb, ka0, kb0, lane);
#pragma unroll
    for (int m = 0; m < 4; m++) d[m] = dn[m];
  }
}

// Latency mode's share of a Goldilocks gadget of BITS-bit digits (BITS < 8): `len` digits
// of the window `win` (the rotated component shifted right to the wave's first digit, keys
// from that digit's row), each packed from registers into ntt_fwd_bytes' layout. swap: as
// gadget_mac_bytes2.
template <int BITS, bool PF>
__device__ __forceinline__ void gadget_mac_window_bytes(u64 (&pa)[16], u64 (&pb)[16],
                                                        const u32 (&win)[16], u32 len,
                                                        const u64* __restrict__ key, u64* lds,
                                                        const u64* twv, int lane, bool swap) {
  constexpr int N = 1024;
  constexpr u32 mask4 = ((1u << BITS) - 1u) * 0x01010101u;
  const int oa = swap ? N : 0, ob = swap ? 0 : N;
#pragma unroll 1
  for (u32 j = 0; j < len; j++) {
    const u64* ka = key + (size_t)(2 * j) * N + oa;
    const u64* kb = key + (size_t)(2 * j) * N + ob;
    u64 ka0[4], kb0[4];
    load4(ka + 4 * lane, ka0);
    load4(kb + 4 * lane, kb0);
    __builtin_amdgcn_sched_barrier(0);
    const u32 sh = j * BITS;
    u32 d[4];
#pragma unroll
    for (int m = 0; m < 4; m++)
      d[m] = pack_low_bytes(win[4 * m] >> sh, win[4 * m + 1] >> sh, win[4 * m + 2] >> sh,
                            win[4 * m + 3] >> sh, mask4);
    u64 x[16];
    ntt_fwd_bytes<PF>(x, d, lds, twv, lane);
    mac_row<Gl, 16>(pa, pb, x, ka, kb, ka0, kb0, lane);
  }
}

// Wave-wide sum mod P of one value per lane.
template <class F>
__device__ __forceinline__ typename DF<F>::W wave_sum(typename DF<F>::W v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    typename DF<F>::W o = __shfl_xor(v, off, 64);
    v = DF<F>::add(v, o);
  }
  return v;
}

// Orders one wave's accumulator (global) and LDS accesses between steps of a gate:
// workgroup-scope release/acquire fences (the wave's own stores complete and become
// visible to its later loads) without a barrier, so the waves of a workgroup, one gate
// each, never wait for one another.
__device__ __forceinline__ void gate_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Goldilocks sets whose blind-rotation digits take the byte-digit path: 8 digits of 8 bits
// (8) or 13 digits of 5 bits (5); 0 for any other gadget.
__host__ __device__ inline int dig_bits(const BootArgs& A) {
  if (!A.is_goldilocks) return 0;
  if (A.br_bits == 8 && A.br_len == 8) return 8;
  if (A.br_bits == 5 && A.br_len == 13) return 5;
  return 0;
}

// ----------------------------------------------------------------------------
// Fused bootstrap: one wavefront per gate, 1, 2 or 4 gates per workgroup (the
// workgroup shares the LDS twiddle tables).
// EvaluationKey::bootstrap (tfhe/src/bfhe/evaluate.rs:42-60):
//   lwe_modulus_switch → BinaryBlindRotationKey::blind_rotate (binary.rs:28-104)
//   → acc.b[0] += Q/8 → KeySwitchingRLWEKey::key_switch_for_rlwe
//   (fhe_core/src/key_switch/rlwe.rs:115-137) → extract_partial_lwe_locally.
// ----------------------------------------------------------------------------
// Step 1 of a gate, by threads tid = 0..nthreads-1 of its workgroup (no barrier inside):
// gate combination and modulus switch of all n+1 coefficients into ms. Inputs are dense
// arrays (gate g at row g) or, for circuits, wire indices into one device-resident wire
// buffer: idx[3g..3g+2] = (input 0, input 1, output).
template <class F>
__device__ __forceinline__ void gate_modswitch(const BootArgs& A, int g, int tid, int nthreads,
                                               unsigned short* ms) {
  using W = typename DF<F>::W;
  constexpr int N = 1024;
  const u32 n1 = A.n + 1;
  const u32* wi = A.idx ? A.idx + 3 * (size_t)g : nullptr;
  const u64* c0 = A.c0 + (size_t)(wi ? wi[0] : g) * n1;
  const u64* c1 = wi ? A.c0 + (size_t)wi[1] * n1 : (A.c1 ? A.c1 + (size_t)g * n1 : nullptr);
  const u64* c2 = (!wi && A.c2) ? A.c2 + (size_t)g * n1 : nullptr;
  for (u32 k = tid; k < n1; k += nthreads) {
    W a = (W)c0[k];
    W b = c1 ? (W)c1[k] : (W)0;
    W c = c2 ? (W)c2[k] : (W)0;
    W v = combine<F>(A.gate, a, b, c);
    ms[k] = (unsigned short)mod_switch((u64)v, A.qf, 2 * N);
  }
}

// Step 2 (after step 1 is visible): ACC = (0, LUT * X^-b) (binary.rs:40-64).
template <class F>
__device__ __forceinline__ void gate_lut_init(const BootArgs& A, int tid, int nthreads,
                                              const unsigned short* ms, typename DF<F>::W* accA,
                                              typename DF<F>::W* accB) {
  using W = typename DF<F>::W;
  constexpr int N = 1024;
  const u32 bms = ms[A.n];
  for (int k = tid; k < N; k += nthreads) {
    W val;
    if (bms == 0) {
      val = lut_at<F>(A, k);
    } else {
      const int nb = 2 * N - (int)bms;
      bool neg;
      int src;
      if (nb <= N) {
        neg = k < nb;
        src = neg ? k - nb + N : k - nb;
      } else {
        const int r = nb - N;
        neg = k >= r;
        src = neg ? k - r : k - r + N;
      }
      val = lut_at<F>(A, src);
      if (neg) val = F::neg_raw(val);
    }
    accA[k] = 0;
    accB[k] = val;
  }
}

// RLWE::add_assign_element_wise (rlwe.rs:287-290) after the inverse transform, raw-word add:
// out[k] = acc[k] + p[e] at k = lane + 64 e. out is acc itself (untraced) or the next history
// slot (traced), a run-time offset the compiler cannot see through, so word by word every load
// would wait behind the previous store: one memory round trip per word. The loads of a group
// of G words go out together, then the group's adds and stores.
template <class F, int E, int G>
__device__ __forceinline__ void acc_add(typename DF<F>::W* out, const typename DF<F>::W* acc,
                                        const typename DF<F>::W (&p)[E], int lane) {
  using W = typename DF<F>::W;
  static_assert(E % G == 0, "whole groups");
#pragma unroll
  for (int e0 = 0; e0 < E; e0 += G) {
    W v[G];
#pragma unroll
    for (int j = 0; j < G; j++) v[j] = acc[lane + 64 * (e0 + j)];
#pragma unroll
    for (int j = 0; j < G; j++) out[lane + 64 * (e0 + j)] = F::add_cy(v[j], p[e0 + j]);
  }
}

// RLWE::mul_monic_monomial_sub_one_inplace (lattice/src/rlwe.rs:437-487) for one component:
// dst[e] = ±acc[k - r mod N] - acc[k] at k = lane + 64e, the ± being the raw-word negation
// (untraced) or the exact negation of the canonical value (traced: the NTT product
// ACC * X^r of binary.rs:158-181).
template <class F, int E>
__device__ __forceinline__ void rotate_sub(const typename DF<F>::W* acc, int r, bool traced,
                                           int lane, typename DF<F>::W (&dst)[E]) {
  using W = typename DF<F>::W;
  constexpr int N = 64 * E;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int k = lane + 64 * e;
    bool neg;
    int src;
    if (r <= N) {
      neg = k < r;
      src = neg ? k - r + N : k - r;
    } else {
      const int rr = r - N;
      neg = k >= rr;
      src = neg ? k - rr : k - rr + N;
    }
    W v = acc[src];
    if constexpr (F::IS_GOLDILOCKS) {
      // untraced: P - v (mod 2^64) is neg_raw(v) for v != 0 and P = 0 (mod P) for v = 0;
      // the raw subtraction keeps the residue of any words, so one canonicalisation ends
      if (neg) v = (u64)F::P - (traced ? F::canon(v) : v);
      dst[e] = F::canon(F::sub(v, acc[k]));
    } else {
      if (neg) v = traced ? F::neg_raw(F::canon(v)) : F::neg_raw(v);
      dst[e] = F::sub_canon(F::canon(v), F::canon(acc[k]));
    }
  }
}

// rotate_sub (untraced) on a component held as cn = [C | M], C[k] = canon(acc[k]) and
// M[k] = canon(neg_raw(acc[k])), 2N words: position k reads cn[(k - r) mod 2N], which is
// M[k - r mod N] when the monomial's sign flips it (k < r for r <= N, k >= r - N above) and
// C[k - r mod N] otherwise.
template <class F, int E>
__device__ __forceinline__ void rotate_sub_cn(const typename DF<F>::W* cn, int r, int lane,
                                              typename DF<F>::W (&dst)[E]) {
  constexpr int N = 64 * E;
  const int t0 = lane - r;
#pragma unroll
  for (int e = 0; e < E; e++)
    dst[e] = F::sub_canon(cn[(t0 + 64 * e) & (2 * N - 1)], cn[lane + 64 * e]);
}

// Step 4 by one wave: acc.b[0] += Q/8 (evaluate.rs:52), key switching and sample extraction.
// NEGA: accA holds canon(neg_raw(acc.a)) (the M half of rotate_sub_cn's layout).
template <class F, bool PF = true, bool NEGA = false>
__device__ __forceinline__ void keyswitch_extract(const BootArgs& A, int g,
                                                  const typename DF<F>::W* accA,
                                                  const typename DF<F>::W* accB,
                                                  typename DF<F>::W* lds,
                                                  const typename DF<F>::W* twl,
                                                  const typename DF<F>::W* itwl, int lane) {
  using W = typename DF<F>::W;
  constexpr int E = 16;
  constexpr int N = 64 * E;
  const u32 n = A.n, n1 = A.n + 1;
  const W* __restrict__ ksk = static_cast<const W*>(A.ksk);
  const u32* wi = A.idx ? A.idx + 3 * (size_t)g : nullptr;
  const W b0 = F::add(accB[0], (W)A.q8);
  W src[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int k = lane + 64 * e;
    // polynomial.neg_assign() (rlwe.rs:1159)
    src[e] = NEGA ? accA[k] : F::canon(F::neg_raw(accA[k]));
  }
  W pa[E], pb[E];
#pragma unroll
  for (int e = 0; e < E; e++) pa[e] = pb[e] = 0;
  gadget_mac<F, E, PF>(pa, pb, src, A.ks_len, A.ks_bits, ksk, lds, twl, lane);
  // b' = coefficient 0 of INTT(init.b + Σ) = b0 + Σ_slots Σ (1/N folded into the KSK).
  W s = 0;
#pragma unroll
  for (int e = 0; e < E; e++) s = DF<F>::add(s, pb[e]);
  s = wave_sum<F>(s);
  const W bout = F::canon(DF<F>::add(F::canon(b0), s));
  ntt_inv_wave<F, E, PF>(pa, lds, itwl, lane);
  // extract_partial_lwe_locally (rlwe.rs:393-402): a' = [A0, -A_{N-1}, ..., -A_1][..n].
  u64* out = A.out + (size_t)(wi ? wi[2] : g) * n1;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int p = lane + 64 * e;
    if (p == 0) {
      out[0] = (u64)pa[e];
    } else {
      const int k = N - p;
      if (k < (int)n) out[k] = (u64)F::canon(F::neg_raw(pa[e]));
    }
  }
  if (lane == 0) out[n] = (u64)bout;
}

// ----------------------------------------------------------------------------
// Fused bootstrap, one wavefront per gate.
// ----------------------------------------------------------------------------
// PF: the transforms request their twiddles a stage ahead (one wave per SIMD); without it
// they read them just before use, which leaves registers for two waves per SIMD.
// DIG (dig_bits of the set, the launcher's choice; 0 for the other sets): 8 or 5, the digit
// loop reads byte digits from the gate's scratch (store_digits8 / store_digits5,
// gadget_mac_dig8). A separate instantiation each, so that the register-held digit source
// of the other sets does not weigh on their register allocation.
template <class F, bool PF = true, int DIG = 0>
__device__ __forceinline__ void bootstrap_gate(const BootArgs& A, int g, typename DF<F>::W* lds,
                                               unsigned short* ms,
                                               const typename DF<F>::W* twl,
                                               const typename DF<F>::W* itwl) {
  using W = typename DF<F>::W;
  constexpr int E = 16;
  constexpr int N = 64 * E;
  const int lane = threadIdx.x & 63;
  const u32 n = A.n;
  const W* __restrict__ bsk = static_cast<const W*>(A.bsk);
  // Untraced: one accumulator per gate, updated in place. Traced (the vfhe path): an
  // (n+1)-slot history per gate; round i reads slot i and writes slot i+1, so the
  // history is the core trace (the acc entering every round, then the final acc).
  const bool traced = A.trace != nullptr;
  W* const scratch = static_cast<W*>(A.acc) + (size_t)g * kScratchWordsPerGate;
  W* accA = traced ? static_cast<W*>(A.trace) + (size_t)g * (n + 1) * 2 * N : scratch;
  W* accB = accA + N;
  uint4* const dig = reinterpret_cast<uint4*>(scratch + 2 * N);
  static_assert(DIG == 0 || F::IS_GOLDILOCKS, "byte digits of 64-bit words");

  gate_modswitch<F>(A, g, lane, 64, ms);
  wave_sync();
  gate_lut_init<F>(A, lane, 64, ms, accA, accB);
  gate_sync();

  // 3. blind rotation: for a_i != 0, ACC += ExtProd((X^a_i - 1) ACC, RGSW(s_i)).
  const size_t rgsw_words = (size_t)2 * A.br_len * 2 * N;
  const size_t step = traced ? 2 * N : 0;
  for (u32 i = 0; i < n; i++, accA += step, accB += step) {
    const int r = ms[i];
    W* outA = accA + step;
    W* outB = accB + step;
    if (r == 0) {  // binary.rs:70 skips the round; the vfhe path adds a zero product
      if (traced) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const int k = lane + 64 * e;
          outA[k] = accA[k];
          outB[k] = accB[k];
        }
        gate_sync();
      }
      continue;
    }
    const W* key = bsk + (size_t)i * rgsw_words;
    W pa[E], pb[E];
#pragma unroll
    for (int e = 0; e < E; e++) pa[e] = pb[e] = 0;
#pragma unroll 1
    for (int comp = 0; comp < 2; comp++) {
      W dst[E];
      rotate_sub<F, E>(comp ? accB : accA, r, traced, lane, dst);
      // c_neg_s_m for the a half, c_m for the b half (rlwe.rs:579-593).
      const W* ckey = key + (size_t)comp * A.br_len * 2 * N;
      if constexpr (DIG != 0) {
        if constexpr (DIG == 8)
          store_digits8(dst, dig, lane);
        else
          store_digits5(dst, dig, lane);
        gate_sync();
        gadget_mac_dig8<E, PF>(pa, pb, dig, A.br_len, ckey, lds, twl, lane);
      } else {
        gadget_mac<F, E, PF>(pa, pb, dst, A.br_len, A.br_bits, ckey, lds, twl, lane);
      }
    }
    // NTTRLWE::inverse_transform_inplace then RLWE::add_assign_element_wise
    // (rlwe.rs:765-777, 287-290): raw-word add of the canonical product.
    // groups of 16 words for Goldilocks with byte digits, 4 otherwise (4096 gates at (4, 2),
    // word by word / 4 / 16: Goldilocks 200.6 / 196.0 / 195.1 ms; Zama 304.1 / 297.7 / 298.2 ms
    // (8: 297.7); BabyBear 57.0 / 53.3 / 53.3 ms (8: 53.4))
    constexpr int kAccGroup = F::IS_GOLDILOCKS && DIG == 8 ? 16 : 4;
    ntt_inv_wave<F, E, PF>(pa, lds, itwl, lane);
    acc_add<F, E, kAccGroup>(outA, accA, pa, lane);
    ntt_inv_wave<F, E, PF>(pb, lds, itwl, lane);
    acc_add<F, E, kAccGroup>(outB, accB, pb, lane);
    gate_sync();
  }
  keyswitch_extract<F, PF>(A, g, accA, accB, lds, twl, itwl, lane);
}

// ----------------------------------------------------------------------------
// Latency mode: one gate per workgroup of WAVES = 4 or 8 wavefronts. Wave w takes component
// c = w / Q of the rotated accumulator (Q = WAVES / 2 waves per component) and the q = w % Q
// share of its gadget digits, [q l / Q, (q + 1) l / Q), accumulating a partial NTT-domain
// product. The partials meet in LDS (with 8 waves, waves w and w + 4 first add theirs
// pairwise, a half each). Then every wave finishes quarters of the result (sum of the
// partials, inverse NTT, accumulator add) with ntt_inv_quarter / ntt_inv_cross, exchanging
// through the waves' LDS buffers, `ldsw` words apart: with 8 waves, waves 0-3 a quarter of
// the a half each and waves 4-7 of the b half; with 4, wave w quarter w of both halves.
// Wave 0 then key-switches. With 8 waves each SIMD holds two waves of the gate, and a
// wave's digit window (at most 4 digits) fits 32 bits; DIG (dig_bits, 8 waves): a wave's
// digits go through the byte-digit transform (two 8-bit digits: gadget_mac_bytes2; 5-bit
// digits: gadget_mac_window_bytes).
// The accumulator stays in LDS for the whole gate, so a round's last barrier waits for LDS
// stores only and the next rotation reads LDS. With 8 waves each wave keeps the raw words
// of the 16 positions it updates in registers for the whole gate, and LDS holds every
// component as rotate_sub_cn's [C | M] (accA = part, slot 0; accB = part + 2N, slot 1;
// the slots are written only after the first barrier of a round, the partials of waves w
// and w + 4 meeting through their exchange buffers first). With 4 waves the raw words
// (accA, accB: N words each) have their own LDS. TR: instantiated for traced gates
// (A.trace set; 8 waves only, which the launchers keep to), so the untraced kernels carry
// none of it.
// ----------------------------------------------------------------------------
template <class F, int WAVES, int DIG = 0, bool TR = false>
__device__ __forceinline__ void bootstrap_gate_split(const BootArgs& A, int g,
                                                     typename DF<F>::W* lds, int ldsw,
                                                     unsigned short* ms,
                                                     typename DF<F>::W* part,
                                                     typename DF<F>::W* accA,
                                                     typename DF<F>::W* accB,
                                                     const typename DF<F>::W* twl,
                                                     const typename DF<F>::W* itwl) {
  static_assert(WAVES == 4 || WAVES == 8, "4 or 8 wavefronts per gate");
  static_assert(DIG == 0 || (F::IS_GOLDILOCKS && WAVES == 8), "byte digits with 8 waves");
  using W = typename DF<F>::W;
  constexpr int E = 16;
  constexpr int N = 64 * E;
  constexpr int Q = WAVES / 2;
  // Twiddles requested a stage ahead even with two waves per SIMD: the waves of one gate run
  // in step, so they hide little of each other's LDS latency (256 Goldilocks gates with 8
  // waves: 18.8 -> 17.9 ms).
  constexpr bool PF = true;
  // the wave index as a scalar (wave_id) for BabyBear: 1024 gates 16.5 -> 16.2 ms; Goldilocks
  // (57.0 -> 59.3 ms) and Zama (92.7 -> 96.6 ms) slower with it, so per lane there
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = F::IS_GOLDILOCKS ? tid >> 6 : wave_id();
  const int comp = wave / Q, q = wave % Q;
  const u32 n = A.n;
  const W* __restrict__ bsk = static_cast<const W*>(A.bsk);

  gate_modswitch<F>(A, g, tid, 64 * WAVES, ms);
  __syncthreads();
  gate_lut_init<F>(A, tid, 64 * WAVES, ms, accA, accB);
  __syncthreads();
  // 8 waves: the raw words of positions 64 q + lane + 256 k of component `comp`, which this
  // wave updates at the end of every round, and the [C | M] form of the component. Traced
  // (the vfhe path, 8 waves only): M negates the canonical word (rotate_sub's traced
  // semantics), and every wave writes its raw words to the gate's history slot by slot
  // (slot 0 the initial accumulator, slot i + 1 the one after round i, copied for a
  // skipped round), the layout of bootstrap_gate's traced accumulator.
  const bool traced = TR && WAVES == 8 && A.trace != nullptr;
  W* const hist = traced ? static_cast<W*>(A.trace) + (size_t)g * (n + 1) * 2 * N : nullptr;
  W raw[4];
  if constexpr (WAVES == 8) {
    W* ac = comp ? accB : accA;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int pos = 64 * q + lane + 256 * k;
      raw[k] = ac[pos];
      ac[pos] = F::canon(raw[k]);
      ac[N + pos] = F::canon(F::neg_raw(traced ? F::canon(raw[k]) : raw[k]));
      if (traced) hist[comp * N + pos] = raw[k];
    }
    __syncthreads();
  }

  const size_t rgsw_words = (size_t)2 * A.br_len * 2 * N;
  // this wave's digits: [j0, j1) of its component's gadget, window qd of Q. With 8 waves
  // the b component's windows run in reverse order, so that when the windows differ in
  // length (Zama: 3, 3, 3, 4 digits) the SIMD running waves w and w + 4 does not get both
  // long ones.
  const int qd = (WAVES == 8 && comp) ? Q - 1 - q : q;
  const u32 j0 = (u32)qd * A.br_len / Q, j1 = (u32)(qd + 1) * A.br_len / Q;
  // partial slots [4][a, b][e][lane]: with 4 waves wave w stores slot w; with 8 waves
  // wave w < 4 the a half of slot w and wave w + 4 its b half, each the sum of both waves'
  // partials
  W* mine = part + (size_t)(wave & 3) * 2 * N;
  for (u32 i = 0; i < n; i++) {
    const int r = ms[i];
    if (r == 0) {  // uniform across the workgroup
      if constexpr (WAVES == 8) {
        if (traced) {
#pragma unroll
          for (int k = 0; k < 4; k++)
            hist[(size_t)(i + 1) * 2 * N + comp * N + 64 * q + lane + 256 * k] = raw[k];
        }
      }
      continue;
    }
    const W* key = bsk + (size_t)i * rgsw_words + (size_t)comp * A.br_len * 2 * N;
    W pa[E], pb[E];
#pragma unroll
    for (int e = 0; e < E; e++) pa[e] = pb[e] = 0;
    {
      W dst[E];
      const W* ac = comp ? accB : accA;
      if constexpr (WAVES == 8)
        rotate_sub_cn<F, E>(ac, r, lane, dst);
      else
        rotate_sub<F, E>(ac, r, false, lane, dst);
      // digits j0..j1-1: shift the digit window by j0 digits, keys from row j0
      const W* kj = key + (size_t)j0 * 2 * N;
      if constexpr (DIG == 8) {
        gadget_mac_bytes2<PF>(pa, pb, dst, j0, kj, lds, twl, lane, wave >= 4);
      } else if constexpr (WAVES == 8) {
        u32 win[E];
#pragma unroll
        for (int e = 0; e < E; e++) win[e] = (u32)(dst[e] >> (j0 * A.br_bits));
        if constexpr (DIG != 0)
          gadget_mac_window_bytes<DIG, PF>(pa, pb, win, j1 - j0, kj, lds, twl, lane, wave >= 4);
        else
          gadget_mac<F, E, PF>(pa, pb, win, j1 - j0, A.br_bits, kj, lds, twl, lane, wave >= 4);
      } else {
        if (j0 > 0) {
#pragma unroll
          for (int e = 0; e < E; e++) dst[e] >>= j0 * A.br_bits;
        }
        gadget_mac<F, E, PF>(pa, pb, dst, j1 - j0, A.br_bits, kj, lds, twl, lane);
      }
    }
    if constexpr (WAVES == 8) {
      // waves w and w + 4 meet in slot w: pa holds the half a wave keeps (the a products
      // for waves 0-3, the b products for waves 4-7, whose keys were swapped) and pb the
      // half it hands over in its own exchange buffer
#pragma unroll
      for (int e = 0; e < E; e++) lds[e * 64 + lane] = pb[e];
      __syncthreads();
      const W* take = lds + (ptrdiff_t)(wave < 4 ? 4 : -4) * ldsw;
      W* put = mine + (wave < 4 ? 0 : N);
#pragma unroll
      for (int e = 0; e < E; e++) put[e * 64 + lane] = DF<F>::add(pa[e], take[e * 64 + lane]);
    } else {
      // partial products to LDS, [poly][e][lane]
#pragma unroll
      for (int e = 0; e < E; e++) {
        mine[e * 64 + lane] = pa[e];
        mine[N + e * 64 + lane] = pb[e];
      }
    }
    __syncthreads();
    if constexpr (WAVES == 8) {
      // waves 0-3 finish the a half, waves 4-7 the b half, a quarter of the slots each:
      // sum of the four partials, the four-wave inverse transform, accumulator add
      const int c = wave >> 2, qw = wave & 3;
      W x[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int e = 4 * qw + k;
        W v = part[(size_t)c * N + e * 64 + lane];
#pragma unroll
        for (int w = 1; w < 4; w++) v = DF<F>::add(v, part[(size_t)w * 2 * N + c * N + e * 64 + lane]);
        x[k] = v;
      }
      ntt_inv_quarter<F>(x, lds, itwl, lane, qw);
      __syncthreads();
      ntt_inv_cross<F>(x, lds - (ptrdiff_t)qw * ldsw, ldsw, itwl, lane, qw);
      W* acc = c ? accB : accA;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int pos = 64 * qw + lane + 256 * k;
        raw[k] = F::add_cy(raw[k], x[k]);
        acc[pos] = F::canon(raw[k]);
        acc[N + pos] = F::canon(F::neg_raw(traced ? F::canon(raw[k]) : raw[k]));
        if (traced) hist[(size_t)(i + 1) * 2 * N + c * N + pos] = raw[k];
      }
    } else {
      // wave w finishes quarter w of both halves: the a quarter in words [0, 320) of its
      // LDS buffer, the b quarter in [320, 640)
      const int qw = wave;
      W xa[4], xb[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int e = 4 * qw + k;
        W va = part[(size_t)e * 64 + lane], vb = part[(size_t)N + e * 64 + lane];
#pragma unroll
        for (int w = 1; w < 4; w++) {
          va = DF<F>::add(va, part[(size_t)w * 2 * N + e * 64 + lane]);
          vb = DF<F>::add(vb, part[(size_t)w * 2 * N + N + e * 64 + lane]);
        }
        xa[k] = va, xb[k] = vb;
      }
      ntt_inv_quarter<F>(xa, lds, itwl, lane, qw);
      ntt_inv_quarter<F>(xb, lds + 320, itwl, lane, qw);
      __syncthreads();
      ntt_inv_cross<F>(xa, lds - (ptrdiff_t)qw * ldsw, ldsw, itwl, lane, qw);
      ntt_inv_cross<F>(xb, lds - (ptrdiff_t)qw * ldsw + 320, ldsw, itwl, lane, qw);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int pos = 64 * qw + lane + 256 * k;
        accA[pos] = F::add_cy(accA[pos], xa[k]);
        accB[pos] = F::add_cy(accB[pos], xb[k]);
      }
    }
    __syncthreads();
  }
  if (wave == 0) {
    // b0 = acc.b[0] + Q/8 is canonicalised before use, so C[0] serves for the raw word
    if constexpr (WAVES == 8)
      keyswitch_extract<F, PF, true>(A, g, accA + N, accB, lds, twl, itwl, lane);
    else
      keyswitch_extract<F, PF>(A, g, accA, accB, lds, twl, itwl, lane);
  }
}

// ----------------------------------------------------------------------------
// Pair mode (untraced; Goldilocks with byte-digit gadgets, BabyBear with at most 4 digits):
// one gate on two workgroups of 8 wavefronts on two CUs. Workgroups pair up in the order
// they start: each takes a ticket from the launch's counter (pair_ticket), and tickets 2g
// (c = 0) and 2g + 1 (c = 1) carry gate g, so a workgroup only ever waits for one that is
// running or for the next one of its launch to start, whatever else holds the CUs.
// Workgroup c owns accumulator component c: it rotates it ([C | M] in LDS, as
// bootstrap_gate_split keeps it), wave w runs digit row w of it (DIG = 8: one digit per
// wave, half of a latency-mode wave's rows; DIG = 5: one or two of the 13), and after the
// balanced partial sums its waves hold the
// quarters of both output halves. The half of the other component goes to the partner
// through global memory (scratch, [sender][parity][quarter][e][lane], two parities by
// exchange count), quarter by quarter, each with an agent-scope release flag; the waves of
// half c wait for the partner's quarters, add them, and finish half c (cooperative inverse,
// accumulator update) as latency mode does. Both workgroups skip the same rounds (each
// computes the modulus switch), so exchange e pairs up. Workgroup 1 hands the b component's
// word 0 to workgroup 0, which key-switches and writes the output.
// The exchange runs on agent-scope relaxed atomics, which the compiler emits as sc1 loads
// and stores, coherent across the XCDs' L2s: the sender stores its quarter, waits for the
// stores to complete (s_waitcnt vmcnt(0)) and stores the flag; the receiver polls the flag
// and then loads the quarter, both past its L2. Agent-scope release/acquire fences (L2
// write-back and invalidation every round) cost 6 ms at 64 gates.
// Waits poll with relaxed agent-scope loads and end after A.pair_wait_ticks ticks of the
// device's constant-rate clock (0: at once, without polling); a timeout records the
// launch's epoch in flag word 10, which ends the waits of both workgroups, and in the
// launch's status word, which the host reads after the launch and reports as an error; a
// wave whose wait failed waits no more (its outputs are wrong), so every wave reaches the
// end after at most one timeout.
// ----------------------------------------------------------------------------
__device__ __forceinline__ bool pair_wait(const u64* f, u64 want, u64* abort_word, u64 epoch,
                                          const BootArgs& A) {
  if (A.pair_wait_ticks != 0) {
    const u64 t0 = wall_clock64();
    for (;;) {
      if (__hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want) return true;
      if (__hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == epoch)
        return false;
      if (wall_clock64() - t0 > A.pair_wait_ticks) break;
      __builtin_amdgcn_s_sleep(1);
    }
  }
  __hip_atomic_store(abort_word, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(A.pair_status, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return false;
}

// This workgroup's ticket in its launch: the launch's workgroups take consecutive values of
// the counter pair_status[1] from pair_ticket0 on (launches on one counter run one after
// another, and each of the 2 * count workgroups takes one). A ticket outside [0, tickets)
// would mean a counter out of step with the host: the workgroup records the launch as
// failed and touches nothing else (its partner's wait then times out).
__device__ __forceinline__ bool pair_ticket(u64* status, u64 ticket0, u32 tickets, u32 epoch,
                                            u32* slot, u32* t) {
  if (threadIdx.x == 0) {
    const u64 v = __hip_atomic_fetch_add(status + 1, (u64)1, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) - ticket0;
    *slot = v < tickets ? (u32)v : ~0u;
    if (v >= tickets)
      __hip_atomic_store(status, (u64)epoch << 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  *t = *slot;
  return *t != ~0u;
}

__device__ __forceinline__ void pair_put(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 pair_get(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void pair_put(u32* p, u32 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 pair_get(const u32* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// The flag store after every earlier pair_put of the wave has completed.
__device__ __forceinline__ void pair_signal(u64* f, u64 v) {
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __hip_atomic_store(f, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class F, int DIG>
__device__ __forceinline__ void bootstrap_gate_pair(const BootArgs& A, int g, int c,
                                                    typename DF<F>::W* lds, int ldsw,
                                                    unsigned short* ms, typename DF<F>::W* part,
                                                    const typename DF<F>::W* twl,
                                                    const typename DF<F>::W* itwl) {
  static_assert(F::IS_GOLDILOCKS ? (DIG == 8 || DIG == 5) : DIG == 0, "gadget of the field");
  using W = typename DF<F>::W;
  constexpr int E = 16;
  constexpr int N = 64 * E;
  constexpr bool PF = true;
  // the wave index as a scalar (wave_id) for BabyBear and Zama: one NAND 4.80 -> 4.69 and
  // 16.2 -> 15.8 ms; Goldilocks with byte digits 9.6 -> 9.8 ms with it, so per lane there
  constexpr bool kScalarWave = !(F::IS_GOLDILOCKS && DIG == 8);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = kScalarWave ? wave_id() : tid >> 6;
  const int ch = wave >> 2, qw = wave & 3;  // the output half and quarter this wave finishes
  const bool fin = ch == c;                 // ... in this workgroup; the others send theirs
  const u32 n = A.n;
  const W* __restrict__ bsk = static_cast<const W*>(A.bsk);
  W* const xch = static_cast<W*>(A.acc) + (size_t)g * kScratchWordsPerGate;
  u64* const fl = A.pair_flags + (size_t)g * kPairFlagWords;
  const u64 ep = (u64)A.pair_epoch << 32;

  gate_modswitch<F>(A, g, tid, 64 * 8, ms);
  __syncthreads();
  gate_lut_init<F>(A, tid, 64 * 8, ms, part, part + N);  // both components, raw
  __syncthreads();
  // the finishing waves keep the raw words of their positions of component c; [C | M] of
  // component c in part[0, 2N)
  W raw[4];
  if (fin) {
#pragma unroll
    for (int k = 0; k < 4; k++) raw[k] = part[(c ? N : 0) + 64 * qw + lane + 256 * k];
  }
  __syncthreads();
  if (fin) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int pos = 64 * qw + lane + 256 * k;
      part[pos] = F::canon(raw[k]);
      part[N + pos] = F::canon(F::neg_raw(raw[k]));
    }
  }
  __syncthreads();

  const size_t rgsw_words = (size_t)2 * A.br_len * 2 * N;
  W* mine = part + (size_t)(wave & 3) * 2 * N;
  u32 ex = 0;         // exchanges so far
  bool alive = true;  // no wait of this wave has timed out
  for (u32 i = 0; i < n; i++) {
    const int r = ms[i];
    if (r == 0) continue;  // uniform across both workgroups
    const W* key = bsk + (size_t)i * rgsw_words + (size_t)c * A.br_len * 2 * N;
    W pa[E], pb[E];
#pragma unroll
    for (int e = 0; e < E; e++) pa[e] = pb[e] = 0;
    {
      W dst[E];
      rotate_sub_cn<F, E>(part, r, lane, dst);
      if constexpr (!F::IS_GOLDILOCKS) {
        // BabyBear's 4 digits: waves 0-3 one each, so that every SIMD runs one row
        const u32 len = (u32)wave < A.br_len ? 1u : 0u;
        const u32 j0 = len ? (u32)wave : 0u;
        u32 win[E];
#pragma unroll
        for (int e = 0; e < E; e++) win[e] = (u32)(dst[e] >> (j0 * A.br_bits));
        gadget_mac<F, E, PF>(pa, pb, win, len, A.br_bits, key + (size_t)j0 * 2 * N, lds, twl,
                             lane, wave >= 4);
      } else if constexpr (DIG == 8) {
        gadget_mac_bytes2<PF>(pa, pb, dst, (u32)wave, key + (size_t)wave * 2 * N, lds, twl, lane,
                              wave >= 4, 1);
      } else {
        // 13 digits over 8 waves: windows of 1 or 2 digits (1, 2, 1, 2, 2, 1, 2, 2), so that
        // the two waves of a SIMD (w, w + 4) take 3 or 4 rows
        const u32 j0 = (u32)wave * 13 / 8, j1 = (u32)(wave + 1) * 13 / 8;
        u32 win[E];
#pragma unroll
        for (int e = 0; e < E; e++) win[e] = (u32)(dst[e] >> (5 * j0));
        gadget_mac_window_bytes<5, PF>(pa, pb, win, j1 - j0, key + (size_t)j0 * 2 * N, lds, twl,
                                       lane, wave >= 4);
      }
    }
    // balanced partial sums (bootstrap_gate_split with 8 waves)
#pragma unroll
    for (int e = 0; e < E; e++) lds[e * 64 + lane] = pb[e];
    __syncthreads();
    {
      const W* take = lds + (ptrdiff_t)(wave < 4 ? 4 : -4) * ldsw;
      W* put = mine + (wave < 4 ? 0 : N);
#pragma unroll
      for (int e = 0; e < E; e++) put[e * 64 + lane] = DF<F>::add(pa[e], take[e * 64 + lane]);
    }
    __syncthreads();
    W x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = 4 * qw + k;
      W v = part[(size_t)ch * N + e * 64 + lane];
#pragma unroll
      for (int w = 1; w < 4; w++) v = DF<F>::add(v, part[(size_t)w * 2 * N + ch * N + e * 64 + lane]);
      x[k] = v;
    }
    const u64 want = ep | (u64)(ex + 1);
    const size_t par = ex & 1;
    if (!fin) {
      W* to = xch + ((size_t)(c * 2 + par) * 4 + qw) * 256;
#pragma unroll
      for (int k = 0; k < 4; k++) pair_put(to + k * 64 + lane, x[k]);
      pair_signal(fl + c * 4 + qw, want);
    } else {
      alive = alive && pair_wait(fl + (1 - c) * 4 + qw, want, fl + 10, ep, A);
      if (alive) {
        const W* from = xch + ((size_t)((1 - c) * 2 + par) * 4 + qw) * 256;
#pragma unroll
        for (int k = 0; k < 4; k++) x[k] = DF<F>::add(x[k], pair_get(from + k * 64 + lane));
      }
      ntt_inv_quarter<F>(x, lds, itwl, lane, qw);
    }
    __syncthreads();
    if (fin) {
      ntt_inv_cross<F>(x, lds - (ptrdiff_t)qw * ldsw, ldsw, itwl, lane, qw);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int pos = 64 * qw + lane + 256 * k;
        raw[k] = F::add_cy(raw[k], x[k]);
        part[pos] = F::canon(raw[k]);
        part[N + pos] = F::canon(F::neg_raw(raw[k]));
      }
    }
    __syncthreads();
    ex++;
  }
  // key switch in workgroup 0 with the b component's word 0 from workgroup 1 (C[0] serves
  // for the raw word: keyswitch_extract canonicalises b0)
  if (wave == 0) {
    if (c == 1) {
      pair_put(fl + 8, (u64)part[0]);
      pair_signal(fl + 9, ep | 1);
    } else {
      W b0 = 0;
      if (alive && pair_wait(fl + 9, ep | 1, fl + 10, ep, A)) b0 = (W)pair_get(fl + 8);
      part[2 * N + lane] = b0;
      wave_sync();
      keyswitch_extract<F, PF, true>(A, g, part + N, part + 2 * N, lds, twl, itwl, lane);
    }
  }
}

// Copies the forward and inverse twiddle tables (N words each; the Goldilocks forward table
// of the digit transforms, ntt_fwd_digits, has kTwxWords) into LDS for the per-lane reads of
// NTT passes 2 and 3. twl holds kTwxWords words.
template <class W>
__device__ __forceinline__ void stage_twiddles(const void* tw, const void* itw, W* twl, W* itwl,
                                               int n, int tid, int nthreads) {
  const W* a = static_cast<const W*>(tw);
  const W* b = static_cast<const W*>(itw);
  for (int k = tid; k < n; k += nthreads) {
    twl[k] = a[k];
    itwl[k] = b[k];
  }
  // the Goldilocks forward table of the digit transforms has kTwxWords words
  if (sizeof(W) == 8)
    for (int k = n + tid; k < kTwxWords; k += nthreads) twl[k] = a[k];
  __syncthreads();
}

// WPS: waves per SIMD the register allocation targets (1: up to 512 VGPR+AGPR per lane;
// 2: at most 256, 3: at most 168, the excess spilled, so that the waves of a SIMD hide each
// other's LDS and memory latency; 3 suits BabyBear's 32-bit words).
template <class F, int GPW, int WPS, int DIG>
__global__ __launch_bounds__(64 * GPW) __attribute__((amdgpu_waves_per_eu(WPS, WPS)))
void bootstrap_kernel(BootArgs A) {
  using W = typename DF<F>::W;
  __shared__ W lds[GPW][1024 + 64];
  __shared__ unsigned short ms[GPW][1024];
  __shared__ W twl[kTwxWords];
  __shared__ W itwl[1024];
  stage_twiddles<W>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * GPW);
  // The wave index as a scalar (wave_id) for Zama's 5-bit digits only: 4096 gates at (4, 2)
  // 310.6 -> 305.9 ms, while Goldilocks with byte digits (201.9 -> 212.9 ms) and BabyBear
  // (56.8 -> 69.9 ms) ran slower with it (same box, its spill 39 -> 2 VGPRs notwithstanding).
  const int w = DIG == 5 ? wave_id() : (int)(threadIdx.x >> 6);
  const u32 g = blockIdx.x * GPW + w;
  if (g >= A.count) return;  // after the only workgroup barrier
  // Twiddles a stage ahead with one wave per SIMD, and for BabyBear, whose 32-bit words
  // leave the registers for it (4096 gates at (4,2): 55.5 -> 54.9 ms).
  bootstrap_gate<F, WPS == 1 || !F::IS_GOLDILOCKS, DIG>(A, (int)g, lds[w], ms[w], twl, itwl);
}

// Latency-mode launch: one gate per workgroup of WAVES wavefronts (bootstrap_gate_split).
template <class F, int WAVES, int DIG, bool TR>
__global__ __launch_bounds__(64 * WAVES) void bootstrap_split_kernel(BootArgs A) {
  using W = typename DF<F>::W;
  __shared__ W lds[WAVES][1024 + 64];
  __shared__ unsigned short ms[1024];
  __shared__ W twl[kTwxWords];
  __shared__ W itwl[1024];
  __shared__ W part[4 * 2 * 1024];
  __shared__ W accl[WAVES == 4 ? 2 * 1024 : 1];
  stage_twiddles<W>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * WAVES);
  W* const aa = WAVES == 4 ? accl : part;
  W* const ab = WAVES == 4 ? accl + 1024 : part + 2 * 1024;
  bootstrap_gate_split<F, WAVES, DIG, TR>(A, blockIdx.x, lds[threadIdx.x >> 6], 1024 + 64, ms,
                                       part, aa, ab, twl, itwl);
}

// Pair-mode launch of 2 * count workgroups: tickets 2g and 2g + 1 carry gate g
// (bootstrap_gate_pair).
template <class F, int DIG>
__global__ __launch_bounds__(64 * 8) void bootstrap_pair_kernel(BootArgs A) {
  using W = typename DF<F>::W;
  __shared__ W lds[8][1024 + 64];
  __shared__ unsigned short ms[1024];
  __shared__ W twl[kTwxWords];
  __shared__ W itwl[1024];
  __shared__ W part[4 * 2 * 1024];
  __shared__ u32 tk;
  u32 t;
  if (!pair_ticket(A.pair_status, A.pair_ticket0, 2 * A.count, A.pair_epoch, &tk, &t)) return;
  stage_twiddles<W>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * 8);
  bootstrap_gate_pair<F, DIG>(A, (int)(t >> 1), (int)(t & 1), lds[threadIdx.x >> 6], 1024 + 64,
                              ms, part, twl, itwl);
}

// Pair-mode mixed launch (the serving loop): tickets 2b and 2b + 1 (from the counter and
// base of part 0, which every part shares) carry gate b of the launch, which belongs to the
// last part whose first_block <= b (every part BabyBear with
// at most 4 digits or Goldilocks with a byte-digit gadget, untraced, with its own
// pair_flags rows). CLS: 1 every part Goldilocks with 8-bit digits, 2 every part with
// 5-bit digits, 3 every part BabyBear, 0 any of them (runtime branch).
template <int CLS>
__global__ __launch_bounds__(64 * 8) void mixed_pair_kernel(const BootArgs* __restrict__ parts,
                                                            int nparts) {
  __shared__ u64 lds[8][1024 + 64];
  __shared__ unsigned short ms[1024];
  __shared__ u64 twl[kTwxWords];
  __shared__ u64 itwl[1024];
  __shared__ u64 part[4 * 2 * 1024];
  __shared__ u32 tk;
  const BootArgs& L = parts[nparts - 1];  // the launch's gates end with the last part's
  u32 t;
  if (!pair_ticket(parts[0].pair_status, parts[0].pair_ticket0, 2 * (L.first_block + L.count),
                   parts[0].pair_epoch, &tk, &t))
    return;
  const u32 gb = t >> 1;
  int p = 0;
  while (p + 1 < nparts && gb >= parts[p + 1].first_block) p++;
  const BootArgs A = parts[p];
  const int g = (int)(gb - A.first_block), c = (int)(t & 1), w = threadIdx.x >> 6;
  if (CLS == 3 || (CLS == 0 && !A.is_goldilocks)) {
    if constexpr (CLS == 0 || CLS == 3) {
      // BabyBear: the same LDS arrays in 32-bit words
      u32* t = reinterpret_cast<u32*>(twl);
      u32* it = reinterpret_cast<u32*>(itwl);
      stage_twiddles<u32>(A.tw_fwd, A.tw_inv, t, it, 1024, threadIdx.x, 64 * 8);
      // the row stride in 32-bit words: twice the u64 row
      bootstrap_gate_pair<Bb, 0>(A, g, c, reinterpret_cast<u32*>(lds[w]), 2 * (1024 + 64), ms,
                                 reinterpret_cast<u32*>(part), t, it);
    }
    return;
  }
  if constexpr (CLS != 3) {
    stage_twiddles<u64>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * 8);
    if (CLS == 1 || (CLS == 0 && dig_bits(A) == 8))
      bootstrap_gate_pair<Gl, 8>(A, g, c, lds[w], 1024 + 64, ms, part, twl, itwl);
    else
      bootstrap_gate_pair<Gl, 5>(A, g, c, lds[w], 1024 + 64, ms, part, twl, itwl);
  }
}

// Mixed-parameter launch: workgroup b belongs to the last part whose first_block <= b.
// Each part carries its own keys, twiddles, gate kind and I/O pointers, so gates of
// BabyBear, Goldilocks and Zama-Goldilocks run side by side in one grid.
__global__ __launch_bounds__(64) void mixed_kernel(const BootArgs* __restrict__ parts, int nparts) {
  __shared__ u64 lds[1024 + 64];
  __shared__ unsigned short ms[1024];
  __shared__ u64 twl[kTwxWords];
  __shared__ u64 itwl[1024];
  const u32 b = blockIdx.x;
  int p = 0;
  while (p + 1 < nparts && b >= parts[p + 1].first_block) p++;
  const BootArgs A = parts[p];
  const int g = (int)(b - A.first_block);
  if (A.is_goldilocks) {
    stage_twiddles<u64>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64);
    const int db = dig_bits(A);
    if (db == 8)
      bootstrap_gate<Gl, true, 8>(A, g, lds, ms, twl, itwl);
    else if (db == 5)
      bootstrap_gate<Gl, true, 5>(A, g, lds, ms, twl, itwl);
    else
      bootstrap_gate<Gl, true, 0>(A, g, lds, ms, twl, itwl);
  } else {
    u32* t = reinterpret_cast<u32*>(twl);
    u32* it = reinterpret_cast<u32*>(itwl);
    stage_twiddles<u32>(A.tw_fwd, A.tw_inv, t, it, 1024, threadIdx.x, 64);
    bootstrap_gate<Bb>(A, g, reinterpret_cast<u32*>(lds), ms, t, it);
  }
}

// Mixed-parameter launch in the batched kernel's throughput shape: GPW gates of one part
// per workgroup, sharing its twiddle tables, with registers allotted for WPS waves per
// SIMD. Here first_block counts workgroups, so every part starts on a workgroup boundary;
// the waves past a part's count leave after the twiddle barrier.
// CLS restricts the parts to one class so the kernel holds one body and allocates registers
// as the dedicated batched kernel does: 1 Goldilocks with 8-bit digits, 2 the other
// Goldilocks sets, 3 BabyBear; 0 takes any mix (runtime branch).
template <int GPW, int WPS, int CLS>
__global__ __launch_bounds__(64 * GPW) __attribute__((amdgpu_waves_per_eu(WPS, WPS)))
void mixed_batch_kernel(const BootArgs* __restrict__ parts, int nparts) {
  __shared__ u64 lds[GPW][1024 + 64];
  __shared__ unsigned short ms[GPW][1024];
  __shared__ u64 twl[kTwxWords];
  __shared__ u64 itwl[1024];
  const u32 b = blockIdx.x;
  int p = 0;
  while (p + 1 < nparts && b >= parts[p + 1].first_block) p++;
  const BootArgs A = parts[p];
  const int w = threadIdx.x >> 6;
  const u32 g = (b - A.first_block) * GPW + w;
  if (CLS == 1 || CLS == 2 || (CLS == 0 && A.is_goldilocks)) {
    stage_twiddles<u64>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * GPW);
    if (g >= A.count) return;
    if constexpr (CLS == 1)
      bootstrap_gate<Gl, WPS == 1, 8>(A, (int)g, lds[w], ms[w], twl, itwl);
    else if constexpr (CLS == 2)
      bootstrap_gate<Gl, WPS == 1, 5>(A, (int)g, lds[w], ms[w], twl, itwl);
    else if (dig_bits(A) == 8)  // byte-digit gadgets only, as in mixed_split_kernel
      bootstrap_gate<Gl, WPS == 1, 8>(A, (int)g, lds[w], ms[w], twl, itwl);
    else
      bootstrap_gate<Gl, WPS == 1, 5>(A, (int)g, lds[w], ms[w], twl, itwl);
  } else if constexpr (CLS == 0 || CLS == 3) {
    u32* t = reinterpret_cast<u32*>(twl);
    u32* it = reinterpret_cast<u32*>(itwl);
    stage_twiddles<u32>(A.tw_fwd, A.tw_inv, t, it, 1024, threadIdx.x, 64 * GPW);
    if (g >= A.count) return;
    bootstrap_gate<Bb>(A, (int)g, reinterpret_cast<u32*>(lds[w]), ms[w], t, it);
  }
}

// Latency-mode mixed launch: one gate per workgroup of WAVES wavefronts, parts as in
// mixed_kernel; TR (8 waves): some parts are traced.
template <int WAVES, bool TR>
__global__ __launch_bounds__(64 * WAVES) void mixed_split_kernel(
    const BootArgs* __restrict__ parts, int nparts) {
  __shared__ u64 lds[WAVES][1024 + 64];
  __shared__ unsigned short ms[1024];
  __shared__ u64 twl[kTwxWords];
  __shared__ u64 itwl[1024];
  __shared__ u64 part[4 * 2 * 1024];
  __shared__ u64 accl[WAVES == 4 ? 2 * 1024 : 1];
  const u32 b = blockIdx.x;
  int p = 0;
  while (p + 1 < nparts && b >= parts[p + 1].first_block) p++;
  const BootArgs A = parts[p];
  const int g = (int)(b - A.first_block);
  const int w = threadIdx.x >> 6;
  // the accumulator's LDS words (bootstrap_gate_split), in the set's word size
  u64* const aa = WAVES == 4 ? accl : part;
  if (A.is_goldilocks) {
    stage_twiddles<u64>(A.tw_fwd, A.tw_inv, twl, itwl, 1024, threadIdx.x, 64 * WAVES);
    u64* const ab = WAVES == 4 ? aa + 1024 : aa + 2 * 1024;
    // 8 waves: the Goldilocks parts have byte-digit gadgets (mixed_kernel_for sends any
    // other to the one-wave kernel), so the kernel holds two Goldilocks bodies, not three,
    // and allocates registers like the dedicated kernels
    if (WAVES == 8 && dig_bits(A) == 8)
      bootstrap_gate_split<Gl, WAVES, WAVES == 8 ? 8 : 0, TR>(A, g, lds[w], 1024 + 64, ms, part,
                                                              aa, ab, twl, itwl);
    else if (WAVES == 8)
      bootstrap_gate_split<Gl, WAVES, WAVES == 8 ? 5 : 0, TR>(A, g, lds[w], 1024 + 64, ms, part,
                                                              aa, ab, twl, itwl);
    else
      bootstrap_gate_split<Gl, WAVES>(A, g, lds[w], 1024 + 64, ms, part, aa, ab, twl, itwl);
  } else {
    u32* t = reinterpret_cast<u32*>(twl);
    u32* it = reinterpret_cast<u32*>(itwl);
    stage_twiddles<u32>(A.tw_fwd, A.tw_inv, t, it, 1024, threadIdx.x, 64 * WAVES);
    u32* const ba = reinterpret_cast<u32*>(aa);
    u32* const bb = WAVES == 4 ? ba + 1024 : ba + 2 * 1024;
    bootstrap_gate_split<Bb, WAVES, 0, TR>(A, g, reinterpret_cast<u32*>(lds[w]), 2 * (1024 + 64),
                                           ms, reinterpret_cast<u32*>(part), ba, bb, t, it);
  }
}

// ----------------------------------------------------------------------------
// Microbench: negacyclic NTT over `count` polynomials. Forward: natural order in, slot
// order out. Inverse: slot order in, natural order out, times 1/N. Canonical u64 words in
// and out. Persistent waves, WAVES per workgroup sharing one LDS twiddle table: wave w of
// the grid transforms polynomials w, w + W, w + 2W, ... (W = waves in the grid), and loads
// the next one into registers before transforming the current one, so the HBM reads of
// one polynomial overlap the arithmetic of the previous (the Goldilocks kernels hold 256
// VGPRs, two waves per SIMD, too few for wave switching alone to hide the loads).
// HALF (N = 2048): exchanges through half-size buffers (ntt_fwd_wave). PIPE: the
// next-polynomial prefetch (off for N = 2048, whose 64 extra VGPRs would halve the waves).
// ----------------------------------------------------------------------------
template <class F, int E, bool INV>
__device__ __forceinline__ void ntt_load(const u64* __restrict__ d, int lane, typename DF<F>::W (&x)[E]) {
  using W = typename DF<F>::W;
  if constexpr (!INV) {
#pragma unroll
    for (int e = 0; e < E; e++) x[e] = (W)d[lane + 64 * e];
  } else {
#pragma unroll
    for (int q = 0; q < E / 4; q++) {
      const ulonglong2* s = reinterpret_cast<const ulonglong2*>(d + 4 * (lane + 64 * q));
      const ulonglong2 a = s[0], b = s[1];
      x[4 * q] = (W)a.x;
      x[4 * q + 1] = (W)a.y;
      x[4 * q + 2] = (W)b.x;
      x[4 * q + 3] = (W)b.y;
    }
  }
}

template <class F, int E, int WAVES, bool HALF, bool PIPE, bool INV>
__global__ __launch_bounds__(64 * WAVES) void ntt_kernel(u64* __restrict__ data, size_t count,
                                                         const void* tw_, typename DF<F>::W ninv) {
  using W = typename DF<F>::W;
  constexpr int N = 64 * E;
  constexpr int BUF = HALF ? 1024 + 64 : N + N / 16;
  __shared__ W lds_all[WAVES][BUF];
  __shared__ W twl[N];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const W* tw = static_cast<const W*>(tw_);
  for (int k = threadIdx.x; k < N; k += 64 * WAVES) twl[k] = tw[k];
  __syncthreads();
  const size_t stride = (size_t)gridDim.x * WAVES;
  size_t poly = (size_t)blockIdx.x * WAVES + wave;
  if (poly >= count) return;
  W* lds = lds_all[wave];
  W x[E], nx[PIPE ? E : 1];
  ntt_load<F, E, INV>(data + poly * N, lane, x);
  for (; poly < count; poly += stride) {
    const bool more = poly + stride < count;  // wave-uniform
    if constexpr (PIPE) {
      if (more) ntt_load<F, E, INV>(data + (poly + stride) * N, lane, nx);
      sched_fence();
    }
    u64* d = data + poly * N;
    if constexpr (!INV) {
      ntt_fwd_wave<F, E, false, true, HALF>(x, lds, twl, lane);
#pragma unroll
      for (int q = 0; q < E / 4; q++) {
        ulonglong2* o = reinterpret_cast<ulonglong2*>(d + 4 * (lane + 64 * q));
        o[0] = make_ulonglong2((u64)DF<F>::canon(x[4 * q]), (u64)DF<F>::canon(x[4 * q + 1]));
        o[1] = make_ulonglong2((u64)DF<F>::canon(x[4 * q + 2]), (u64)DF<F>::canon(x[4 * q + 3]));
      }
    } else {
      ntt_inv_wave<F, E, true, HALF>(x, lds, twl, lane);
#pragma unroll
      for (int e = 0; e < E; e++) d[lane + 64 * e] = (u64)DF<F>::canon(DF<F>::mul(x[e], ninv));
    }
    if (more) {
      if constexpr (PIPE) {
#pragma unroll
        for (int e = 0; e < E; e++) x[e] = nx[e];
      } else {
        ntt_load<F, E, INV>(data + (poly + stride) * N, lane, x);
      }
    }
  }
}

// ----------------------------------------------------------------------------
// Microbench: external product of `count` RLWEs (a then b, 2N canonical words each)
// with one RGSW in the device layout. One wave per RLWE, 4 waves per workgroup.
// ----------------------------------------------------------------------------
template <class F>
__global__ __launch_bounds__(256) void extprod_kernel(u64* __restrict__ rlwe, size_t count,
                                                      const void* rgsw_, u32 len, u32 bits,
                                                      const void* tw_, const void* itw_) {
  using W = typename DF<F>::W;
  constexpr int E = 16, N = 1024;
  __shared__ W lds_all[4][N + N / 16];
  __shared__ W twl[kTwxWords];
  __shared__ W itwl[N];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  stage_twiddles<W>(tw_, itw_, twl, itwl, N, threadIdx.x, 256);
  const size_t idx = (size_t)blockIdx.x * 4 + wave;
  if (idx >= count) return;
  const W* rgsw = static_cast<const W*>(rgsw_);
  W* lds = lds_all[wave];
  u64* d = rlwe + idx * 2 * N;
  W pa[E], pb[E];
#pragma unroll
  for (int e = 0; e < E; e++) pa[e] = pb[e] = 0;
#pragma unroll 1
  for (int comp = 0; comp < 2; comp++) {
    W src[E];
#pragma unroll
    for (int e = 0; e < E; e++) src[e] = (W)d[comp * N + lane + 64 * e];
    gadget_mac<F, E>(pa, pb, src, len, bits, rgsw + (size_t)comp * len * 2 * N, lds, twl, lane);
  }
  ntt_inv_wave<F, E>(pa, lds, itwl, lane);
  ntt_inv_wave<F, E>(pb, lds, itwl, lane);
#pragma unroll
  for (int e = 0; e < E; e++) {
    d[lane + 64 * e] = (u64)pa[e];
    d[N + lane + 64 * e] = (u64)pb[e];
  }
}

// ----------------------------------------------------------------------------
// Microbench: peak rate of the Goldilocks multiply-accumulate of the MAC stage (per
// lane and gl_mac2: two Montgomery products and two modular adds) on register-resident
// operands, 4 independent accumulator pairs per lane. One word per lane is stored so
// nothing is elided.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(64) void modmul_peak_kernel(u64* __restrict__ out, u32 iters, u64 seed) {
  const u64 tid = (u64)blockIdx.x * 64 + threadIdx.x;
  u64 x[4], pa[4], pb[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    x[i] = Gl::canon(seed * (tid + 1) + 0x9E3779B97F4A7C15ull * (i + 1));
    pa[i] = Gl::canon(x[i] ^ 0x5555AAAA5555AAAAull);
    pb[i] = Gl::canon(x[i] * 3);
  }
  const u64 ka = Gl::canon(seed ^ 0x0123456789ABCDEFull), kb = Gl::canon(seed + 0xFEDCBA98ull);
  for (u32 it = 0; it < iters; it++) {
#pragma unroll
    for (int i = 0; i < 4; i++) gl_mac2(pa[i], pb[i], x[i], ka, kb);
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = pa[i];
  }
  out[tid] = pa[0] ^ pa[1] ^ pa[2] ^ pa[3] ^ pb[0] ^ pb[1] ^ pb[2] ^ pb[3];
}

// ----------------------------------------------------------------------------
// Evaluator::not (tfhe/src/bfhe/evaluate.rs:98-104) on `count` LWEs of n1 canonical words:
// out = -c, then out.b += Q/4, canonical out. One thread per word.
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void lwe_not_kernel(const u64* __restrict__ in,
                                                      u64* __restrict__ out, u32 count, u32 n1,
                                                      u64 q, int is_goldilocks) {
  const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= (size_t)count * n1) return;
  const u64 x = in[k];
  u64 v = x == 0 ? 0 : q - x;
  if (k % n1 == n1 - 1) {
    const u64 q4 = q >> 2;
    v = is_goldilocks ? Gl::canon(Gl::add(v, q4)) : (v + q4 >= q ? v + q4 - q : v + q4);
  }
  out[k] = v;
}

// ----------------------------------------------------------------------------
// Launchers
// ----------------------------------------------------------------------------
hipError_t launch_lwe_not(const u64* in, u64* out, u32 count, u32 n1, u64 q, int is_goldilocks,
                          hipStream_t s) {
  const size_t words = (size_t)count * n1;
  if (words == 0) return hipSuccess;
  hipLaunchKernelGGL(lwe_not_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s, in,
                     out, count, n1, q, is_goldilocks);
  return hipGetLastError();
}
template <class F, int DIG>
static hipError_t launch_bootstrap_f(const BootArgs& a, int gpw, int wps, hipStream_t s) {
  const dim3 g1(a.count), g2((a.count + 1) / 2), g4((a.count + 3) / 4);
  if (gpw == 1 && wps == 1)
    hipLaunchKernelGGL((bootstrap_kernel<F, 1, 1, DIG>), g1, dim3(64), 0, s, a);
  else if (gpw == 2 && wps == 1)
    hipLaunchKernelGGL((bootstrap_kernel<F, 2, 1, DIG>), g2, dim3(128), 0, s, a);
  else if (gpw == 4 && wps == 1)
    hipLaunchKernelGGL((bootstrap_kernel<F, 4, 1, DIG>), g4, dim3(256), 0, s, a);
  else if (gpw == 2 && wps == 2)
    hipLaunchKernelGGL((bootstrap_kernel<F, 2, 2, DIG>), g2, dim3(128), 0, s, a);
  else if (gpw == 4 && wps == 2)
    hipLaunchKernelGGL((bootstrap_kernel<F, 4, 2, DIG>), g4, dim3(256), 0, s, a);
  else if (gpw == 4 && wps == 3)
    hipLaunchKernelGGL((bootstrap_kernel<F, 4, 3, DIG>), g4, dim3(256), 0, s, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_bootstrap_pair(const BootArgs& a, hipStream_t s) {
  const int db = dig_bits(a);
  const bool bb = !a.is_goldilocks && a.br_len <= 4;  // one digit row per wave 0-3
  if ((db != 8 && db != 5 && !bb) || a.trace || a.idx || !a.pair_flags || !a.pair_status || a.count == 0)
    return hipErrorInvalidValue;
  const BootArgs& arg = a;
  const dim3 grid(2 * a.count), block(64 * 8);
  if (bb)
    hipLaunchKernelGGL((bootstrap_pair_kernel<Bb, 0>), grid, block, 0, s, arg);
  else if (db == 8)
    hipLaunchKernelGGL((bootstrap_pair_kernel<Gl, 8>), grid, block, 0, s, arg);
  else
    hipLaunchKernelGGL((bootstrap_pair_kernel<Gl, 5>), grid, block, 0, s, arg);
  return hipGetLastError();
}

template <bool TR>
static void launch_split8(const BootArgs& a, int is_goldilocks, hipStream_t s) {
  const int db = is_goldilocks ? dig_bits(a) : 0;
  if (db == 8)
    hipLaunchKernelGGL((bootstrap_split_kernel<Gl, 8, 8, TR>), dim3(a.count), dim3(512), 0, s, a);
  else if (db == 5)
    hipLaunchKernelGGL((bootstrap_split_kernel<Gl, 8, 5, TR>), dim3(a.count), dim3(512), 0, s, a);
  else if (is_goldilocks)
    hipLaunchKernelGGL((bootstrap_split_kernel<Gl, 8, 0, TR>), dim3(a.count), dim3(512), 0, s, a);
  else
    hipLaunchKernelGGL((bootstrap_split_kernel<Bb, 8, 0, TR>), dim3(a.count), dim3(512), 0, s, a);
}

hipError_t launch_bootstrap_split(const BootArgs& a, int is_goldilocks, int waves,
                                  hipStream_t s) {
  if (waves != 4 && waves != 8) return hipErrorInvalidValue;
  if (waves == 4) {
    if (a.trace) return hipErrorInvalidValue;  // traced gates need 8 waves
    if (is_goldilocks)
      hipLaunchKernelGGL((bootstrap_split_kernel<Gl, 4, 0, false>), dim3(a.count), dim3(256), 0,
                         s, a);
    else
      hipLaunchKernelGGL((bootstrap_split_kernel<Bb, 4, 0, false>), dim3(a.count), dim3(256), 0,
                         s, a);
  } else if (a.trace) {
    launch_split8<true>(a, is_goldilocks, s);
  } else {
    launch_split8<false>(a, is_goldilocks, s);
  }
  return hipGetLastError();
}

hipError_t launch_bootstrap(const BootArgs& a, int is_goldilocks, int gpw, int wps,
                            hipStream_t s) {
  if (!is_goldilocks) return launch_bootstrap_f<Bb, 0>(a, gpw, wps, s);
  switch (dig_bits(a)) {
    case 8: return launch_bootstrap_f<Gl, 8>(a, gpw, wps, s);
    case 5: return launch_bootstrap_f<Gl, 5>(a, gpw, wps, s);
    default: return launch_bootstrap_f<Gl, 0>(a, gpw, wps, s);
  }
}

hipError_t launch_mixed_split(const BootArgs* parts_dev, int nparts, u32 total_gates, int waves,
                              bool traced, hipStream_t s) {
  if (waves != 4 && waves != 8) return hipErrorInvalidValue;
  if (waves == 4) {
    if (traced) return hipErrorInvalidValue;  // traced gates need 8 waves
    hipLaunchKernelGGL((mixed_split_kernel<4, false>), dim3(total_gates), dim3(256), 0, s,
                       parts_dev, nparts);
  } else if (traced) {
    hipLaunchKernelGGL((mixed_split_kernel<8, true>), dim3(total_gates), dim3(512), 0, s,
                       parts_dev, nparts);
  } else {
    hipLaunchKernelGGL((mixed_split_kernel<8, false>), dim3(total_gates), dim3(512), 0, s,
                       parts_dev, nparts);
  }
  return hipGetLastError();
}

hipError_t launch_mixed_pair(const BootArgs* parts_dev, int nparts, u32 total_gates, int cls,
                             hipStream_t s) {
  if (total_gates == 0) return hipErrorInvalidValue;
  const dim3 grid(2 * total_gates), block(64 * 8);
  switch (cls) {
    case 1: hipLaunchKernelGGL(mixed_pair_kernel<1>, grid, block, 0, s, parts_dev, nparts); break;
    case 2: hipLaunchKernelGGL(mixed_pair_kernel<2>, grid, block, 0, s, parts_dev, nparts); break;
    case 3: hipLaunchKernelGGL(mixed_pair_kernel<3>, grid, block, 0, s, parts_dev, nparts); break;
    default: hipLaunchKernelGGL(mixed_pair_kernel<0>, grid, block, 0, s, parts_dev, nparts);
  }
  return hipGetLastError();
}

hipError_t launch_mixed(const BootArgs* parts_dev, int nparts, u32 total_gates, hipStream_t s) {
  hipLaunchKernelGGL(mixed_kernel, dim3(total_gates), dim3(64), 0, s, parts_dev, nparts);
  return hipGetLastError();
}

hipError_t launch_mixed_batch(const BootArgs* parts_dev, int nparts, u32 total_blocks, int cls,
                              hipStream_t s) {
  const dim3 grid(total_blocks), block(64 * kMixedBatchGpw);
  switch (cls) {
    case 1:
      hipLaunchKernelGGL((mixed_batch_kernel<kMixedBatchGpw, 2, 1>), grid, block, 0, s, parts_dev,
                         nparts);
      break;
    case 2:
      hipLaunchKernelGGL((mixed_batch_kernel<kMixedBatchGpw, 2, 2>), grid, block, 0, s, parts_dev,
                         nparts);
      break;
    case 3:
      hipLaunchKernelGGL((mixed_batch_kernel<kMixedBatchGpw, 2, 3>), grid, block, 0, s, parts_dev,
                         nparts);
      break;
    default:
      hipLaunchKernelGGL((mixed_batch_kernel<kMixedBatchGpw, 2, 0>), grid, block, 0, s, parts_dev,
                         nparts);
  }
  return hipGetLastError();
}

template <class F, int E, int WAVES, bool HALF>
static void launch_ntt_f(int inverse, u64* data, size_t count, const void* tw, u64 ninv,
                         u32 max_blocks, hipStream_t s) {
  using W = typename DF<F>::W;
  constexpr bool PIPE = false;
  const size_t need = (count + WAVES - 1) / WAVES;
  const dim3 grid((unsigned)(need < max_blocks ? need : max_blocks)), block(64 * WAVES);
  if (inverse)
    hipLaunchKernelGGL((ntt_kernel<F, E, WAVES, HALF, PIPE, true>), grid, block, 0, s, data,
                       count, tw, (W)ninv);
  else
    hipLaunchKernelGGL((ntt_kernel<F, E, WAVES, HALF, PIPE, false>), grid, block, 0, s, data,
                       count, tw, (W)ninv);
}

// Launch shapes measured on MI355X (4096 Goldilocks polynomials, kernel us, forward /
// inverse): N = 1024 one polynomial per wave, 4 waves per workgroup: 22.5 / 25.4 (8 persistent
// waves per workgroup with the next polynomial prefetched: 23.2 / 26.2, since at two waves
// per SIMD the VALU, not the loads, sets the pace; two 4-wave workgroups per CU, each wave
// two polynomials, the next one copied into a per-wave LDS stage by global_load_lds_dwordx4
// during the current transform, no registers held: 25.6 / 29.0 against 24.2 / 27.6 in the
// same run); N = 2048 with half-size exchange buffers,
// two 4-wave workgroups per CU: 45.0 / 53.6 (full buffers, one per CU: 62 / 73).
hipError_t launch_ntt(int is_goldilocks, int inverse, u32 log_n, u64* data, size_t count,
                      const void* tw, u64 ninv, u32 cus, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const u32 all = 0xFFFFFFFFu;
  if (is_goldilocks) {
    if (log_n == 10)
      launch_ntt_f<Gl, 16, 4, false>(inverse, data, count, tw, ninv, all, s);
    else if (log_n == 11)
      launch_ntt_f<Gl, 32, 4, true>(inverse, data, count, tw, ninv, 2 * cus, s);
    else
      return hipErrorInvalidValue;
  } else {
    if (log_n == 10)
      launch_ntt_f<Bb, 16, 4, false>(inverse, data, count, tw, ninv, all, s);
    else if (log_n == 11)
      launch_ntt_f<Bb, 32, 4, true>(inverse, data, count, tw, ninv, 6 * cus, s);
    else
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_modmul_peak(u64* out, u32 waves, u32 iters, u64 seed, hipStream_t s) {
  hipLaunchKernelGGL(modmul_peak_kernel, dim3(waves), dim3(64), 0, s, out, iters, seed);
  return hipGetLastError();
}

hipError_t launch_extprod(int is_goldilocks, u64* rlwe, size_t count, const void* rgsw, u32 len,
                          u32 bits, const void* tw, const void* itw, hipStream_t s) {
  dim3 grid((unsigned)((count + 3) / 4)), block(256);
  if (is_goldilocks)
    hipLaunchKernelGGL(extprod_kernel<Gl>, grid, block, 0, s, rlwe, count, rgsw, len, bits, tw, itw);
  else
    hipLaunchKernelGGL(extprod_kernel<Bb>, grid, block, 0, s, rlwe, count, rgsw, len, bits, tw, itw);
  return hipGetLastError();
}

}  // namespace hb
