// CPU oracle for TFHE NAND gate bootstrapping — TEST INFRASTRUCTURE ONLY.
// See oracle.hpp for scope, conventions and what pins each part.
#include "oracle.hpp"

#include <algorithm>
#include <atomic>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>

namespace orc {

// tfhe/src/bfhe/parameters.rs:45-99; vfhe/src/bfhe/parameters.rs:26-46 (ZAMA).
// decompose_len = ceil(MODULUS_BITS / bits): goldilocks/mod.rs:68-75, baby_bear/mod.rs:72-79.
bool get_params(int set, Params* p) {
  auto ceil_div = [](u32 a, u32 b) { return (a + b - 1) / b; };
  Params r{};
  r.set = set;
  r.n = 728;
  r.N = 1024;
  r.t = 4;
  switch (set) {
    case BABYBEAR:
      r.is_goldilocks = 0;
      r.br_bits = 8;
      r.ks_bits = 5;
      r.q = Bb::P;
      r.br_len = ceil_div(Bb::MODULUS_BITS, 8);
      r.ks_len = ceil_div(Bb::MODULUS_BITS, 5);
      break;
    case GOLDILOCKS:
      r.is_goldilocks = 1;
      r.br_bits = 8;
      r.ks_bits = 5;
      r.q = Gl::P;
      r.br_len = ceil_div(Gl::MODULUS_BITS, 8);
      r.ks_len = ceil_div(Gl::MODULUS_BITS, 5);
      break;
    case ZAMA:
      r.is_goldilocks = 1;
      r.br_bits = 5;
      r.ks_bits = 5;
      r.q = Gl::P;
      r.br_len = ceil_div(Gl::MODULUS_BITS, 5);
      r.ks_len = ceil_div(Gl::MODULUS_BITS, 5);
      break;
    default:
      return false;
  }
  *p = r;
  return true;
}

// ---------------------------------------------------------------------------
// Plain modular helpers used only to build tables (canonical values).
// ---------------------------------------------------------------------------
static u64 mulmod(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
static u64 powmod(u64 b, u64 e, u64 m) {
  u64 r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod(r, b, m);
    b = mulmod(b, b, m);
    e >>= 1;
  }
  return r;
}
static u32 bitrev(u32 x, u32 bits) {
  u32 r = 0;
  for (u32 i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// ---------------------------------------------------------------------------
// Negacyclic NTT in the documented convention (oracle.hpp). The reference calls
// concrete-ntt Plan::{fwd, inv, normalize} (algebra/src/transformation/concrete.rs:51-61,
// 138-148); this is a plain Cooley-Tukey / Gentleman-Sande restatement of the
// same transform.
// ---------------------------------------------------------------------------
template <class F>
struct NttTable {
  using W = typename F::word;
  u32 N = 0, logN = 0;
  u64 psi = 0;  // canonical
  std::vector<W> psi_rev, psi_inv_rev;
  W n_inv{};
};

template <class F>
static NttTable<F> make_table(u32 logN) {
  NttTable<F> T;
  T.logN = logN;
  T.N = 1u << logN;
  const u64 P = F::P;
  u64 psi = powmod(F::GENERATOR, (P - 1) / (2ull * T.N), P);
  T.psi = psi;
  u64 psi_inv = powmod(psi, P - 2, P);
  T.psi_rev.resize(T.N);
  T.psi_inv_rev.resize(T.N);
  for (u32 k = 0; k < T.N; k++) {
    u32 e = bitrev(k, logN);
    T.psi_rev[k] = F::from_value(powmod(psi, e, P));
    T.psi_inv_rev[k] = F::from_value(powmod(psi_inv, e, P));
  }
  T.n_inv = F::from_value(powmod(T.N, P - 2, P));
  return T;
}

template <class F>
static const NttTable<F>& table(u32 logN) {
  static std::mutex mu;
  static std::map<u32, std::unique_ptr<NttTable<F>>> cache;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find(logN);
  if (it == cache.end()) {
    it = cache.emplace(logN, std::make_unique<NttTable<F>>(make_table<F>(logN))).first;
  }
  return *it->second;
}

// One stage of butterflies of distance t over m blocks, block i with twiddle S[i]. The
// stages with t < 8 are instantiated with a compile-time distance so that the block body
// unrolls and the compiler can vectorise across blocks (same arithmetic, same words).
template <class F, u32 TT>
static inline void ct_stage(typename F::word* a, const typename F::word* S, u32 m, u32 t) {
  using W = typename F::word;
  const u32 tt = TT ? TT : t;
  for (u32 i = 0; i < m; i++) {
    const W s = S[i];
    W* x = a + 2 * i * tt;
    W* y = x + tt;
    for (u32 j = 0; j < tt; j++) {
      W U = x[j];
      W V = F::mul(y[j], s);
      x[j] = F::add(U, V);
      y[j] = F::sub(U, V);
    }
  }
}
template <class F, u32 TT>
static inline void gs_stage(typename F::word* a, const typename F::word* S, u32 h, u32 t) {
  using W = typename F::word;
  const u32 tt = TT ? TT : t;
  for (u32 i = 0; i < h; i++) {
    const W s = S[i];
    W* x = a + 2 * i * tt;
    W* y = x + tt;
    for (u32 j = 0; j < tt; j++) {
      W U = x[j], V = y[j];
      x[j] = F::add(U, V);
      y[j] = F::mul(F::sub(U, V), s);
    }
  }
}

template <class F>
static void ntt_fwd(typename F::word* a, const NttTable<F>& T) {
  const u32 N = T.N;
  u32 t = N;
  for (u32 m = 1; m < N; m <<= 1) {
    t >>= 1;
    const auto* S = T.psi_rev.data() + m;
    if (t == 1) ct_stage<F, 1>(a, S, m, t);
    else if (t == 2) ct_stage<F, 2>(a, S, m, t);
    else if (t == 4) ct_stage<F, 4>(a, S, m, t);
    else ct_stage<F, 0>(a, S, m, t);
  }
  for (u32 j = 0; j < N; j++) a[j] = F::canon(a[j]);
}

template <class F>
static void ntt_inv(typename F::word* a, const NttTable<F>& T) {
  const u32 N = T.N;
  u32 t = 1;
  for (u32 m = N; m > 1; m >>= 1) {
    u32 h = m >> 1;
    const auto* S = T.psi_inv_rev.data() + h;
    if (t == 1) gs_stage<F, 1>(a, S, h, t);
    else if (t == 2) gs_stage<F, 2>(a, S, h, t);
    else if (t == 4) gs_stage<F, 4>(a, S, h, t);
    else gs_stage<F, 0>(a, S, h, t);
    t <<= 1;
  }
  for (u32 j = 0; j < N; j++) a[j] = F::canon(F::mul(a[j], T.n_inv));
}

// Schoolbook negacyclic product c = a*b mod (X^N + 1), canonical in and out.
template <class F>
static void negacyclic_schoolbook(const typename F::word* a, const typename F::word* b,
                                  typename F::word* c, u32 N) {
  using W = typename F::word;
  std::vector<W> acc(N, F::from_value(0));
  for (u32 i = 0; i < N; i++) {
    if (F::value(a[i]) == 0) continue;
    for (u32 j = 0; j < N; j++) {
      W p = F::mul(a[i], b[j]);
      u32 k = i + j;
      if (k < N)
        acc[k] = F::add(acc[k], p);
      else
        acc[k - N] = F::sub(acc[k - N], p);
    }
  }
  for (u32 k = 0; k < N; k++) c[k] = F::canon(acc[k]);
}

// ---------------------------------------------------------------------------
// Decomposition: Polynomial::decompose_lsb_bits_inplace
// (algebra/src/polynomial/univariate/native_polynomial.rs:340-348) with
// decompose_lsb_bits_at (goldilocks/mod.rs:123-127, baby_bear/mod.rs:122-127).
// ---------------------------------------------------------------------------
template <class F>
static inline void decompose_lsb_bits_at(typename F::word& self, typename F::word& dst, u64 mask,
                                         u32 bits) {
  u64 value = F::value(self);
  dst = F::from_value(value & mask);
  self = F::from_value(value >> bits);
}

// ---------------------------------------------------------------------------
// Per-thread scratch (lattice/src/utils/space.rs counterparts).
// ---------------------------------------------------------------------------
template <class F>
struct Scratch {
  using W = typename F::word;
  std::vector<W> poly, digit, med_a, med_b, ext_a, ext_b, acc_a, acc_b, lut;
  explicit Scratch(u32 N)
      : poly(N), digit(N), med_a(N), med_b(N), ext_a(N), ext_b(N), acc_a(N), acc_b(N), lut(N) {}
};

// One gadget half: med += Σ_j NTT(digit_j(src)) ⊙ key_j, key_j = (a_j, b_j) of length N each.
// NTTGadgetRLWE::mul_polynomial_inplace_fast (lattice/src/gadget/rlwe.rs:396-419) and
// NTTRLWE::add_assign_gadget_rlwe_mul_polynomial_fast (lattice/src/rlwe.rs:1035-1055),
// each step via ntt_add_mul_assign_fast (ntt_polynomial.rs:620-629) = self + a*b
// (algebra/src/field/mod.rs:178-180).
template <class F>
static void gadget_mac(const typename F::word* src, const typename F::word* key, u32 len,
                       u32 bits, Scratch<F>& s, const NttTable<F>& T) {
  using W = typename F::word;
  const u32 N = T.N;
  const u64 mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
  std::copy(src, src + N, s.poly.begin());
  for (u32 j = 0; j < len; j++) {
    for (u32 k = 0; k < N; k++) decompose_lsb_bits_at<F>(s.poly[k], s.digit[k], mask, bits);
    ntt_fwd<F>(s.digit.data(), T);
    const W* ka = key + (size_t)(2 * j) * N;
    const W* kb = ka + N;
    for (u32 k = 0; k < N; k++) {
      s.med_a[k] = F::add(s.med_a[k], F::mul(ka[k], s.digit[k]));
      s.med_b[k] = F::add(s.med_b[k], F::mul(kb[k], s.digit[k]));
    }
  }
}

// RLWE::mul_assign_ntt_rgsw (lattice/src/rlwe.rs:573-596): in-place external product of
// (in_a, in_b) with one NTT-domain RGSW laid out as [c_neg_s_m, c_m][j < len][a, b][N].
template <class F>
static void external_product(typename F::word* io_a, typename F::word* io_b,
                             const typename F::word* rgsw, u32 len, u32 bits, Scratch<F>& s,
                             const NttTable<F>& T) {
  const u32 N = T.N;
  std::fill(s.med_a.begin(), s.med_a.end(), F::from_value(0));
  std::fill(s.med_b.begin(), s.med_b.end(), F::from_value(0));
  gadget_mac<F>(io_a, rgsw, len, bits, s, T);                        // c_neg_s_m × a
  gadget_mac<F>(io_b, rgsw + (size_t)2 * len * N, len, bits, s, T);  // c_m × b
  // NTTRLWE::inverse_transform_inplace (lattice/src/rlwe.rs:765-777)
  std::copy(s.med_a.begin(), s.med_a.end(), io_a);
  std::copy(s.med_b.begin(), s.med_b.end(), io_b);
  ntt_inv<F>(io_a, T);
  ntt_inv<F>(io_b, T);
}

// RLWE::mul_monic_monomial_sub_one_inplace (lattice/src/rlwe.rs:437-487): x = y*X^r - y.
template <class F>
static void monomial_sub_one(const typename F::word* y, typename F::word* x, u32 N, u32 r) {
  if (r <= N) {
    u32 nsr = N - r;
    for (u32 k = 0; k < r; k++) x[k] = F::neg(y[nsr + k]);
    for (u32 k = r; k < N; k++) x[k] = y[k - r];
  } else {
    u32 rr = r - N;
    u32 nsr = N - rr;
    for (u32 k = 0; k < rr; k++) x[k] = y[nsr + k];
    for (u32 k = rr; k < N; k++) x[k] = F::neg(y[k - rr]);
  }
  for (u32 k = 0; k < N; k++) x[k] = F::sub(x[k], y[k]);
}

// lwe_modulus_switch (fhe_core/src/modulus_switch.rs:12-38), IEEE f64, round half away.
static inline u32 mod_switch_one(u64 v, u64 q, u64 two_n) {
  double vf = (double)v;
  double r = std::round(vf * (double)two_n / (double)q);
  u64 t = (u64)r;
  return (u32)(t < two_n ? t : t - two_n);
}

// BinaryBlindRotationKey::blind_rotate (fhe_core/src/blind_rotation/rlwe/binary.rs:28-104).
// bsk: raw words, [i < n][c_neg_s_m, c_m][j < br_len][a, b][N]. lut: raw words.
template <class F>
static void blind_rotate(const Params& p, const typename F::word* bsk, const typename F::word* lut,
                         const u32* ms_a, u32 ms_b, Scratch<F>& s, const NttTable<F>& T) {
  const u32 N = p.N;
  std::copy(lut, lut + N, s.lut.begin());
  if (ms_b != 0) {  // binary.rs:40-62
    u32 neg_b = 2 * N - ms_b;
    if (neg_b <= N) {
      std::rotate(s.lut.begin(), s.lut.end() - neg_b, s.lut.end());
      for (u32 k = 0; k < neg_b; k++) s.lut[k] = F::neg(s.lut[k]);
    } else {
      u32 r = neg_b - N;
      std::rotate(s.lut.begin(), s.lut.end() - r, s.lut.end());
      for (u32 k = r; k < N; k++) s.lut[k] = F::neg(s.lut[k]);
    }
  }
  std::fill(s.acc_a.begin(), s.acc_a.end(), F::from_value(0));
  std::copy(s.lut.begin(), s.lut.end(), s.acc_b.begin());
  const size_t rgsw_words = (size_t)2 * p.br_len * 2 * N;
  for (u32 i = 0; i < p.n; i++) {
    u32 a_i = ms_a[i];
    if (a_i == 0) continue;  // binary.rs:70
    monomial_sub_one<F>(s.acc_a.data(), s.ext_a.data(), N, a_i);
    monomial_sub_one<F>(s.acc_b.data(), s.ext_b.data(), N, a_i);
    external_product<F>(s.ext_a.data(), s.ext_b.data(), bsk + i * rgsw_words, p.br_len, p.br_bits,
                        s, T);
    for (u32 k = 0; k < N; k++) {  // RLWE::add_assign_element_wise (rlwe.rs:287-290)
      s.acc_a[k] = F::add(s.acc_a[k], s.ext_a[k]);
      s.acc_b[k] = F::add(s.acc_b[k], s.ext_b[k]);
    }
  }
}

// KeySwitchingRLWEKey::key_switch_for_rlwe (fhe_core/src/key_switch/rlwe.rs:115-137) with
// key_switch_inner SubAMulS (:250-286) → NTTRLWE::sub_assign_gadget_rlwe_mul_polynomial_inplace_fast
// (lattice/src/rlwe.rs:1147-1166) → RLWE::from (rlwe.rs:78-86) → extract_partial_lwe_locally
// (rlwe.rs:393-402). Only the n.next_power_of_two() == N layout (one gadget, no chunk
// reversal, rlwe.rs:123-132 not taken), which covers all three parameter sets.
// ksk: raw words [j < ks_len][a, b][N]. out: n+1 raw words (a then b).
template <class F>
static void key_switch(const Params& p, const typename F::word* ksk, const typename F::word* acc_a,
                       typename F::word b0, typename F::word* out, Scratch<F>& s,
                       const NttTable<F>& T) {
  using W = typename F::word;
  const u32 N = p.N;
  std::fill(s.med_a.begin(), s.med_a.end(), F::from_value(0));
  std::fill(s.med_b.begin(), s.med_b.end(), b0);
  // polynomial.neg_assign() (rlwe.rs:1159): raw negation of the accumulator's a.
  std::vector<W> negated(N);
  for (u32 k = 0; k < N; k++) negated[k] = F::neg(acc_a[k]);
  gadget_mac<F>(negated.data(), ksk, p.ks_len, p.ks_bits, s, T);
  ntt_inv<F>(s.med_a.data(), T);
  ntt_inv<F>(s.med_b.data(), T);
  // extract_partial_lwe_locally: a[1..].reverse(); a[1..] = -a[1..]; truncate(n); b = b[0].
  std::vector<W> a(s.med_a.begin(), s.med_a.end());
  std::reverse(a.begin() + 1, a.end());
  for (u32 k = 1; k < N; k++) a[k] = F::neg(a[k]);
  for (u32 k = 0; k < p.n; k++) out[k] = a[k];
  out[p.n] = s.med_b[0];
}

// EvaluationKey::bootstrap (tfhe/src/bfhe/evaluate.rs:42-60). c: raw words (n+1).
template <class F>
static void bootstrap(const Params& p, const typename F::word* bsk, const typename F::word* ksk,
                      const typename F::word* c, const typename F::word* lut,
                      typename F::word* out, Scratch<F>& s, const NttTable<F>& T) {
  std::vector<u32> ms(p.n + 1);
  for (u32 k = 0; k <= p.n; k++) ms[k] = mod_switch_one(F::value(c[k]), p.q, 2ull * p.N);
  blind_rotate<F>(p, bsk, lut, ms.data(), ms[p.n], s, T);
  // acc.b_mut()[0] += Q::new(Q::MODULUS_VALUE >> 3)  (evaluate.rs:52)
  typename F::word b0 = F::add(s.acc_b[0], F::from_value(p.q >> 3));
  key_switch<F>(p, ksk, s.acc_a.data(), b0, out, s, T);
}

// ---------------------------------------------------------------------------
// Traced twin: vfhe::EvaluationKey::bootstrap (vfhe/src/bfhe/evaluate.rs:47-139) with
// blind_rotate_w_trace (fhe_core/src/blind_rotation/rlwe/binary.rs:106-204),
// mul_assign_ntt_rgsw_w_trace (lattice/src/rlwe.rs:604-632) and
// key_switch_for_rlwe_w_trace (fhe_core/src/key_switch/rlwe.rs:143-198, 290-350).
// Differences from the untraced path: the LUT rotation and (X^a_i)·ACC go through NTT
// products (exact, no raw-word negation), and every round runs (no a_i == 0 skip).
//
// Trace records hold canonical values. Round record (R = (22 + 12*br_len)*N + 2 words):
//   input_acc [a, b, NTT a, NTT b] 4N | output_acc 4N | monomial [poly, NTT] 2N |
//   monomial representation [degree, coefficient] 2 | monomial_times_acc 4N |
//   external_product_input 4N | 2*br_len hadamards [bit poly, bit NTT, rlwe NTT a, b,
//   rlwe poly a, b] 6N each | sum_prod [NTT a, NTT b, poly a, poly b] 4N.
// Tail record: ms input (n+1) | ms output (n+1) | ks decomposed poly N |
//   ks_len hadamards 6N each | ks sum_prod 4N | sample extraction input N, output N.
// (trace/src/acc_trace.rs:104-138, basic_ops/hadamard_trace.rs:329-394,
//  basic_ops/rlwe_trace.rs:146-218, modulus_switching_trace.rs:100-108, pbs_trace.rs:8-13.)
// ---------------------------------------------------------------------------
template <class F>
static void put_poly(u64* dst, const typename F::word* src, u32 N) {
  for (u32 k = 0; k < N; k++) dst[k] = F::value(src[k]);
}
template <class F>
static void put_poly_ntt(u64* dst, const typename F::word* src, u32 N, const NttTable<F>& T) {
  std::vector<typename F::word> t(src, src + N);
  ntt_fwd<F>(t.data(), T);
  put_poly<F>(dst, t.data(), N);
}
template <class F>
static void put_rlwe(u64* dst, const typename F::word* a, const typename F::word* b, u32 N,
                     const NttTable<F>& T) {
  put_poly<F>(dst, a, N);
  put_poly<F>(dst + N, b, N);
  put_poly_ntt<F>(dst + 2 * N, a, N, T);
  put_poly_ntt<F>(dst + 3 * N, b, N, T);
}

static size_t trace_round_words(const Params& p) {
  return (size_t)(22 + 12 * p.br_len) * p.N + 2;
}
static size_t trace_tail_words(const Params& p) {
  return (size_t)2 * (p.n + 1) + (size_t)p.N + (size_t)6 * p.ks_len * p.N + 4 * p.N + 2 * p.N;
}

// Traced gadget product: med += Σ_j NTT(digit_j(src)) ⊙ key_j, recording each hadamard
// (mul_polynomial_inplace_fast_w_traces, lattice/src/gadget/rlwe.rs:426-457).
template <class F>
static void gadget_mac_traced(const typename F::word* src, const typename F::word* key, u32 len,
                              u32 bits, Scratch<F>& s, const NttTable<F>& T, u64* had) {
  using W = typename F::word;
  const u32 N = T.N;
  const u64 mask = (1ull << bits) - 1;
  std::copy(src, src + N, s.poly.begin());
  for (u32 j = 0; j < len; j++) {
    for (u32 k = 0; k < N; k++) decompose_lsb_bits_at<F>(s.poly[k], s.digit[k], mask, bits);
    u64* h = had ? had + (size_t)6 * j * N : nullptr;
    if (h) put_poly<F>(h, s.digit.data(), N);
    ntt_fwd<F>(s.digit.data(), T);
    if (h) put_poly<F>(h + N, s.digit.data(), N);
    const W* ka = key + (size_t)(2 * j) * N;
    const W* kb = ka + N;
    for (u32 k = 0; k < N; k++) {
      s.med_a[k] = F::add(s.med_a[k], F::mul(ka[k], s.digit[k]));
      s.med_b[k] = F::add(s.med_b[k], F::mul(kb[k], s.digit[k]));
    }
    if (h) {  // HadamardTrace::append_rlwe_ntt: the key row and its inverse transform
      put_poly<F>(h + 2 * N, ka, N);
      put_poly<F>(h + 3 * N, kb, N);
      std::vector<W> ia(ka, ka + N), ib(kb, kb + N);
      ntt_inv<F>(ia.data(), T);
      ntt_inv<F>(ib.data(), T);
      put_poly<F>(h + 4 * N, ia.data(), N);
      put_poly<F>(h + 5 * N, ib.data(), N);
    }
  }
}

// rounds: sorted indices of the rounds to record into rec (one round record each).
// sink, when given, also receives every round's record (round index, record).
using RoundSink = std::function<void(u32, const u64*)>;
template <class F>
static void bootstrap_traced(const Params& p, const typename F::word* bsk,
                             const typename F::word* ksk, const typename F::word* c,
                             const typename F::word* lut, typename F::word* out,
                             const u32* rounds, u32 n_rounds, u64* rec, u64* tail, Scratch<F>& s,
                             const NttTable<F>& T, const RoundSink* sink = nullptr) {
  using W = typename F::word;
  const u32 N = p.N, n = p.n;
  const size_t R = trace_round_words(p);
  std::vector<u64> every(sink ? R : 0);
  // lwe_modulus_switch_w_trace (fhe_core/src/modulus_switch.rs:41-86)
  std::vector<u32> ms(n + 1);
  for (u32 k = 0; k <= n; k++) {
    ms[k] = mod_switch_one(F::value(c[k]), p.q, 2ull * N);
    if (tail) tail[k] = F::value(c[k]), tail[n + 1 + k] = ms[k];
  }
  // LUT * X^-b through NTT products (binary.rs:126-143)
  std::vector<W> l(lut, lut + N), m(N);
  if (ms[n] != 0) {
    u32 neg_b = 2 * N - ms[n];
    std::fill(m.begin(), m.end(), F::from_value(0));
    if (neg_b < N) m[neg_b] = F::from_value(1);
    else m[neg_b - N] = F::from_value(p.q - 1);
    ntt_fwd<F>(m.data(), T);
    ntt_fwd<F>(l.data(), T);
    for (u32 k = 0; k < N; k++) l[k] = F::mul(l[k], m[k]);
    ntt_inv<F>(l.data(), T);
  }
  std::fill(s.acc_a.begin(), s.acc_a.end(), F::from_value(0));
  std::copy(l.begin(), l.end(), s.acc_b.begin());
  const size_t rgsw_words = (size_t)2 * p.br_len * 2 * N;
  std::vector<W> na(N), nb(N), pa(N), pb(N);
  u32 ri = 0;
  for (u32 i = 0; i < n; i++) {
    u64* r = (ri < n_rounds && rounds[ri] == i) ? rec + (size_t)(ri++) * R : nullptr;
    if (!r && sink) r = every.data();
    if (r) put_rlwe<F>(r, s.acc_a.data(), s.acc_b.data(), N, T);
    // product = ACC * X^a_i via NTT (binary.rs:158-178)
    std::copy(s.acc_a.begin(), s.acc_a.end(), na.begin());
    std::copy(s.acc_b.begin(), s.acc_b.end(), nb.begin());
    ntt_fwd<F>(na.data(), T);
    ntt_fwd<F>(nb.data(), T);
    const u32 a_i = ms[i];
    std::fill(m.begin(), m.end(), F::from_value(0));
    u32 deg;
    u64 coef;
    if (a_i < N) deg = a_i, coef = 1;
    else deg = a_i - N, coef = p.q - 1;
    m[deg] = F::from_value(coef);
    if (r) {
      put_poly<F>(r + 8 * N, m.data(), N);
      r[10 * N] = deg;
      r[10 * N + 1] = coef;
    }
    ntt_fwd<F>(m.data(), T);
    if (r) put_poly<F>(r + 9 * N, m.data(), N);
    for (u32 k = 0; k < N; k++) pa[k] = F::mul(na[k], m[k]), pb[k] = F::mul(nb[k], m[k]);
    ntt_inv<F>(pa.data(), T);
    ntt_inv<F>(pb.data(), T);
    if (r) put_rlwe<F>(r + 10 * N + 2, pa.data(), pb.data(), N, T);
    // external_product.sub_assign_element_wise(&acc) (binary.rs:180)
    for (u32 k = 0; k < N; k++) {
      s.ext_a[k] = F::sub(pa[k], s.acc_a[k]);
      s.ext_b[k] = F::sub(pb[k], s.acc_b[k]);
    }
    if (r) put_rlwe<F>(r + 14 * N + 2, s.ext_a.data(), s.ext_b.data(), N, T);
    std::fill(s.med_a.begin(), s.med_a.end(), F::from_value(0));
    std::fill(s.med_b.begin(), s.med_b.end(), F::from_value(0));
    const W* key = bsk + i * rgsw_words;
    u64* had = r ? r + 18 * N + 2 : nullptr;
    gadget_mac_traced<F>(s.ext_a.data(), key, p.br_len, p.br_bits, s, T, had);
    gadget_mac_traced<F>(s.ext_b.data(), key + (size_t)2 * p.br_len * N, p.br_len, p.br_bits, s,
                         T, had ? had + (size_t)6 * p.br_len * N : nullptr);
    u64* sp = r ? r + (size_t)(18 + 12 * p.br_len) * N + 2 : nullptr;
    if (sp) put_poly<F>(sp, s.med_a.data(), N), put_poly<F>(sp + N, s.med_b.data(), N);
    std::copy(s.med_a.begin(), s.med_a.end(), s.ext_a.begin());
    std::copy(s.med_b.begin(), s.med_b.end(), s.ext_b.begin());
    ntt_inv<F>(s.ext_a.data(), T);
    ntt_inv<F>(s.ext_b.data(), T);
    if (sp) put_poly<F>(sp + 2 * N, s.ext_a.data(), N), put_poly<F>(sp + 3 * N, s.ext_b.data(), N);
    for (u32 k = 0; k < N; k++) {
      s.acc_a[k] = F::add(s.acc_a[k], s.ext_a[k]);
      s.acc_b[k] = F::add(s.acc_b[k], s.ext_b[k]);
    }
    if (r) put_rlwe<F>(r + 4 * N, s.acc_a.data(), s.acc_b.data(), N, T);
    if (sink) (*sink)(i, r);
  }
  // acc.b[0] += Q/8 (vfhe/src/bfhe/evaluate.rs:91) and the traced key switch.
  W b0 = F::add(s.acc_b[0], F::from_value(p.q >> 3));
  std::fill(s.med_a.begin(), s.med_a.end(), F::from_value(0));
  std::fill(s.med_b.begin(), s.med_b.end(), F::from_value(0));
  std::vector<W> negated(N);
  for (u32 k = 0; k < N; k++) negated[k] = F::neg(s.acc_a[k]);
  u64* t = tail ? tail + 2 * (n + 1) : nullptr;
  if (t) put_poly<F>(t, negated.data(), N);  // decomposed_polys[0].poly (its ntt stays zero)
  gadget_mac_traced<F>(negated.data(), ksk, p.ks_len, p.ks_bits, s, T, t ? t + N : nullptr);
  // sum_prod = init - init_temp: the MAC alone, since init was (0, b0) (rlwe.rs:335-341)
  u64* sp = t ? t + N + (size_t)6 * p.ks_len * N : nullptr;
  std::vector<W> sa(s.med_a), sb(s.med_b);
  if (sp) {
    put_poly<F>(sp, sa.data(), N);
    put_poly<F>(sp + N, sb.data(), N);
    ntt_inv<F>(sa.data(), T);
    ntt_inv<F>(sb.data(), T);
    put_poly<F>(sp + 2 * N, sa.data(), N);
    put_poly<F>(sp + 3 * N, sb.data(), N);
  }
  for (u32 k = 0; k < N; k++) s.med_b[k] = F::add(s.med_b[k], b0);
  ntt_inv<F>(s.med_a.data(), T);
  ntt_inv<F>(s.med_b.data(), T);
  // extract_partial_lwe_locally_w_trace (lattice/src/rlwe.rs:406-434)
  std::vector<W> a(s.med_a.begin(), s.med_a.end());
  if (sp) put_poly<F>(sp + 4 * N, a.data(), N);
  std::reverse(a.begin() + 1, a.end());
  for (u32 k = 1; k < N; k++) a[k] = F::neg(a[k]);
  if (sp) put_poly<F>(sp + 5 * N, a.data(), N);
  for (u32 k = 0; k < n; k++) out[k] = a[k];
  out[n] = s.med_b[0];
}

// LUT builders (tfhe/src/bfhe/evaluate.rs:267-465), lut_step = 1.
enum Lut { LUT_NAND = 0, LUT_AND_MAJ = 1, LUT_OR = 2, LUT_NOR = 3, LUT_XOR = 4, LUT_XNOR = 5 };
template <class F>
static void build_lut(const Params& p, int kind, typename F::word* v) {
  const u32 N = p.N;
  u64 q8 = p.q >> 3;
  typename F::word pos = F::from_value(q8), negv = F::from_value(p.q - q8);
  auto fill = [&](u32 mid, typename F::word lo, typename F::word hi) {
    for (u32 k = 0; k < mid; k++) v[k] = lo;
    for (u32 k = mid; k < N; k++) v[k] = hi;
  };
  switch (kind) {
    case LUT_NAND: fill((N >> 1) + (N >> 2), pos, negv); break;     // :267-284
    case LUT_AND_MAJ: fill((N >> 1) + (N >> 2), negv, pos); break;  // :287-304
    case LUT_OR: fill(N >> 2, negv, pos); break;                    // :334-352, :375-400
    case LUT_NOR: fill(N >> 2, pos, negv); break;                   // :355-372
    case LUT_XOR: fill(N >> 1, negv, pos); break;                   // :403-420, :443-465
    case LUT_XNOR: fill(N >> 1, pos, negv); break;                  // :423-440
    default: fill(N, F::from_value(0), F::from_value(0));
  }
}

// ---------------------------------------------------------------------------
// Evaluation-key handle: raw-word copies of the keys in the field's representation.
// ---------------------------------------------------------------------------
struct EkBase {
  Params p;
  virtual ~EkBase() {}
};
template <class F>
struct Ek : EkBase {
  using W = typename F::word;
  std::vector<W> bsk, ksk;
};

template <class F>
static EkBase* make_ek(const Params& p, const u64* bsk, const u64* ksk) {
  auto* e = new Ek<F>();
  e->p = p;
  size_t nb = (size_t)p.n * 2 * p.br_len * 2 * p.N;
  size_t nk = (size_t)p.ks_len * 2 * p.N;
  e->bsk.resize(nb);
  e->ksk.resize(nk);
  for (size_t i = 0; i < nb; i++) e->bsk[i] = F::from_value(bsk[i]);
  for (size_t i = 0; i < nk; i++) e->ksk[i] = F::from_value(ksk[i]);
  return e;
}

static u32 log2u(u32 x) {
  u32 r = 0;
  while ((1u << r) < x) r++;
  return r;
}

// Gate pre-processing (tfhe/src/bfhe/evaluate.rs:98-263) on raw words.
enum Gate { G_NAND = 0, G_AND = 1, G_OR = 2, G_NOR = 3, G_XOR = 4, G_XNOR = 5, G_MAJ = 6,
            G_MUX = 8 };
template <class F>
static int gate_combine(const Params& p, int gate, const typename F::word* c0,
                        const typename F::word* c1, const typename F::word* c2,
                        typename F::word* c, int* lut_kind) {
  const u32 n1 = p.n + 1;
  switch (gate) {
    case G_NAND: case G_AND: case G_OR: case G_NOR:  // add_component_wise_ref (lattice/src/lwe.rs:72-81)
      for (u32 k = 0; k < n1; k++) c[k] = F::add(c0[k], c1[k]);
      *lut_kind = gate == G_NAND ? LUT_NAND : gate == G_AND ? LUT_AND_MAJ : gate == G_OR ? LUT_OR : LUT_NOR;
      return 0;
    case G_XOR: case G_XNOR:  // sub_component_wise_ref + double_inplace (evaluate.rs:180-208)
      for (u32 k = 0; k < n1; k++) {
        typename F::word d = F::sub(c0[k], c1[k]);
        c[k] = F::add(d, d);
      }
      *lut_kind = gate == G_XOR ? LUT_XOR : LUT_XNOR;
      return 0;
    case G_MAJ:  // evaluate.rs:219-231
      if (!c2) return -1;
      for (u32 k = 0; k < n1; k++) c[k] = F::add(F::add(c0[k], c1[k]), c2[k]);
      *lut_kind = LUT_AND_MAJ;
      return 0;
  }
  return -1;
}

// Evaluator::not (tfhe/src/bfhe/evaluate.rs:98-104): LWE::neg (component-wise field neg),
// then b += Q/4 (field add).
template <class F>
static void lwe_not(const Params& p, const typename F::word* c, typename F::word* out) {
  const u32 n1 = p.n + 1;
  for (u32 k = 0; k < n1; k++) out[k] = F::neg(c[k]);
  out[p.n] = F::add(out[p.n], F::from_value(p.q >> 2));
}

template <class F>
static void run_gates(const Ek<F>& e, int gate, size_t count, const u64* c0s, const u64* c1s,
                      const u64* c2s, u64* outs, int threads) {
  const Params& p = e.p;
  const u32 n1 = p.n + 1;
  const NttTable<F>& T = table<F>(log2u(p.N));
  std::atomic<size_t> next{0};
  auto worker = [&]() {
    Scratch<F> s(p.N);
    std::vector<typename F::word> a(n1), b(n1), c2(n1), c(n1), lut(p.N), out(n1);
    for (;;) {
      size_t g = next.fetch_add(1);
      if (g >= count) break;
      for (u32 k = 0; k < n1; k++) {
        a[k] = F::from_value(c0s[g * n1 + k]);
        b[k] = F::from_value(c1s[g * n1 + k]);
        if (c2s) c2[k] = F::from_value(c2s[g * n1 + k]);
      }
      int kind = 0;
      if (gate == G_MUX) {
        // Evaluator::mux (evaluate.rs:245-263): t0 = and(c0, c1), t1 = and(not(c0), c2),
        // t0 += t1 (add_inplace_component_wise), then the OR-LUT bootstrap of t0.
        std::vector<typename F::word> nc0(n1), t0(n1), t1(n1);
        lwe_not<F>(p, a.data(), nc0.data());
        gate_combine<F>(p, G_AND, a.data(), b.data(), nullptr, c.data(), &kind);
        build_lut<F>(p, kind, lut.data());
        bootstrap<F>(p, e.bsk.data(), e.ksk.data(), c.data(), lut.data(), t0.data(), s, T);
        gate_combine<F>(p, G_AND, nc0.data(), c2.data(), nullptr, c.data(), &kind);
        bootstrap<F>(p, e.bsk.data(), e.ksk.data(), c.data(), lut.data(), t1.data(), s, T);
        for (u32 k = 0; k < n1; k++) c[k] = F::add(t0[k], t1[k]);
        build_lut<F>(p, LUT_OR, lut.data());
      } else {
        gate_combine<F>(p, gate, a.data(), b.data(), c2s ? c2.data() : nullptr, c.data(), &kind);
        build_lut<F>(p, kind, lut.data());
      }
      bootstrap<F>(p, e.bsk.data(), e.ksk.data(), c.data(), lut.data(), out.data(), s, T);
      for (u32 k = 0; k < n1; k++) outs[g * n1 + k] = F::value(out[k]);
    }
  };
  if (threads <= 1) {
    worker();
  } else {
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(worker);
    for (auto& th : pool) th.join();
  }
}

// ---------------------------------------------------------------------------
// Finalized PBSTrace containers, as the prover consumes them after
// ModulusSwitchingTrace::finalize(n + 1) and BlindRotationTrace::finalize(n)
// (vfhe/examples/zk_binary_pbs/zk_nand_goldilocks.rs:97-105). Restated from:
//   the append order of blind_rotate_w_trace (fhe_core/src/blind_rotation/rlwe/binary.rs:150-200)
//     into AccTrace (trace/src/acc_trace.rs:124-155) and SumHadamardTrace
//     (trace/src/basic_ops/hadamard_trace.rs:329-394);
//   AccTrace::finalize (trace/src/acc_trace.rs:157-202): input_acc_permuted is a clone of
//     output_acc whose poly planes' last row gets += initial - final (its ntt planes stay
//     output_acc's), then every container is zero padded to 2^log_num_round rows and
//     permutation_info = PermutationInfo::new_rotation_left(log_num_round, n, 1)
//     (trace/src/basic_ops/row_perm_trace.rs:41-52);
//   RLWETrace / PolynomialTrace / MonomialTrace::finalize (basic_ops/rlwe_trace.rs:103-109,
//     129-135, 184-192): padding only when the row count is not a power of two, by
//     (2^log_num - num) rows;
//   HadamardTrace / SumHadamardTrace::finalize (basic_ops/hadamard_trace.rs:339-342, 396-401);
//   ModulusSwitchingTrace::new / finalize (trace/src/modulus_switching_trace.rs:77-133).
// Words are canonical field values. Plane order follows each struct's field order:
//   RLWETrace [poly.0, poly.1, ntt.0, ntt.1], PolynomialTrace [poly, ntt],
//   MonomialTrace [degree, coefficient], HadamardTrace [bit.poly, bit.ntt, rlwe.poly.0,
//   rlwe.poly.1, rlwe.ntt.0, rlwe.ntt.1]; each plane is rows x N, row-major.
// ---------------------------------------------------------------------------
enum TracePart {
  TP_INITIAL_ACC = 0, TP_FINAL_ACC = 1, TP_INPUT_ACC = 2, TP_OUTPUT_ACC = 3,
  TP_INPUT_ACC_PERMUTED = 4, TP_MONOMIAL = 5, TP_MONOMIAL_REPR = 6, TP_MONOMIAL_TIMES_ACC = 7,
  TP_EXTERNAL_PRODUCT_INPUT = 8, TP_HADAMARD = 9, TP_SUM_PROD = 10, TP_PERMUTATION = 11,
  TP_MODULUS_SWITCH = 12, TP_COUNT = 13
};

static bool is_pow2(size_t x) { return x && !(x & (x - 1)); }

// One vector-of-rows container with the trace crate's append/finalize rules.
struct Planes {
  u32 width;                       // words per row (N, or 1 for MonomialTrace)
  u32 log_num;                     // rows after padding = 2^log_num
  std::vector<std::vector<u64>> v;  // one vector per plane
  Planes(u32 planes, u32 w, u32 lg) : width(w), log_num(lg), v(planes) {}
  void append(u32 plane, const u64* src) { v[plane].insert(v[plane].end(), src, src + width); }
  void finalize(size_t num) {
    if (is_pow2(num)) return;
    const size_t zeros = (((size_t)1 << log_num) - num) * width;
    for (auto& x : v) x.resize(x.size() + zeros, 0);
  }
  u64* write(u64* dst) const {
    for (const auto& x : v) dst = std::copy(x.begin(), x.end(), dst);
    return dst;
  }
};

static size_t pbs_part_words(const Params& p, int part) {
  const size_t N = p.N, R = (size_t)1 << log2u(p.n), M = (size_t)1 << log2u(p.n + 1);
  switch (part) {
    case TP_INITIAL_ACC: case TP_FINAL_ACC: return 4 * N;
    case TP_INPUT_ACC: case TP_OUTPUT_ACC: case TP_INPUT_ACC_PERMUTED:
    case TP_MONOMIAL_TIMES_ACC: case TP_EXTERNAL_PRODUCT_INPUT: case TP_SUM_PROD:
      return 4 * R * N;
    case TP_MONOMIAL: return 2 * R * N;
    case TP_MONOMIAL_REPR: return 2 * R;
    case TP_HADAMARD: return 6 * R * N;
    case TP_PERMUTATION: return R;
    case TP_MODULUS_SWITCH: return 4 * M + 3;
    default: return 0;
  }
}

// ModulusSwitchingTrace::new + finalize(n + 1) on the recorded input / output
// (modulus_switching_trace.rs:77-133). Field division by 2q and q is exact since 2q | Q-1.
static void ms_finalize(const Params& p, const u64* ms_in, const u64* ms_out, u64* out) {
  const u64 Q = p.q, q2 = 2ull * p.N;
  const u32 num = p.n + 1, lg = log2u(num);
  const u64 minus_one = Q - 1;
  const u64 k = mulmod(minus_one, powmod(2 * q2 % Q, Q - 2, Q), Q);  // (-1) / (q + q)
  const u64 two_k = mulmod(minus_one, powmod(q2 % Q, Q - 2, Q), Q);  // (-1) / q
  auto add = [Q](u64 a, u64 b) { return (u64)(((u128)a + b) % Q); };
  auto sub = [Q](u64 a, u64 b) { return a >= b ? a - b : (u64)((u128)a + Q - b); };
  const size_t M = (size_t)1 << lg;
  std::vector<u64> in(ms_in, ms_in + num), o(ms_out, ms_out + num), wit(num), helper(num);
  for (u32 i = 0; i < num; i++) {
    wit[i] = o[i] == 0 ? q2 : o[i];                        // b' in [1, q]
    const u64 two_b_minus_1 = sub(add(wit[i], wit[i]), 1);
    helper[i] = sub(sub(in[i], mulmod(two_b_minus_1, k, Q)), 1);  // e = a - (2b'-1)k - 1
  }
  in.resize(M, 0);
  o.resize(M, 0);
  wit.resize(M, q2);
  helper.resize(M, k);
  u64* d = out;
  for (const auto* x : {&in, &o, &wit, &helper}) d = std::copy(x->begin(), x->end(), d);
  d[0] = q2;
  d[1] = k;
  d[2] = two_k;
}

// Record offsets inside one traced round record (see the traced-bootstrap layout above).
template <class F>
static void pbs_trace_part(const Params& p, const typename F::word* bsk,
                           const typename F::word* ksk, const typename F::word* c,
                           const typename F::word* lut, int part, u32 index, u64* out,
                           Scratch<F>& s, const NttTable<F>& T) {
  const u32 N = p.N, n = p.n, L = p.br_len, lg = log2u(n);
  std::vector<typename F::word> o(n + 1);
  if (part == TP_PERMUTATION) {  // PermutationInfo::new_rotation_left(log_num_round, n, 1)
    std::vector<u64> table((size_t)1 << lg);
    for (size_t i = 0; i < table.size(); i++) table[i] = i;
    std::rotate(table.begin(), table.begin() + 1, table.begin() + n);
    std::copy(table.begin(), table.end(), out);
    return;
  }
  std::vector<u64> tail(part == TP_MODULUS_SWITCH ? trace_tail_words(p) : 0);
  if (part == TP_MODULUS_SWITCH) {
    bootstrap_traced<F>(p, bsk, ksk, c, lut, o.data(), nullptr, 0, nullptr, tail.data(), s, T);
    ms_finalize(p, tail.data(), tail.data() + n + 1, out);
    return;
  }
  const u32 planes = part == TP_HADAMARD ? 6 : part == TP_MONOMIAL || part == TP_MONOMIAL_REPR ? 2 : 4;
  Planes acc(planes, part == TP_MONOMIAL_REPR ? 1 : N, lg);
  Planes initial(4, N, 0), final_acc(4, N, 0);
  const size_t had_off = 18 * (size_t)N + 2, sp_off = (size_t)(18 + 12 * L) * N + 2;
  auto rlwe4 = [&](Planes& dst, const u64* r4) {
    for (u32 k = 0; k < 4; k++) dst.append(k, r4 + (size_t)k * N);
  };
  RoundSink sink = [&](u32 i, const u64* r) {
    if (i == 0) rlwe4(initial, r);              // append_acc_initial
    if (i + 1 == n) rlwe4(final_acc, r + 4 * N);  // append_acc_output
    switch (part) {
      case TP_INPUT_ACC: rlwe4(acc, r); break;
      case TP_OUTPUT_ACC: case TP_INPUT_ACC_PERMUTED: rlwe4(acc, r + 4 * N); break;
      case TP_MONOMIAL: acc.append(0, r + 8 * N), acc.append(1, r + 9 * N); break;
      case TP_MONOMIAL_REPR: acc.append(0, r + 10 * N), acc.append(1, r + 10 * N + 1); break;
      case TP_MONOMIAL_TIMES_ACC: rlwe4(acc, r + 10 * N + 2); break;
      case TP_EXTERNAL_PRODUCT_INPUT: rlwe4(acc, r + 14 * N + 2); break;
      case TP_HADAMARD: {  // record [bit, NTT bit, key NTT a, b, key poly a, b]
        const u64* h = r + had_off + (size_t)6 * index * N;
        const u32 order[6] = {0, 1, 4, 5, 2, 3};
        for (u32 k = 0; k < 6; k++) acc.append(k, h + (size_t)order[k] * N);
        break;
      }
      case TP_SUM_PROD: {  // record [NTT a, NTT b, poly a, poly b]
        const u64* sp = r + sp_off;
        const u32 order[4] = {2, 3, 0, 1};
        for (u32 k = 0; k < 4; k++) acc.append(k, sp + (size_t)order[k] * N);
        break;
      }
      default: break;
    }
  };
  bootstrap_traced<F>(p, bsk, ksk, c, lut, o.data(), nullptr, 0, nullptr, nullptr, s, T, &sink);
  if (part == TP_INPUT_ACC_PERMUTED) {  // last row of the poly planes += initial - final
    const size_t last = (size_t)(n - 1) * N;
    for (u32 h = 0; h < 2; h++)
      for (u32 k = 0; k < N; k++) {
        u64& a = acc.v[h][last + k];
        const typename F::word d =
            F::sub(F::from_value(initial.v[h][k]), F::from_value(final_acc.v[h][k]));
        a = F::value(F::add(F::from_value(a), d));
      }
  }
  if (part == TP_INITIAL_ACC) { initial.write(out); return; }   // finalize(1): no padding
  if (part == TP_FINAL_ACC) { final_acc.write(out); return; }
  acc.finalize(n);
  acc.write(out);
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C entry points (loaded with ctypes by tests/ and bench.py's CPU baseline).
// All words crossing this boundary are canonical u64 values.
// ---------------------------------------------------------------------------
extern "C" {

int orc_get_params(int set, Params* out) { return get_params(set, out) ? 0 : -1; }

// Number of Goldilocks negations of a raw word above P on the calling thread since the
// last reset.
u64 orc_raw_neg_events(int reset) {
  u64 v = Gl::raw_neg_events();
  if (reset) Gl::raw_neg_events() = 0;
  return v;
}

// Field ops on raw words, element-wise over arrays (for closed-form tests).
// op: 0 add, 1 sub, 2 neg, 3 mul, 4 value, 5 reduce128(hi:lo given as x=lo,y=hi)
int orc_gl_op(int op, const u64* x, const u64* y, u64* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    switch (op) {
      case 0: out[i] = Gl::add(x[i], y[i]); break;
      case 1: out[i] = Gl::sub(x[i], y[i]); break;
      case 2: out[i] = Gl::neg(x[i]); break;
      case 3: out[i] = Gl::mul(x[i], y[i]); break;
      case 4: out[i] = Gl::value(x[i]); break;
      case 5: out[i] = Gl::reduce128(((u128)y[i] << 64) | x[i]); break;
      default: return -1;
    }
  }
  return 0;
}
// BabyBear ops on canonical values (converted to/from Montgomery words internally).
int orc_bb_op(int op, const u64* x, const u64* y, u64* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    u32 a = Bb::from_value(x[i]), b = y ? Bb::from_value(y[i]) : 0;
    u32 r;
    switch (op) {
      case 0: r = Bb::add(a, b); break;
      case 1: r = Bb::sub(a, b); break;
      case 2: r = Bb::neg(a); break;
      case 3: r = Bb::mul(a, b); break;
      case 4: r = a; break;
      default: return -1;
    }
    out[i] = Bb::value(r);
  }
  return 0;
}

// NTT over `count` polynomials of length 2^logN; canonical words in and out.
int orc_ntt(int is_goldilocks, int inverse, u32 logN, u64* data, size_t count) {
  if (is_goldilocks) {
    const auto& T = table<Gl>(logN);
    for (size_t c = 0; c < count; c++) {
      u64* a = data + c * T.N;
      inverse ? ntt_inv<Gl>(a, T) : ntt_fwd<Gl>(a, T);
    }
  } else {
    const auto& T = table<Bb>(logN);
    std::vector<u32> w(T.N);
    for (size_t c = 0; c < count; c++) {
      u64* a = data + c * T.N;
      for (u32 k = 0; k < T.N; k++) w[k] = Bb::from_value(a[k]);
      inverse ? ntt_inv<Bb>(w.data(), T) : ntt_fwd<Bb>(w.data(), T);
      for (u32 k = 0; k < T.N; k++) a[k] = Bb::value(w[k]);
    }
  }
  return 0;
}

// psi (the primitive 2N-th root the convention uses), canonical.
u64 orc_ntt_psi(int is_goldilocks, u32 logN) {
  return is_goldilocks ? table<Gl>(logN).psi : table<Bb>(logN).psi;
}

int orc_negacyclic_schoolbook(int is_goldilocks, u32 N, const u64* a, const u64* b, u64* c) {
  if (is_goldilocks) {
    negacyclic_schoolbook<Gl>(a, b, c, N);
  } else {
    std::vector<u32> x(N), y(N), z(N);
    for (u32 k = 0; k < N; k++) x[k] = Bb::from_value(a[k]), y[k] = Bb::from_value(b[k]);
    negacyclic_schoolbook<Bb>(x.data(), y.data(), z.data(), N);
    for (u32 k = 0; k < N; k++) c[k] = Bb::value(z[k]);
  }
  return 0;
}

// Modulus switch of an (n+1)-word LWE to integers in [0, 2N).
int orc_mod_switch(const Params* p, const u64* c, u32* out) {
  for (u32 k = 0; k <= p->n; k++) out[k] = mod_switch_one(c[k], p->q, 2ull * p->N);
  return 0;
}

// (X^r - 1) * y with raw-word semantics; y and x are raw words (Goldilocks) or canonical
// values (BabyBear).
int orc_monomial_sub_one(int is_goldilocks, u32 N, u32 r, const u64* y, u64* x) {
  if (is_goldilocks) {
    monomial_sub_one<Gl>(y, x, N, r);
  } else {
    std::vector<u32> a(N), b(N);
    for (u32 k = 0; k < N; k++) a[k] = Bb::from_value(y[k]);
    monomial_sub_one<Bb>(a.data(), b.data(), N, r);
    for (u32 k = 0; k < N; k++) x[k] = Bb::value(b[k]);
  }
  return 0;
}

// One external product RLWE (io_a, io_b) × NTT-domain RGSW [2][len][2][N].
int orc_external_product(int is_goldilocks, u32 N, u32 bits, u32 len, const u64* rgsw, u64* io_a,
                         u64* io_b) {
  u32 logN = log2u(N);
  size_t words = (size_t)2 * len * 2 * N;
  if (is_goldilocks) {
    Scratch<Gl> s(N);
    external_product<Gl>(io_a, io_b, rgsw, len, bits, s, table<Gl>(logN));
  } else {
    Scratch<Bb> s(N);
    std::vector<u32> k(words), a(N), b(N);
    for (size_t i = 0; i < words; i++) k[i] = Bb::from_value(rgsw[i]);
    for (u32 i = 0; i < N; i++) a[i] = Bb::from_value(io_a[i]), b[i] = Bb::from_value(io_b[i]);
    external_product<Bb>(a.data(), b.data(), k.data(), len, bits, s, table<Bb>(logN));
    for (u32 i = 0; i < N; i++) io_a[i] = Bb::value(a[i]), io_b[i] = Bb::value(b[i]);
  }
  return 0;
}

void* orc_ek_create(const Params* p, const u64* bsk, const u64* ksk) {
  return p->is_goldilocks ? make_ek<Gl>(*p, bsk, ksk) : make_ek<Bb>(*p, bsk, ksk);
}
void orc_ek_destroy(void* e) { delete static_cast<EkBase*>(e); }

// Blind rotation only: ms = n+1 switched integers; lut canonical; acc_out = 2N raw words
// (a then b). Goldilocks words are returned raw (possibly in [P, 2^64)).
int orc_blind_rotate(void* ek, const u32* ms, const u64* lut, u64* acc_out) {
  EkBase* b = static_cast<EkBase*>(ek);
  const Params& p = b->p;
  const u32 N = p.N;
  if (p.is_goldilocks) {
    auto* e = static_cast<Ek<Gl>*>(b);
    Scratch<Gl> s(N);
    blind_rotate<Gl>(p, e->bsk.data(), lut, ms, ms[p.n], s, table<Gl>(log2u(N)));
    std::copy(s.acc_a.begin(), s.acc_a.end(), acc_out);
    std::copy(s.acc_b.begin(), s.acc_b.end(), acc_out + N);
  } else {
    auto* e = static_cast<Ek<Bb>*>(b);
    Scratch<Bb> s(N);
    std::vector<u32> l(N);
    for (u32 k = 0; k < N; k++) l[k] = Bb::from_value(lut[k]);
    blind_rotate<Bb>(p, e->bsk.data(), l.data(), ms, ms[p.n], s, table<Bb>(log2u(N)));
    for (u32 k = 0; k < N; k++) acc_out[k] = Bb::value(s.acc_a[k]), acc_out[N + k] = Bb::value(s.acc_b[k]);
  }
  return 0;
}

// Key switch + extraction from an accumulator (2N raw words) whose b[0] already
// carries the +Q/8 shift. out: n+1 canonical words.
int orc_key_switch(void* ek, const u64* acc, u64* out) {
  EkBase* b = static_cast<EkBase*>(ek);
  const Params& p = b->p;
  const u32 N = p.N;
  if (p.is_goldilocks) {
    auto* e = static_cast<Ek<Gl>*>(b);
    Scratch<Gl> s(N);
    std::vector<u64> o(p.n + 1);
    key_switch<Gl>(p, e->ksk.data(), acc, acc[N], o.data(), s, table<Gl>(log2u(N)));
    for (u32 k = 0; k <= p.n; k++) out[k] = Gl::value(o[k]);
  } else {
    auto* e = static_cast<Ek<Bb>*>(b);
    Scratch<Bb> s(N);
    std::vector<u32> a(N), o(p.n + 1);
    for (u32 k = 0; k < N; k++) a[k] = Bb::from_value(acc[k]);
    key_switch<Bb>(p, e->ksk.data(), a.data(), Bb::from_value(acc[N]), o.data(), s,
                   table<Bb>(log2u(N)));
    for (u32 k = 0; k <= p.n; k++) out[k] = Bb::value(o[k]);
  }
  return 0;
}

// Bootstrap with an arbitrary LUT (canonical words): c (n+1) → out (n+1).
int orc_bootstrap(void* ek, const u64* c, const u64* lut, u64* out) {
  EkBase* b = static_cast<EkBase*>(ek);
  const Params& p = b->p;
  const u32 N = p.N, n1 = p.n + 1;
  if (p.is_goldilocks) {
    auto* e = static_cast<Ek<Gl>*>(b);
    Scratch<Gl> s(N);
    std::vector<u64> o(n1);
    bootstrap<Gl>(p, e->bsk.data(), e->ksk.data(), c, lut, o.data(), s, table<Gl>(log2u(N)));
    for (u32 k = 0; k < n1; k++) out[k] = Gl::value(o[k]);
  } else {
    auto* e = static_cast<Ek<Bb>*>(b);
    Scratch<Bb> s(N);
    std::vector<u32> cc(n1), l(N), o(n1);
    for (u32 k = 0; k < n1; k++) cc[k] = Bb::from_value(c[k]);
    for (u32 k = 0; k < N; k++) l[k] = Bb::from_value(lut[k]);
    bootstrap<Bb>(p, e->bsk.data(), e->ksk.data(), cc.data(), l.data(), o.data(), s,
                  table<Bb>(log2u(N)));
    for (u32 k = 0; k < n1; k++) out[k] = Bb::value(o[k]);
  }
  return 0;
}

// Sizes (words) of one round record and of the tail record of the traced bootstrap.
int orc_trace_sizes(const Params* p, u64* round_words, u64* tail_words) {
  *round_words = trace_round_words(*p);
  *tail_words = trace_tail_words(*p);
  return 0;
}

// Traced bootstrap (the vfhe path) with an arbitrary LUT; records the listed rounds
// (sorted, < n) into rec and the modulus-switch / key-switch / extraction values into
// tail (either may be NULL). out: n+1 canonical words.
int orc_bootstrap_traced(void* ek, const u64* c, const u64* lut, u64* out, const u32* rounds,
                         u32 n_rounds, u64* rec, u64* tail) {
  EkBase* b = static_cast<EkBase*>(ek);
  const Params& p = b->p;
  const u32 N = p.N, n1 = p.n + 1;
  if (p.is_goldilocks) {
    auto* e = static_cast<Ek<Gl>*>(b);
    Scratch<Gl> s(N);
    std::vector<u64> o(n1);
    bootstrap_traced<Gl>(p, e->bsk.data(), e->ksk.data(), c, lut, o.data(), rounds, n_rounds, rec,
                         tail, s, table<Gl>(log2u(N)));
    for (u32 k = 0; k < n1; k++) out[k] = Gl::value(o[k]);
  } else {
    auto* e = static_cast<Ek<Bb>*>(b);
    Scratch<Bb> s(N);
    std::vector<u32> cc(n1), l(N), o(n1);
    for (u32 k = 0; k < n1; k++) cc[k] = Bb::from_value(c[k]);
    for (u32 k = 0; k < N; k++) l[k] = Bb::from_value(lut[k]);
    bootstrap_traced<Bb>(p, e->bsk.data(), e->ksk.data(), cc.data(), l.data(), o.data(), rounds,
                         n_rounds, rec, tail, s, table<Bb>(log2u(N)));
    for (u32 k = 0; k < n1; k++) out[k] = Bb::value(o[k]);
  }
  return 0;
}

// Words of one finalized PBSTrace container (part ids in the container comment above);
// 0 for an unknown part.
int orc_pbs_trace_part_words(const Params* p, int part, u64* words) {
  *words = pbs_part_words(*p, part);
  return *words ? 0 : -1;
}

// One finalized PBSTrace container of the traced bootstrap of c with lut. index selects
// the hadamard (0 <= index < 2*br_len: digit index % br_len of half index / br_len).
int orc_pbs_trace_part(void* ek, const u64* c, const u64* lut, int part, u32 index, u64* out) {
  EkBase* b = static_cast<EkBase*>(ek);
  const Params& p = b->p;
  const u32 N = p.N, n1 = p.n + 1;
  if (!pbs_part_words(p, part) || (part == TP_HADAMARD && index >= 2 * p.br_len)) return -1;
  if (p.is_goldilocks) {
    auto* e = static_cast<Ek<Gl>*>(b);
    Scratch<Gl> s(N);
    pbs_trace_part<Gl>(p, e->bsk.data(), e->ksk.data(), c, lut, part, index, out, s,
                       table<Gl>(log2u(N)));
  } else {
    auto* e = static_cast<Ek<Bb>*>(b);
    Scratch<Bb> s(N);
    std::vector<u32> cc(n1), l(N);
    for (u32 k = 0; k < n1; k++) cc[k] = Bb::from_value(c[k]);
    for (u32 k = 0; k < N; k++) l[k] = Bb::from_value(lut[k]);
    pbs_trace_part<Bb>(p, e->bsk.data(), e->ksk.data(), cc.data(), l.data(), part, index, out,
                       s, table<Bb>(log2u(N)));
  }
  return 0;
}

// Batched gates on a thread pool (one gate per task), like the reference's rayon
// caller pattern (tfhe/examples/bfhe.rs:139-148). gate: 0 nand, 1 and, 2 or, 3 nor,
// 4 xor, 5 xnor, 6 majority (c2s required).
int orc_gate_batch(void* ek, int gate, size_t count, const u64* c0s, const u64* c1s,
                   const u64* c2s, u64* outs, int threads) {
  EkBase* b = static_cast<EkBase*>(ek);
  if (b->p.is_goldilocks)
    run_gates<Gl>(*static_cast<Ek<Gl>*>(b), gate, count, c0s, c1s, c2s, outs, threads);
  else
    run_gates<Bb>(*static_cast<Ek<Bb>*>(b), gate, count, c0s, c1s, c2s, outs, threads);
  return 0;
}

// LUT for a gate kind (0 nand, 1 and/majority, 2 or, 3 nor, 4 xor, 5 xnor), canonical.
int orc_build_lut(const Params* p, int kind, u64* out) {
  if (p->is_goldilocks) {
    build_lut<Gl>(*p, kind, out);
  } else {
    std::vector<u32> v(p->N);
    build_lut<Bb>(*p, kind, v.data());
    for (u32 k = 0; k < p->N; k++) out[k] = Bb::value(v[k]);
  }
  return 0;
}

// encode/decode (fhe_core/src/plaintext.rs:64-94) and SecretKeyPack::decrypt_with_noise
// (fhe_core/src/secret_key.rs:161-188). lwe_sk: n canonical words.
int orc_decrypt_with_noise(const Params* p, const u64* lwe_sk, const u64* c, u64* msg,
                           u64* noise) {
  const u32 n = p->n;
  double tf = (double)p->t, qf = (double)p->q;
  auto encode = [&](u64 m) { return (u64)std::round((double)m * (qf / tf)); };
  auto decode = [&](u64 v) {
    u64 r = (u64)std::round((double)v * tf / qf);
    return r < p->t ? r : r - p->t;
  };
  if (p->is_goldilocks) {
    u64 acc = 0;  // goldilocks_ntt.rs:162-169: fold x*y + acc
    for (u32 k = 0; k < n; k++) acc = Gl::add(Gl::mul(c[k], lwe_sk[k]), acc);
    u64 plain = Gl::sub(c[n], acc);
    u64 m = decode(Gl::value(plain));
    u64 pp = Gl::from_value(encode(m));
    *msg = m;
    *noise = std::min(Gl::value(Gl::sub(plain, pp)), Gl::value(Gl::sub(pp, plain)));
  } else {
    u32 acc = 0;
    for (u32 k = 0; k < n; k++)
      acc = Bb::add(Bb::mul(Bb::from_value(c[k]), Bb::from_value(lwe_sk[k])), acc);
    u32 plain = Bb::sub(Bb::from_value(c[n]), acc);
    u64 m = decode(Bb::value(plain));
    u32 pp = Bb::from_value(encode(m));
    *msg = m;
    *noise = std::min(Bb::value(Bb::sub(plain, pp)), Bb::value(Bb::sub(pp, plain)));
  }
  return 0;
}

}  // extern "C"
