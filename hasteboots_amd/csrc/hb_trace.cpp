// Trace expansion for the traced bootstrap (the reference's vfhe path).
//
// The GPU emits a compact core trace per gate: the accumulator RLWE entering each of the
// n blind-rotation rounds plus the final one. Everything else the reference's trace
// crate records is a deterministic function of that core, the keys and the switched
// input, and is rebuilt here on demand:
//   AccTrace (trace/src/acc_trace.rs:104-138): input/output acc, monomial, its
//     representation, monomial_times_acc, external_product_input, each with its NTT;
//   SumHadamardTrace (trace/src/basic_ops/hadamard_trace.rs:329-394): per digit the bit
//     polynomial and its NTT, the key row and its inverse transform; sum_prod;
//   ModulusSwitchingTrace input/output (modulus_switching_trace.rs:100-108);
//   KeySwitchingTrace decomposed poly + hadamards + sum_prod (key_switch/rlwe.rs:290-350);
//   sample-extraction RowPermTrace input/output (lattice/src/rlwe.rs:406-434);
//   and, through hb_trace_part, the finalized containers the prover takes after
//   BlindRotationTrace::finalize(n) / ModulusSwitchingTrace::finalize(n + 1)
//   (trace/src/acc_trace.rs:157-202, basic_ops/rlwe_trace.rs:103-192,
//   basic_ops/hadamard_trace.rs:339-401, basic_ops/row_perm_trace.rs:41-52,
//   modulus_switching_trace.rs:77-133).
// Record layout (canonical words), per round:
//   input_acc [a, b, NTT a, NTT b] 4N | output_acc 4N | monomial [poly, NTT] 2N |
//   [degree, coefficient] 2 | monomial_times_acc 4N | external_product_input 4N |
//   2*br_len hadamards [bit, NTT bit, key a, key b, INTT key a, INTT key b] 6N each |
//   sum_prod [NTT a, NTT b, a, b] 4N.
// Tail: ms input n+1 | ms output n+1 | ks decomposed poly N | ks_len hadamards 6N each |
//   ks sum_prod 4N | sample extraction input N, output N.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "hb_field.hpp"
#include "hb_internal.hpp"

using namespace hb;

namespace {

uint32_t ilog2(uint32_t x) {
  uint32_t r = 0;
  while ((1u << r) < x) r++;
  return r;
}

struct Ctx {
  const ParamSet& p;
  HostNtt ntt;
  explicit Ctx(const ParamSet& ps) : p(ps) { ntt.init(ps, ilog2(ps.N)); }
  uint64_t canon(uint64_t x) const { return p.is_goldilocks ? Gl::canon(x) : x; }
  uint64_t neg(uint64_t x) const { return x == 0 ? 0 : p.q - x; }
  void put(uint64_t* dst, const uint64_t* src) const { std::memcpy(dst, src, p.N * 8); }
  void put_ntt(uint64_t* dst, const uint64_t* src) const {
    put(dst, src);
    ntt.fwd(dst);
  }
  void put_intt(uint64_t* dst, const uint64_t* src) const {
    put(dst, src);
    ntt.inv(dst);
  }
  void put_rlwe(uint64_t* dst, const uint64_t* a, const uint64_t* b) const {
    put(dst, a);
    put(dst + p.N, b);
    put_ntt(dst + 2 * p.N, a);
    put_ntt(dst + 3 * p.N, b);
  }
  // Σ_j NTT(digit_j(src)) ⊙ key_j into (ma, mb), recording hadamards (6N each) if had.
  void gadget(const uint64_t* src, const uint64_t* key, uint32_t len, uint32_t bits,
              uint64_t* ma, uint64_t* mb, uint64_t* had) const {
    const uint32_t N = p.N;
    std::vector<uint64_t> d(N);
    const uint64_t mask = (1ull << bits) - 1;
    for (uint32_t j = 0; j < len; j++) {
      for (uint32_t k = 0; k < N; k++) d[k] = (src[k] >> (j * bits)) & mask;
      uint64_t* h = had ? had + (size_t)6 * j * N : nullptr;
      if (h) put(h, d.data());
      ntt.fwd(d.data());
      if (h) put(h + N, d.data());
      const uint64_t* ka = key + (size_t)2 * j * N;
      const uint64_t* kb = ka + N;
      for (uint32_t k = 0; k < N; k++) {
        ma[k] = ntt.add(ma[k], ntt.mul(d[k], ka[k]));
        mb[k] = ntt.add(mb[k], ntt.mul(d[k], kb[k]));
      }
      if (h) {
        put(h + 2 * N, ka);
        put(h + 3 * N, kb);
        put_intt(h + 4 * N, ka);
        put_intt(h + 5 * N, kb);
      }
    }
  }
};

uint64_t combine_host(const ParamSet& p, int gate, uint64_t a, uint64_t b, uint64_t c) {
  HostNtt o;
  o.p = p;
  switch (gate) {
    case GATE_XOR:
    case GATE_XNOR: {
      uint64_t d = o.sub(a, b);
      return o.add(d, d);
    }
    case GATE_MAJORITY: return o.add(o.add(a, b), c);
    case GATE_BOOTSTRAP: return a;
    default: return o.add(a, b);
  }
}

// Combined gate input and its modulus switch (fhe_core/src/modulus_switch.rs:41-86): v the
// combined words, ms the switched values in [0, 2N).
void switch_input(const ParamSet& p, int gate, const uint64_t* c0, const uint64_t* c1,
                  const uint64_t* c2, uint64_t* v, uint32_t* ms) {
  const uint32_t N = p.N;
  for (uint32_t k = 0; k <= p.n; k++) {
    v[k] = combine_host(p, gate, c0[k], c1 ? c1[k] : 0, c2 ? c2[k] : 0);
    double r = std::round((double)v[k] * (double)(2 * N) / (double)p.q);
    uint64_t t = (uint64_t)r;
    ms[k] = (uint32_t)(t >= 2 * N ? t - 2 * N : t);
  }
}

int check_trace_args(const hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1,
                     const uint64_t* core) {
  if (!ek || !c0 || !core) return fail(HB_ERR_ARG, "null argument");
  if (gate < 0 || gate > GATE_BOOTSTRAP) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (gate != GATE_BOOTSTRAP && !c1) return fail(HB_ERR_ARG, "gate needs two inputs");
  return HB_OK;
}

// Words of one finalized container; 0 for an unknown part.
uint64_t part_words(const ParamSet& p, int part) {
  const uint64_t N = p.N, R = 1ull << ilog2(p.n), M = 1ull << ilog2(p.n + 1);
  switch (part) {
    case HB_TP_INITIAL_ACC: case HB_TP_FINAL_ACC: return 4 * N;
    case HB_TP_INPUT_ACC: case HB_TP_OUTPUT_ACC: case HB_TP_INPUT_ACC_PERMUTED:
    case HB_TP_MONOMIAL_TIMES_ACC: case HB_TP_EXTERNAL_PRODUCT_INPUT: case HB_TP_SUM_PROD:
      return 4 * R * N;
    case HB_TP_MONOMIAL: return 2 * R * N;
    case HB_TP_MONOMIAL_REPR: return 2 * R;
    case HB_TP_HADAMARD: return 6 * R * N;
    case HB_TP_PERMUTATION: return R;
    case HB_TP_MODULUS_SWITCH: return 4 * M + 3;
    default: return 0;
  }
}

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t m) { return (uint64_t)((__uint128_t)a * b % m); }
uint64_t invmod(uint64_t a, uint64_t m) {  // a^(m-2), m prime
  uint64_t r = 1, e = m - 2;
  for (a %= m; e; e >>= 1, a = mulmod(a, a, m))
    if (e & 1) r = mulmod(r, a, m);
  return r;
}

}  // namespace

extern "C" {

int hb_trace_sizes(int set, uint64_t* round_words, uint64_t* tail_words) {
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  if (round_words) *round_words = (uint64_t)(22 + 12 * p.br_len) * p.N + 2;
  if (tail_words)
    *tail_words = (uint64_t)2 * (p.n + 1) + p.N + (uint64_t)6 * p.ks_len * p.N + 6 * p.N;
  return HB_OK;
}

int hb_trace_expand(const hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1,
                    const uint64_t* c2, const uint64_t* core, const uint32_t* rounds,
                    uint32_t n_rounds, uint64_t* rec, uint64_t* tail) {
  if (int rc = check_trace_args(ek, gate, c0, c1, core)) return rc;
  const ParamSet& p = ek->p;
  const uint32_t N = p.N, n = p.n, L = p.br_len;
  for (uint32_t i = 1; i < n_rounds; i++)
    if (!rounds || rounds[i] <= rounds[i - 1]) return fail(HB_ERR_ARG, "rounds must increase");
  if (n_rounds && (!rec || !rounds || rounds[n_rounds - 1] >= n))
    return fail(HB_ERR_ARG, "bad round list");
  Ctx X(p);
  uint64_t round_words = 0, tail_words = 0;
  hb_trace_sizes(p.set, &round_words, &tail_words);
  // modulus switch of the combined input (fhe_core/src/modulus_switch.rs:41-86)
  std::vector<uint32_t> ms(n + 1);
  std::vector<uint64_t> v(n + 1);
  switch_input(p, gate, c0, c1, c2, v.data(), ms.data());
  if (tail)
    for (uint32_t k = 0; k <= n; k++) tail[k] = v[k], tail[n + 1 + k] = ms[k];
  std::vector<uint64_t> acc(2 * N), nxt(2 * N), prod(2 * N), ext(2 * N), ma(N), mb(N), m(N);
  const size_t rgsw = (size_t)2 * L * 2 * N;
  for (uint32_t t = 0; t < n_rounds; t++) {
    const uint32_t i = rounds[t];
    uint64_t* r = rec + (size_t)t * round_words;
    for (uint32_t k = 0; k < 2 * N; k++) {
      acc[k] = X.canon(core[(size_t)i * 2 * N + k]);
      nxt[k] = X.canon(core[(size_t)(i + 1) * 2 * N + k]);
    }
    X.put_rlwe(r, acc.data(), acc.data() + N);
    X.put_rlwe(r + 4 * N, nxt.data(), nxt.data() + N);
    // monomial X^a_i (binary.rs:158-176)
    const uint32_t a_i = ms[i];
    const uint32_t deg = a_i < N ? a_i : a_i - N;
    const uint64_t coef = a_i < N ? 1 : p.q - 1;
    std::fill(m.begin(), m.end(), 0);
    m[deg] = coef;
    X.put(r + 8 * N, m.data());
    X.put_ntt(r + 9 * N, m.data());
    r[10 * N] = deg;
    r[10 * N + 1] = coef;
    // monomial_times_acc = ACC * X^a_i, exact negacyclic rotation of canonical values
    for (uint32_t c = 0; c < 2; c++)
      for (uint32_t k = 0; k < N; k++) {
        uint32_t s = (k + 2 * N - a_i) % (2 * N);
        uint64_t v = acc[c * N + (s % N)];
        prod[c * N + k] = (s >= N) ? X.neg(v) : v;
      }
    X.put_rlwe(r + 10 * N + 2, prod.data(), prod.data() + N);
    for (uint32_t k = 0; k < 2 * N; k++) ext[k] = X.ntt.sub(prod[k], acc[k]);
    X.put_rlwe(r + 14 * N + 2, ext.data(), ext.data() + N);
    std::fill(ma.begin(), ma.end(), 0);
    std::fill(mb.begin(), mb.end(), 0);
    const uint64_t* key = ek->bsk_host.data() + (size_t)i * rgsw;
    uint64_t* had = r + 18 * N + 2;
    X.gadget(ext.data(), key, L, p.br_bits, ma.data(), mb.data(), had);
    X.gadget(ext.data() + N, key + (size_t)2 * L * N, L, p.br_bits, ma.data(), mb.data(),
             had + (size_t)6 * L * N);
    uint64_t* sp = r + (size_t)(18 + 12 * L) * N + 2;
    X.put(sp, ma.data());
    X.put(sp + N, mb.data());
    X.put_intt(sp + 2 * N, ma.data());
    X.put_intt(sp + 3 * N, mb.data());
    // consistency: the product recorded by the GPU (next acc - acc) is INTT(sum_prod)
    for (uint32_t k = 0; k < N; k++)
      if (X.ntt.sub(nxt[k], acc[k]) != sp[2 * N + k] || X.ntt.sub(nxt[N + k], acc[N + k]) != sp[3 * N + k])
        return fail(HB_ERR_STATE, "core trace inconsistent at round %u coefficient %u", i, k);
  }
  if (tail) {
    // key switch (key_switch/rlwe.rs:290-350): raw negation of the final acc.a
    const uint64_t* fin = core + (size_t)n * 2 * N;
    std::vector<uint64_t> neg(N);
    for (uint32_t k = 0; k < N; k++)
      neg[k] = X.canon(p.is_goldilocks ? Gl::neg_raw(fin[k]) : X.neg(fin[k]));
    uint64_t* t = tail + 2 * (n + 1);
    X.put(t, neg.data());
    std::fill(ma.begin(), ma.end(), 0);
    std::fill(mb.begin(), mb.end(), 0);
    X.gadget(neg.data(), ek->ksk_host.data(), p.ks_len, p.ks_bits, ma.data(), mb.data(), t + N);
    uint64_t* sp = t + N + (size_t)6 * p.ks_len * N;
    X.put(sp, ma.data());
    X.put(sp + N, mb.data());
    X.put_intt(sp + 2 * N, ma.data());
    X.put_intt(sp + 3 * N, mb.data());
    // sample extraction input = INTT(init.a) = INTT(sum_prod a); output permuted
    X.put(sp + 4 * N, sp + 2 * N);
    uint64_t* o = sp + 5 * N;
    o[0] = sp[4 * N];
    for (uint32_t k = 1; k < N; k++) o[k] = X.neg(sp[4 * N + (N - k)]);
  }
  return HB_OK;
}

int hb_trace_part_words(int set, int part, uint64_t* words) {
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  if (!words) return fail(HB_ERR_ARG, "null argument");
  *words = part_words(p, part);
  return *words ? HB_OK : fail(HB_ERR_ARG, "unknown trace part %d", part);
}

// One finalized container, built straight from the core trace (acc rows entering each
// round, and the last one): rows are the n rounds, zero padded to R = 2^ceil(log2 n).
int hb_trace_part(const hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1,
                  const uint64_t* c2, const uint64_t* core, int part, uint32_t index,
                  uint64_t* out) {
  if (int rc = check_trace_args(ek, gate, c0, c1, core)) return rc;
  const ParamSet& p = ek->p;
  const uint32_t N = p.N, n = p.n, L = p.br_len;
  const uint64_t words = part_words(p, part);
  if (!words) return fail(HB_ERR_ARG, "unknown trace part %d", part);
  if (!out) return fail(HB_ERR_ARG, "null output");
  if (part == HB_TP_HADAMARD && index >= 2 * L)
    return fail(HB_ERR_ARG, "hadamard index %u out of range [0, %u)", index, 2 * L);
  std::memset(out, 0, words * 8);
  const size_t R = (size_t)1 << ilog2(n), RN = R * N;
  std::vector<uint32_t> ms(n + 1);
  std::vector<uint64_t> v(n + 1);
  switch_input(p, gate, c0, c1, c2, v.data(), ms.data());
  Ctx X(p);
  auto acc_row = [&](uint32_t i, uint32_t half, uint64_t* dst) {
    for (uint32_t k = 0; k < N; k++) dst[k] = X.canon(core[(size_t)i * 2 * N + half * N + k]);
  };
  // RLWETrace planes [poly.0, poly.1, ntt.0, ntt.1] of row r from (a, b)
  auto put_row = [&](uint64_t* base, size_t plane, size_t r, const uint64_t* a, const uint64_t* b) {
    X.put(base + r * N, a);
    X.put(base + plane + r * N, b);
    X.put_ntt(base + 2 * plane + r * N, a);
    X.put_ntt(base + 3 * plane + r * N, b);
  };
  std::vector<uint64_t> a(N), b(N), pa(N), pb(N);
  switch (part) {
    case HB_TP_INITIAL_ACC:
    case HB_TP_FINAL_ACC: {
      const uint32_t i = part == HB_TP_INITIAL_ACC ? 0 : n;
      acc_row(i, 0, a.data());
      acc_row(i, 1, b.data());
      put_row(out, N, 0, a.data(), b.data());
      return HB_OK;
    }
    case HB_TP_PERMUTATION:  // PermutationInfo::new_rotation_left(log R, n, 1)
      for (size_t r = 0; r < R; r++) out[r] = r < n ? (r + 1) % n : r;
      return HB_OK;
    case HB_TP_MODULUS_SWITCH: {  // ModulusSwitchingTrace::new + finalize(n + 1)
      const uint64_t Q = p.q, q2 = 2ull * N, M = 1ull << ilog2(n + 1);
      const uint64_t k = mulmod(Q - 1, invmod(2 * q2, Q), Q), two_k = mulmod(Q - 1, invmod(q2, Q), Q);
      for (uint64_t i = 0; i < M; i++) {
        const bool real = i <= n;
        const uint64_t in = real ? X.canon(v[i]) : 0, o = real ? ms[i] : 0;
        const uint64_t wit = o == 0 ? q2 : o;
        const uint64_t t = mulmod((2 * wit + Q - 1) % Q, k, Q);  // (2b' - 1) k
        out[i] = in;
        out[M + i] = o;
        out[2 * M + i] = wit;
        out[3 * M + i] = real ? X.ntt.sub(X.ntt.sub(in, t), 1) : k;
      }
      out[4 * M] = q2;
      out[4 * M + 1] = k;
      out[4 * M + 2] = two_k;
      return HB_OK;
    }
    default: break;
  }
  const size_t rgsw = (size_t)2 * L * 2 * N;
  const uint64_t mask = (1ull << p.br_bits) - 1;
  for (uint32_t r = 0; r < n; r++) {
    acc_row(r, 0, a.data());
    acc_row(r, 1, b.data());
    const uint32_t a_i = ms[r];
    const uint32_t deg = a_i < N ? a_i : a_i - N;
    const uint64_t coef = a_i < N ? 1 : p.q - 1;
    switch (part) {
      case HB_TP_INPUT_ACC: put_row(out, RN, r, a.data(), b.data()); break;
      case HB_TP_OUTPUT_ACC:
      case HB_TP_INPUT_ACC_PERMUTED:
      case HB_TP_SUM_PROD: {
        acc_row(r + 1, 0, pa.data());
        acc_row(r + 1, 1, pb.data());
        if (part == HB_TP_SUM_PROD) {  // poly = next - acc; its NTT is the NTT-domain sum
          for (uint32_t k = 0; k < N; k++) pa[k] = X.ntt.sub(pa[k], a[k]), pb[k] = X.ntt.sub(pb[k], b[k]);
        }
        put_row(out, RN, r, pa.data(), pb.data());
        break;
      }
      case HB_TP_MONOMIAL:
        out[(size_t)r * N + deg] = coef;
        X.put_ntt(out + RN + (size_t)r * N, out + (size_t)r * N);
        break;
      case HB_TP_MONOMIAL_REPR:
        out[r] = deg;
        out[R + r] = coef;
        break;
      case HB_TP_MONOMIAL_TIMES_ACC:
      case HB_TP_EXTERNAL_PRODUCT_INPUT:
      case HB_TP_HADAMARD: {
        for (uint32_t k = 0; k < N; k++) {  // X^a_i * ACC, exact negacyclic rotation
          const uint32_t s = (k + 2 * N - a_i) % (2 * N);
          pa[k] = s >= N ? X.neg(a[s % N]) : a[s % N];
          pb[k] = s >= N ? X.neg(b[s % N]) : b[s % N];
        }
        if (part == HB_TP_MONOMIAL_TIMES_ACC) {
          put_row(out, RN, r, pa.data(), pb.data());
          break;
        }
        for (uint32_t k = 0; k < N; k++) pa[k] = X.ntt.sub(pa[k], a[k]), pb[k] = X.ntt.sub(pb[k], b[k]);
        if (part == HB_TP_EXTERNAL_PRODUCT_INPUT) {
          put_row(out, RN, r, pa.data(), pb.data());
          break;
        }
        // hadamard index = half * L + digit: the digit polynomial of ext[half], the key row
        const uint32_t half = index / L, dig = index % L;
        const uint64_t* src = half ? pb.data() : pa.data();
        uint64_t* bit = out + (size_t)r * N;
        for (uint32_t k = 0; k < N; k++) bit[k] = (src[k] >> (dig * p.br_bits)) & mask;
        X.put_ntt(out + RN + (size_t)r * N, bit);
        const uint64_t* ka = ek->bsk_host.data() + (size_t)r * rgsw + (size_t)half * 2 * L * N +
                             (size_t)2 * dig * N;
        for (uint32_t k = 0; k < N; k++) a[k] = X.canon(ka[k]), b[k] = X.canon(ka[N + k]);
        X.put_intt(out + 2 * RN + (size_t)r * N, a.data());
        X.put_intt(out + 3 * RN + (size_t)r * N, b.data());
        X.put(out + 4 * RN + (size_t)r * N, a.data());
        X.put(out + 5 * RN + (size_t)r * N, b.data());
        break;
      }
      default: break;
    }
  }
  if (part == HB_TP_INPUT_ACC_PERMUTED) {  // last poly row += initial - final
    const size_t last = (size_t)(n - 1) * N;
    acc_row(0, 0, a.data());
    acc_row(0, 1, b.data());
    acc_row(n, 0, pa.data());
    acc_row(n, 1, pb.data());
    for (uint32_t k = 0; k < N; k++) {
      out[last + k] = X.ntt.add(out[last + k], X.ntt.sub(a[k], pa[k]));
      out[RN + last + k] = X.ntt.add(out[RN + last + k], X.ntt.sub(b[k], pb[k]));
    }
  }
  return HB_OK;
}

}  // extern "C"
