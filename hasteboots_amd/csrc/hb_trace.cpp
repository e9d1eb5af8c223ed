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
//   sample-extraction RowPermTrace input/output (lattice/src/rlwe.rs:406-434).
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
  if (!ek || !c0 || !core) return fail(HB_ERR_ARG, "null argument");
  if (gate < 0 || gate > GATE_BOOTSTRAP) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (gate != GATE_BOOTSTRAP && !c1) return fail(HB_ERR_ARG, "gate needs two inputs");
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
  for (uint32_t k = 0; k <= n; k++) {
    uint64_t v = combine_host(p, gate, c0[k], c1 ? c1[k] : 0, c2 ? c2[k] : 0);
    double r = std::round((double)v * (double)(2 * N) / (double)p.q);
    uint64_t t = (uint64_t)r;
    ms[k] = (uint32_t)(t >= 2 * N ? t - 2 * N : t);
    if (tail) tail[k] = v, tail[n + 1 + k] = ms[k];
  }
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

}  // extern "C"
