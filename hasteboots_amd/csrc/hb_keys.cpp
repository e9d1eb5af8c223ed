// Host-side key generation, encryption and decryption for the gate API.
#include "hb_keys.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

#include "hb_field.hpp"
#include "hb_sample.hpp"

namespace hb {

// ------------------------------ ChaCha20 ------------------------------------
// Sequential keystream over chacha20_block (secret keys and encryption); the evaluation
// keys use the counter-based samplers of hb_sample.hpp.
ChaCha20::ChaCha20(uint64_t seed, uint32_t kind, uint32_t index) {
  derive_key(seed, key_);
  nonce_[0] = kind;
  nonce_[1] = index;
  nonce_[2] = 0x68623031u;  // domain tag
}

void ChaCha20::refill() {
  chacha20_block(key_, counter_, nonce_[0], nonce_[1], nonce_[2], buf_);
  counter_++;
  pos_ = 0;
}

uint32_t ChaCha20::next_u32() {
  if (pos_ >= 16) refill();
  return buf_[pos_++];
}
uint64_t ChaCha20::next_u64() {
  uint64_t lo = next_u32();
  uint64_t hi = next_u32();
  return lo | (hi << 32);
}
double ChaCha20::next_unit() { return (double)((next_u64() >> 11) + 1) * 0x1.0p-53; }

// ------------------------------ samplers ------------------------------------
// FieldUniformSampler (algebra/src/random.rs:11-35) by rejection.
static uint64_t sample_uniform(ChaCha20& r, const ParamSet& p) {
  for (;;) {
    uint64_t x = p.is_goldilocks ? r.next_u64() : (uint64_t)(r.next_u32() >> 1);
    if (x < p.q) return x;
  }
}
// FieldDiscreteGaussianSampler (algebra/src/random.rs:156-170): N(0, sigma) cut at 6 sigma,
// rounded, negative values mapped to Q - |v|. Normal deviates by Box-Muller.
static uint64_t sample_gaussian(ChaCha20& r, double sigma, uint64_t q) {
  const double two_pi = 6.283185307179586476925286766559;
  for (;;) {
    double u1 = r.next_unit(), u2 = r.next_unit();
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(two_pi * u2) * sigma;
    if (std::fabs(z) < 6.0 * sigma) {
      double rz = std::round(z);
      if (rz < 0.0) return q - (uint64_t)(-rz);
      return (uint64_t)rz;
    }
  }
}
// sample_binary_field_vec (algebra/src/utils/sample.rs:6-27): one u32 per 32 entries.
static void sample_binary(ChaCha20& r, uint64_t* v, uint32_t len) {
  uint32_t i = 0;
  for (; i + 32 <= len; i += 32) {
    uint32_t w = r.next_u32();
    for (int b = 0; b < 32; b++) v[i + b] = (w >> b) & 1u;
  }
  uint32_t w = r.next_u32();
  for (int b = 0; i < len; i++, b++) v[i] = (w >> b) & 1u;
}
// sample_ternary_field_vec (sample.rs:30-51): {0, 0, 1, -1}[r & 3], one u32 per 16 entries.
static void sample_ternary(ChaCha20& r, uint64_t* v, uint32_t len, uint64_t q) {
  const uint64_t table[4] = {0, 0, 1, q - 1};
  uint32_t i = 0;
  for (; i + 16 <= len; i += 16) {
    uint32_t w = r.next_u32();
    for (int b = 0; b < 16; b++) v[i + b] = table[(w >> (2 * b)) & 3u];
  }
  uint32_t w = r.next_u32();
  for (int b = 0; i < len; i++, b++) v[i] = table[(w >> (2 * b)) & 3u];
}

// ------------------------------ host NTT ------------------------------------
static uint64_t mulmod_slow(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)((unsigned __int128)a * b % m);
}
static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = mulmod_slow(r, b, m);
    b = mulmod_slow(b, b, m);
    e >>= 1;
  }
  return r;
}
static uint32_t bitrev(uint32_t x, uint32_t bits) {
  uint32_t r = 0;
  for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

void HostNtt::init(const ParamSet& ps, uint32_t log_n) {
  p = ps;
  logN = log_n;
  N = 1u << log_n;
  const uint64_t g = ps.is_goldilocks ? Gl::GENERATOR : Bb::GENERATOR;
  psi = powmod(g, (ps.q - 1) / (2ull * N), ps.q);
  uint64_t psi_inv = powmod(psi, ps.q - 2, ps.q);
  psi_rev.resize(N);
  psi_inv_rev.resize(N);
  for (uint32_t k = 0; k < N; k++) {
    uint32_t e = bitrev(k, log_n);
    psi_rev[k] = powmod(psi, e, ps.q);
    psi_inv_rev[k] = powmod(psi_inv, e, ps.q);
  }
  n_inv = powmod(N, ps.q - 2, ps.q);
}
uint64_t HostNtt::add(uint64_t a, uint64_t b) const {
  unsigned __int128 s = (unsigned __int128)a + b;
  return (uint64_t)(s >= p.q ? s - p.q : s);
}
uint64_t HostNtt::sub(uint64_t a, uint64_t b) const { return a >= b ? a - b : a + (p.q - b); }
uint64_t HostNtt::mul(uint64_t a, uint64_t b) const {
  if (p.is_goldilocks) return Gl::canon(Gl::mul(a, b));
  return (a * b) % p.q;
}
void HostNtt::fwd(uint64_t* a) const {
  uint32_t t = N;
  for (uint32_t m = 1; m < N; m <<= 1) {
    t >>= 1;
    for (uint32_t i = 0; i < m; i++) {
      const uint64_t S = psi_rev[m + i];
      uint64_t* x = a + 2 * i * t;
      uint64_t* y = x + t;
      for (uint32_t j = 0; j < t; j++) {
        uint64_t V = mul(y[j], S);
        y[j] = sub(x[j], V);
        x[j] = add(x[j], V);
      }
    }
  }
}
void HostNtt::inv(uint64_t* a) const {
  uint32_t t = 1;
  for (uint32_t m = N; m > 1; m >>= 1) {
    uint32_t h = m >> 1;
    for (uint32_t i = 0; i < h; i++) {
      const uint64_t S = psi_inv_rev[h + i];
      uint64_t* x = a + 2 * i * t;
      uint64_t* y = x + t;
      for (uint32_t j = 0; j < t; j++) {
        uint64_t U = x[j], V = y[j];
        x[j] = add(U, V);
        y[j] = mul(sub(U, V), S);
      }
    }
    t <<= 1;
  }
  for (uint32_t j = 0; j < N; j++) a[j] = mul(a[j], n_inv);
}

// ------------------------------ keys ----------------------------------------
static uint32_t ilog2(uint32_t x) {
  uint32_t r = 0;
  while ((1u << r) < x) r++;
  return r;
}

// SecretKeyPack::new (fhe_core/src/secret_key.rs:67-89): binary LWE key, ternary ring key.
bool secret_key_create(int set, uint64_t seed, SecretKey* sk) {
  ParamSet p;
  if (!param_set(set, &p)) return false;
  sk->p = p;
  sk->seed = seed;
  sk->lwe_sk.assign(p.n, 0);
  sk->ring_sk.assign(p.N, 0);
  ChaCha20 r1(seed, STREAM_LWE_SK, 0);
  sample_binary(r1, sk->lwe_sk.data(), p.n);
  ChaCha20 r2(seed, STREAM_RING_SK, 0);
  sample_ternary(r2, sk->ring_sk.data(), p.N, p.q);
  HostNtt ntt;
  ntt.init(p, ilog2(p.N));
  sk->ring_sk_ntt = sk->ring_sk;
  ntt.fwd(sk->ring_sk_ntt.data());
  sk->enc_rng = ChaCha20(seed, STREAM_ENCRYPT, 0);
  return true;
}

static uint64_t encode(uint64_t m, const ParamSet& p) {  // plaintext.rs:64-75
  return (uint64_t)std::round((double)m * ((double)p.q / (double)p.t));
}
static uint64_t decode(uint64_t c, const ParamSet& p) {  // plaintext.rs:86-93
  uint64_t r = (uint64_t)std::round((double)c * (double)p.t / (double)p.q);
  return r < p.t ? r : r - p.t;
}

// SecretKeyPack::encrypt (secret_key.rs:129-150) with
// LWE::generate_random_zero_sample_field (lattice/src/lwe.rs:389-404).
void encrypt(SecretKey& sk, uint32_t bit, uint64_t* out) {
  const ParamSet& p = sk.p;
  HostNtt ops;
  ops.p = p;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < p.n; k++) {
    uint64_t a = sample_uniform(sk.enc_rng, p);
    out[k] = a;
    if (sk.lwe_sk[k]) acc = ops.add(acc, a);
  }
  uint64_t e = sample_gaussian(sk.enc_rng, p.lwe_sigma, p.q);
  uint64_t b = ops.add(acc, e);
  out[p.n] = ops.add(b, encode(bit, p));
}

// decrypt_with_noise (secret_key.rs:161-188).
void decrypt(const SecretKey& sk, const uint64_t* c, uint32_t* bit, uint64_t* noise) {
  const ParamSet& p = sk.p;
  HostNtt ops;
  ops.p = p;
  uint64_t acc = 0;
  for (uint32_t k = 0; k < p.n; k++)
    if (sk.lwe_sk[k]) acc = ops.add(acc, c[k]);
  uint64_t plain = ops.sub(c[p.n], acc);
  uint64_t m = decode(plain, p);
  uint64_t fresh = encode(m, p) % p.q;
  if (bit) *bit = (uint32_t)m;
  if (noise) *noise = std::min(ops.sub(plain, fresh), ops.sub(fresh, plain));
}

uint64_t bsk_words(const ParamSet& p) { return (uint64_t)p.n * 2 * p.br_len * 2 * p.N; }
uint64_t ksk_words(const ParamSet& p) { return (uint64_t)p.ks_len * 2 * p.N; }

// NTTRLWE::generate_random_{zero,value}_sample (lattice/src/rlwe.rs:1215-1252):
// a uniform in the NTT domain, e Gaussian in the coefficient domain (+value at X^0),
// b = NTT(e) + a ⊙ s. Coefficient k draws from block `first + k` of stream (kind, row).
static void rlwe_sample(const uint32_t key[8], uint32_t kind, uint32_t row, uint32_t first,
                        const HostNtt& ntt, const uint64_t* s_ntt, double sigma, uint64_t value,
                        uint64_t* a, uint64_t* b) {
  const ParamSet& p = ntt.p;
  for (uint32_t k = 0; k < ntt.N; k++) key_coefficient(key, kind, row, first + k, p.q, sigma, a + k, b + k);
  b[0] = ntt.add(b[0], value);
  ntt.fwd(b);
  for (uint32_t k = 0; k < ntt.N; k++) b[k] = ntt.add(b[k], ntt.mul(a[k], s_ntt[k]));
}

// BinaryBlindRotationKey::generate (fhe_core/src/blind_rotation/rlwe/binary.rs:207-238):
// RGSW(0) or RGSW(1) per LWE key bit, via NTTRGSW::generate_random_{zero,one}_sample
// (lattice/src/rgsw.rs:343-392) and NTTGadgetRLWE generators (gadget/rlwe.rs:478-551).
// Sample (c, j) of key bit i draws blocks [(c L + j) N, (c L + j + 1) N) of stream (BSK, i).
void generate_bsk(const SecretKey& sk, uint64_t* bsk) {
  const ParamSet& p = sk.p;
  HostNtt ntt;
  ntt.init(p, ilog2(p.N));
  const uint32_t N = p.N, L = p.br_len;
  uint32_t key[8];
  derive_key(sk.seed, key);
  unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto work = [&](unsigned tid) {
    for (uint32_t i = tid; i < p.n; i += nth) {
      const bool one = sk.lwe_sk[i] != 0;
      for (uint32_t c = 0; c < 2; c++) {
        uint64_t basis_power = 1;
        for (uint32_t j = 0; j < L; j++) {
          uint64_t* a = bsk + ((((uint64_t)i * 2 + c) * L + j) * 2) * N;
          uint64_t* b = a + N;
          // c = 1 (c_m): value sample of B^j when s_i = 1 (gadget/rlwe.rs:494-527).
          uint64_t value = (one && c == 1) ? basis_power : 0;
          rlwe_sample(key, STREAM_BSK, i, (c * L + j) * N, ntt, sk.ring_sk_ntt.data(),
                      p.ring_sigma, value, a, b);
          // c = 0 (c_neg_s_m): zero sample with a += B^j when s_i = 1 (gadget/rlwe.rs:530-551).
          if (one && c == 0)
            for (uint32_t k = 0; k < N; k++) a[k] = ntt.add(a[k], basis_power);
          basis_power = basis_power << p.br_bits;
        }
      }
    }
  };
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nth; t++) pool.emplace_back(work, t);
  for (auto& t : pool) t.join();
}

// KeySwitchingRLWEKey::generate (fhe_core/src/key_switch/rlwe.rs:37-112), branch
// n.next_power_of_two() == N: RLWE samples under NTT(s_lwe ‖ 0), b += B_ks^i · NTT(s_ring).
// Sample i draws blocks [0, N) of stream (KSK, i).
void generate_ksk(const SecretKey& sk, uint64_t* ksk) {
  const ParamSet& p = sk.p;
  HostNtt ntt;
  ntt.init(p, ilog2(p.N));
  const uint32_t N = p.N;
  uint32_t key[8];
  derive_key(sk.seed, key);
  std::vector<uint64_t> lwe_ntt(N, 0);
  for (uint32_t k = 0; k < p.n; k++) lwe_ntt[k] = sk.lwe_sk[k];
  ntt.fwd(lwe_ntt.data());
  std::vector<uint64_t> s = sk.ring_sk_ntt;
  const uint64_t basis = 1ull << p.ks_bits;
  for (uint32_t i = 0; i < p.ks_len; i++) {
    uint64_t* a = ksk + (uint64_t)i * 2 * N;
    uint64_t* b = a + N;
    rlwe_sample(key, STREAM_KSK, i, 0, ntt, lwe_ntt.data(), p.ks_sigma, 0, a, b);
    for (uint32_t k = 0; k < N; k++) b[k] = ntt.add(b[k], s[k]);
    if (i + 1 < p.ks_len)
      for (uint32_t k = 0; k < N; k++) s[k] = ntt.mul(s[k], basis);
  }
}

}  // namespace hb
