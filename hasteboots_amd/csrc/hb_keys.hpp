// Host-side key material: seeded PRG, host NTT (keygen only), secret keys,
// evaluation-key generation, encryption and decryption.
#pragma once
#include <cstdint>
#include <vector>

#include "hb_params.hpp"

namespace hb {

// ChaCha20 keystream (RFC 8439 block function) used as the seeded CSPRNG.
// The reference seeds an AES-128-CTR PRG from OS entropy (algebra/src/utils/prg.rs:38-108);
// this engine derives every stream from (seed, stream kind, index) so keys are reproducible.
class ChaCha20 {
 public:
  ChaCha20(uint64_t seed, uint32_t kind, uint32_t index);
  uint32_t next_u32();
  uint64_t next_u64();
  // Uniform double in (0, 1].
  double next_unit();

 private:
  void refill();
  uint32_t key_[8];
  uint32_t nonce_[3];
  uint32_t counter_ = 0;
  uint32_t buf_[16];
  int pos_ = 16;
};

enum StreamKind : uint32_t {
  STREAM_LWE_SK = 1,
  STREAM_RING_SK = 2,
  STREAM_BSK = 3,
  STREAM_KSK = 4,
  STREAM_ENCRYPT = 5
};

// Host negacyclic NTT on canonical values (same convention as the device kernels).
struct HostNtt {
  ParamSet p;
  uint32_t N = 0, logN = 0;
  uint64_t psi = 0, n_inv = 0;
  std::vector<uint64_t> psi_rev, psi_inv_rev;  // canonical
  void init(const ParamSet& ps, uint32_t log_n);
  uint64_t add(uint64_t a, uint64_t b) const;
  uint64_t sub(uint64_t a, uint64_t b) const;
  uint64_t mul(uint64_t a, uint64_t b) const;
  void fwd(uint64_t* a) const;
  void inv(uint64_t* a) const;
};

struct SecretKey {
  ParamSet p;
  uint64_t seed = 0;
  std::vector<uint64_t> lwe_sk;       // n values in {0, 1}
  std::vector<uint64_t> ring_sk;      // N values in {0, 1, Q-1}
  std::vector<uint64_t> ring_sk_ntt;  // NTT(ring_sk)
  ChaCha20 enc_rng{0, STREAM_ENCRYPT, 0};
};

bool secret_key_create(int set, uint64_t seed, SecretKey* sk);
void encrypt(SecretKey& sk, uint32_t bit, uint64_t* out);
void decrypt(const SecretKey& sk, const uint64_t* c, uint32_t* bit, uint64_t* noise);

// Canonical key layouts: BSK [n][2][br_len][2][N], KSK [ks_len][2][N].
uint64_t bsk_words(const ParamSet& p);
uint64_t ksk_words(const ParamSet& p);
void generate_bsk(const SecretKey& sk, uint64_t* bsk);
void generate_ksk(const SecretKey& sk, uint64_t* ksk);

}  // namespace hb
