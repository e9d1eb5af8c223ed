// Parameter sets of the gate-bootstrapping path.
// BABYBEAR / GOLDILOCKS / ZAMA restate tfhe/src/bfhe/parameters.rs:45-99
// (ZAMA_BINARY_128_BITS_PARAMETERS is identical to vfhe's ZAMA_GOLDILOCKS_PARAMETERS,
// vfhe/src/bfhe/parameters.rs:26-46). Parameter validation follows
// fhe_core/src/parameter.rs:111-172 (N a power of two, 2N | Q-1, t a power of two).
#pragma once
#include <cstdint>

namespace hb {

enum SetId : int { SET_BABYBEAR = 0, SET_GOLDILOCKS = 1, SET_ZAMA = 2, SET_COUNT = 3 };

struct ParamSet {
  int set;
  int is_goldilocks;
  uint32_t n;        // LWE dimension
  uint32_t N;        // ring dimension
  uint32_t t;        // plaintext modulus
  uint32_t br_bits;  // blind-rotation gadget basis bits
  uint32_t br_len;   // ceil(bits(Q) / br_bits)
  uint32_t ks_bits;  // key-switching gadget basis bits
  uint32_t ks_len;   // ceil(bits(Q) / ks_bits)
  uint64_t q;
  double lwe_sigma;
  double ring_sigma;
  double ks_sigma;
};

inline bool param_set(int set, ParamSet* out) {
  auto ceil_div = [](uint32_t a, uint32_t b) { return (a + b - 1) / b; };
  ParamSet p{};
  p.set = set;
  p.n = 728;
  p.N = 1024;
  p.t = 4;
  switch (set) {
    case SET_BABYBEAR:
      p.is_goldilocks = 0;
      p.q = 0x78000001ull;
      p.br_bits = 8;
      p.ks_bits = 5;
      p.br_len = ceil_div(31, p.br_bits);
      p.ks_len = ceil_div(31, p.ks_bits);
      p.lwe_sigma = 11000.0;
      p.ring_sigma = 41.9;
      p.ks_sigma = 11000.0;
      break;
    case SET_GOLDILOCKS:
    case SET_ZAMA:
      p.is_goldilocks = 1;
      p.q = 0xFFFFFFFF00000001ull;
      p.br_bits = set == SET_ZAMA ? 5 : 8;
      p.ks_bits = 5;
      p.br_len = ceil_div(64, p.br_bits);
      p.ks_len = ceil_div(64, p.ks_bits);
      p.lwe_sigma = 2.9 * (double)(1ull << 45);
      p.ring_sigma = 2.3 * (double)(1ull << 37);
      p.ks_sigma = 2.9 * (double)(1ull << 45);
      break;
    default:
      return false;
  }
  *out = p;
  return true;
}

}  // namespace hb
