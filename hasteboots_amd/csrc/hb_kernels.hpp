// Kernel-side argument block and launchers (hb_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace hb {

enum GateKind : int {
  GATE_NAND = 0,
  GATE_AND = 1,
  GATE_OR = 2,
  GATE_NOR = 3,
  GATE_XOR = 4,
  GATE_XNOR = 5,
  GATE_MAJORITY = 6,
  GATE_BOOTSTRAP = 7,  // c0 is the ciphertext, `lut` the test polynomial
  GATE_MUX = 8         // host-side sequence: NOT, two ANDs, OR-LUT bootstrap of the sum
};

// Words of the Goldilocks forward table of the digit transforms (the twisted transform of
// hb_ntt.hpp: 16 blocks of 64 twist factors at a 68-word stride, 16 + 32 pass-3 twiddles);
// BootArgs::tw_fwd of the Goldilocks sets.
constexpr int kTwxWords = 16 * 68 + 48;

struct BootArgs {
  const void* bsk;     // device layout, see hb_kernels.hip
  const void* ksk;
  const void* tw_fwd;  // psi^bitrev(k), N words (BabyBear); Goldilocks: kTwxWords words
  const void* tw_inv;  // psi^-bitrev(k), N words
  const uint64_t* c0;  // count*(n+1) canonical words
  const uint64_t* c1;  // may be null for GATE_BOOTSTRAP
  const uint64_t* c2;  // majority only
  const uint64_t* lut; // null: built from (lut_mid, lut_lo, lut_hi)
  uint64_t* out;       // count*(n+1) canonical words
  void* acc;           // count * kScratchWordsPerGate words of scratch
  double qf;           // (double)Q, as the reference's `as f64`
  uint64_t q8;         // Q >> 3 in the device word form
  uint64_t lut_lo, lut_hi;
  uint32_t lut_mid;
  uint32_t n, br_len, br_bits, ks_len, ks_bits;
  uint32_t count;
  int gate;
  int is_goldilocks;
  uint32_t first_block;  // mixed launches: first workgroup index of this part
  const uint32_t* idx;   // optional: per gate (in0, in1, out) wire rows of c0 / out
  void* trace;           // optional: traced (vfhe) semantics, count*(n+1)*2N acc history
  uint64_t* pair_flags;  // pair mode: count * kPairFlagWords words, zeroed once
  uint32_t pair_epoch;   // pair mode: this launch's epoch (> every earlier one on pair_flags)
  uint64_t* pair_status;  // pair mode: [0] epoch << 32 of a launch whose wait timed out,
                          // [1] the workgroup ticket counter (pair_ticket)
  uint64_t pair_ticket0;  // pair mode: the counter's value when this launch starts
  uint64_t pair_wait_ticks;  // pair mode: wall-clock ticks before a wait times out (0: at once)
};

// Pair mode's per-gate flag words (launch_bootstrap_pair): [0, 8) the quarter flags of the
// two workgroups, 8 the b component's word 0 for the key switch, 9 its flag, 10 the epoch
// of a launch whose wait timed out.
constexpr int kPairFlagWords = 16;
// Pair-mode wait limit, microseconds: a round's exchange waits microseconds, a workgroup
// waiting for its launch's next workgroup to start waits for a CU to free up.
constexpr uint32_t kPairWaitDefaultUs = 2000000;
// Device words of a pair-mode status (pair_status): timeout epoch, ticket counter.
constexpr int kPairStatusWords = 2;

// Per-gate scratch in BootArgs::acc, in words of the set's size: the accumulator RLWE (2N
// words) then the rotated accumulator's byte-sized gadget digits, byte-transposed (one
// 16-byte word per lane and digit: 8 digits of 8 bits or 13 of 5 bits, at most 2N words).
constexpr size_t kScratchWordsPerGate = 4 * 1024;

// Batched-kernel variants: gpw gates (wavefronts) per workgroup and wps waves per SIMD
// targeted by register allocation; (gpw, wps) in {(1,1), (2,1), (4,1), (2,2), (4,2), (4,3)}.
hipError_t launch_bootstrap(const BootArgs& a, int is_goldilocks, int gpw, int wps,
                            hipStream_t s);
// Pair mode (untraced; Goldilocks with a byte-digit gadget or BabyBear with at most 4
// digits): one gate on two workgroups of 8 waves, on two CUs, one accumulator component
// each, 2 * count workgroups paired in the order they start (pair_flags, pair_epoch,
// pair_status, pair_ticket0 and pair_wait_ticks set; launches on one pair_status run one
// after another).
hipError_t launch_bootstrap_pair(const BootArgs& a, hipStream_t s);
// Latency mode: `waves` (4 or 8) wavefronts per gate (traced gates: 8 only).
hipError_t launch_bootstrap_split(const BootArgs& a, int is_goldilocks, int waves,
                                  hipStream_t s);
// One launch over `nparts` parts (device array of BootArgs, parts in block order).
// Latency-mode mixed launch (`waves` = 4 or 8 wavefronts per gate; `traced`: some part is
// traced, 8 waves only).
hipError_t launch_mixed_split(const BootArgs* parts_dev, int nparts, uint32_t total_gates,
                              int waves, bool traced, hipStream_t s);
hipError_t launch_mixed(const BootArgs* parts_dev, int nparts, uint32_t total_gates,
                        hipStream_t s);
// Pair-mode mixed launch (the parts' first_block count gates; every part untraced,
// Goldilocks with a byte-digit gadget or BabyBear with at most 4 digits, with the pair
// fields set as above and one pair_status): cls 1 every part Goldilocks with 8-bit digits,
// 2 every part 5-bit, 3 every part BabyBear, 0 any.
hipError_t launch_mixed_pair(const BootArgs* parts_dev, int nparts, uint32_t total_gates,
                             int cls, hipStream_t s);
// Throughput-mode mixed launch: kMixedBatchGpw gates of one part per workgroup, two waves
// per SIMD (the batched kernel's (4, 2) shape). first_block counts workgroups here, so
// part k occupies ceil(count_k / kMixedBatchGpw) workgroups from its first_block.
// cls: 1 every part Goldilocks with 8-bit digits, 2 every part another Goldilocks set,
// 3 every part BabyBear (one kernel body each), 0 any mix.
constexpr uint32_t kMixedBatchGpw = 4;
hipError_t launch_mixed_batch(const BootArgs* parts_dev, int nparts, uint32_t total_blocks,
                              int cls, hipStream_t s);
// Sets first_block of parts[0..n) for a launch with `gpw` gates per workgroup (1 for
// launch_mixed and launch_mixed_split) and returns the workgroup count.
inline uint32_t assign_blocks(BootArgs* parts, size_t n, uint32_t gpw) {
  uint32_t b = 0;
  for (size_t k = 0; k < n; k++) {
    parts[k].first_block = b;
    b += (parts[k].count + gpw - 1) / gpw;
  }
  return b;
}
// Evaluation-key generation (hb_keygen.hip): `rows` RLWE samples of N coefficients into
// `keys` (canonical [row][a, b][N] layout), from the counter-based samplers of hb_sample.hpp.
// BSK (is_bsk): row r = (i * 2 + c) * L + j, stream (kind, i), gadget basis 2^basis_bits,
// s = NTT(s_ring). KSK: row r, stream (kind, r), s = NTT(s_lwe ‖ 0), plus B_ks^r NTT(s_ring).
struct KeygenArgs {
  uint32_t key[8];
  uint32_t kind, rows, N, L, basis_bits;
  int is_bsk, is_goldilocks;
  uint64_t q;
  double sigma;
  const uint8_t* lwe_sk;  // device, n bytes (BSK only)
};
// err: rows * N words of device scratch; s_ntt and ring_ntt: N canonical words on device.
hipError_t launch_keygen(const KeygenArgs& a, uint64_t* keys, uint64_t* err, const uint64_t* s_ntt,
                         const uint64_t* ring_ntt, const void* tw_fwd, uint32_t cus, hipStream_t s);
// Canonical key words -> the bootstrap kernels' device form (× scale, Montgomery; u32 words
// for BabyBear).
hipError_t launch_keys_to_device(const uint64_t* src, void* dst, size_t words, uint64_t scale,
                                 int is_goldilocks, uint64_t q, hipStream_t s);
// Evaluator::not on `count` LWEs of n1 canonical words (out may equal in).
hipError_t launch_lwe_not(const uint64_t* in, uint64_t* out, uint32_t count, uint32_t n1,
                          uint64_t q, int is_goldilocks, hipStream_t s);
// Microbench NTT over `count` polynomials on a device with `cus` compute units.
hipError_t launch_ntt(int is_goldilocks, int inverse, uint32_t log_n, uint64_t* data, size_t count,
                      const void* tw, uint64_t ninv, uint32_t cus, hipStream_t s);
// Goldilocks MAC peak microbench: `waves` wavefronts, `iters` x 4 gl_mac2 per lane;
// out holds waves*64 words.
hipError_t launch_modmul_peak(uint64_t* out, uint32_t waves, uint32_t iters, uint64_t seed,
                              hipStream_t s);
hipError_t launch_extprod(int is_goldilocks, uint64_t* rlwe, size_t count, const void* rgsw,
                          uint32_t len, uint32_t bits, const void* tw, const void* itw,
                          hipStream_t s);

}  // namespace hb
