// Engine internals shared by the C-ABI translation units (hb_api.cpp, hb_server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "../../include/hb.h"
#include "hb_kernels.hpp"
#include "hb_keys.hpp"
#include "hb_params.hpp"

namespace hb {

// Records a message for hb_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::hb::fail(HB_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));   \
  } while (0)

// Per-device constant tables: twiddles for log N = 10, 11 in device word form.
struct DeviceTables {
  bool ready = false;
  void* tw[2][2][2] = {};    // [field][logN-10][inverse]
  uint64_t ninv[2][2] = {};  // device-form N^-1 for the microbench inverse
  // Goldilocks, N = 1024: the forward table of the twisted digit transform (kTwxWords
  // words, ntt_fwd_digits), the kernels' BootArgs::tw_fwd for the Goldilocks sets
  void* twx = nullptr;
};

// The process-wide engine: one device, its default stream, twiddle tables and the
// scratch of mixed-parameter launches.
struct Engine {
  std::mutex mu;
  int device = -1;
  hipStream_t stream = nullptr;
  DeviceTables tables;
  // Device address of the fixed scratch array of hb_gate_batch_dev_noalloc on this
  // engine's device (hipGetSymbolAddress at engine set-up).
  void* noalloc_scratch = nullptr;
  // Scratch of mixed-parameter launches (accumulators, the part table the kernel reads
  // throughout its run, host-call staging). mixed_ev is recorded after the last launch
  // that used them; the next user's stream waits for it (and a reallocation synchronises
  // with it), so calls on different streams never share the buffers while a kernel runs.
  void* mixed_acc = nullptr;
  size_t mixed_acc_bytes = 0;
  BootArgs* mixed_parts = nullptr;
  size_t mixed_parts_cap = 0;
  uint64_t* mixed_io = nullptr;
  size_t mixed_io_words = 0;
  hipEvent_t mixed_ev = nullptr;
  bool mixed_ev_used = false;
  // Batched-kernel variant per parameter set (HB_SET_*): gates (wavefronts) per
  // workgroup and waves per SIMD. The default is the variant measured fastest on MI355X
  // for Goldilocks; hb_autotune or the cache file replaces it.
  int gpw[3] = {4, 4, 4};
  int wps[3] = {2, 2, 2};
  bool tuned[3] = {false, false, false};
  // Batches of at most this many gates (one wave per SIMD across the device) run the
  // 1-gate-per-workgroup, 1-wave-per-SIMD variant: spreading them over every SIMD beats
  // doubling up waves on half of them.
  uint32_t simd_slots = 1024;
  // Per parameter set, batches of at most split_max[set] untraced gates run in latency
  // mode (split_waves wavefronts per gate, one workgroup per gate). Defaults
  // (default_split_max): seven gates per CU for Goldilocks and BabyBear, six for Zama, the
  // largest batches where latency mode still wins. Measured on MI355X (256 CUs), kernel ms
  // in latency mode / the default path above the cutoff: Goldilocks 1792 gates 98.5 /
  // 103.8, 2048 112.6 / 104.1; Zama 1536 138.5 / 159.0, 1792 161.7 / 160.4; BabyBear 1792
  // 27.8 / 28.5, 2048 31.5 / 28.5.
  uint32_t split_max[3] = {1792, 1792, 1536};
  // Pair mode (untraced batches within the latency-mode cutoff and at most one gate per
  // two CUs): hb_set_pair_mode switches it; hb_set_pair_wait_limit sets how long a wait
  // lasts before it times out. wall_khz: the rate of the device clock wall_clock64 reads.
  bool pair_mode = true;
  uint32_t pair_wait_us = kPairWaitDefaultUs;
  uint32_t wall_khz = 100000;
  uint64_t pair_wait_ticks() const { return (uint64_t)pair_wait_us * wall_khz / 1000; }
};

// Built-in latency-mode cutoff of a parameter set on a device with `simd_slots` SIMDs.
inline uint32_t default_split_max(int set, uint32_t simd_slots) {
  const uint32_t cus = simd_slots / 4;
  return (set == SET_ZAMA ? 6 : 7) * cus;
}

// Latency-mode cutoff of a launch mixing the parameter sets in the bit mask `sets`: the
// smallest of their cutoffs (0 if any of them has latency mode off).
inline uint32_t mixed_split_cutoff(const Engine& E, int sets) {
  uint32_t c = 0xFFFFFFFFu;
  for (int set = 0; set < SET_COUNT; set++)
    if (sets & (1 << set)) c = E.split_max[set] < c ? E.split_max[set] : c;
  return sets ? c : 0;
}

// Start/stop events of the calling thread's last timed launch (hb_last_kernel_ms reads
// them); created on first use, one pair per thread.
struct LaunchTimer {
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool used = false;
};
LaunchTimer& launch_timer();
// Records the start event on `s`; returns HB_OK or an error code.
int timer_start(hipStream_t s);
int timer_stop(hipStream_t s);

// Wavefronts per gate in latency mode: 8 (two per SIMD, hiding each other's latency) for
// the Goldilocks sets, and for any launch of at most one gate per CU; otherwise 4, for
// BabyBear, whose 4-wave workgroups fit two to a CU. Measured on MI355X (kernel ms, 8 / 4
// waves): Goldilocks 256 gates 17.9 / 19.5, 512 35.8 / 39.0; Zama 256 28.2 / 30.2;
// BabyBear 256 5.9 / 6.0, 512 11.5 / 8.5, 768 17.2 / 13.1. With 8 waves a wave's share of a gadget (at
// most ceil(len / 4) digits) is held in 32 bits, which every parameter set here meets
// (`narrow` false falls back to 4 waves).
inline bool split8_fits(uint32_t br_len, uint32_t br_bits) {
  return (br_len + 3) / 4 * br_bits <= 32;
}
inline int split_waves(const Engine& E, bool all_goldilocks, size_t count, bool narrow = true) {
  if (!narrow) return 4;
  return all_goldilocks || count <= E.simd_slots / 4 ? 8 : 4;
}
// The mixed latency-mode kernel sizes its LDS for 64-bit words whatever the parts, so one
// workgroup fits a CU with 4 waves or with 8: 8 whenever the digit windows allow.
inline int mixed_split_waves(bool narrow) { return narrow ? 8 : 4; }

// Kernel of a mixed launch of `gates` gates over the parameter sets in `sets` (bit mask):
// latency mode up to the sets' smallest cutoff (or up to one gate per SIMD when the launch
// mixes sets: heaviest parts first, the tail workgroups hold the lightest gates), one
// wavefront per gate up to one gate per SIMD, the (4, 2) batched shape above that.
// Traced gates take latency mode only where it runs 8 waves (`narrow`).
enum class MixedKernel { split, one_wave, batch, pair };
// Whether every Goldilocks set in `sets` has a byte-digit blind-rotation gadget (8 digits
// of 8 bits or 13 of 5, as every set of param_set has): the mixed latency-mode and
// throughput kernels hold only those Goldilocks bodies; any other gadget goes to the
// one-wave kernel.
inline bool byte_digit_sets(int sets) {
  for (int set = 0; set < SET_COUNT; set++) {
    ParamSet p;
    if (!(sets & (1 << set)) || !param_set(set, &p) || !p.is_goldilocks) continue;
    if (!((p.br_bits == 8 && p.br_len == 8) || (p.br_bits == 5 && p.br_len == 13))) return false;
  }
  return true;
}
inline MixedKernel mixed_kernel_for(const Engine& E, int sets, uint32_t gates, bool traced,
                                    bool narrow = true, bool allow_pair = false) {
  if (!byte_digit_sets(sets)) return MixedKernel::one_wave;
  // pair mode: at most one gate per two CUs
  if (allow_pair && E.pair_mode && !traced && gates > 0 &&
      gates <= E.simd_slots / 8 && gates <= mixed_split_cutoff(E, sets))
    return MixedKernel::pair;
  const uint32_t cutoff = mixed_split_cutoff(E, sets);
  const bool mixes = (sets & (sets - 1)) != 0;
  // traced gates run the latency-mode kernel with 8 waves only
  if ((!traced || narrow) && (gates <= cutoff || (cutoff > 0 && mixes && gates <= E.simd_slots)))
    return MixedKernel::split;
  return gates <= E.simd_slots ? MixedKernel::one_wave : MixedKernel::batch;
}
inline uint32_t mixed_gpw(MixedKernel k) { return k == MixedKernel::batch ? kMixedBatchGpw : 1; }
// The kernel class of a launch over the parameter sets in `sets` (launch_mixed_batch's cls):
// one class when every part is Goldilocks with 8-bit digits, another Goldilocks set
// (Zama), or BabyBear; 0 when they mix.
inline int mixed_class(int sets) {
  if (sets == (1 << SET_GOLDILOCKS)) return 1;
  if (sets == (1 << SET_ZAMA)) return 2;
  if (sets == (1 << SET_BABYBEAR)) return 3;
  return 0;
}
inline hipError_t launch_mixed_kernel(MixedKernel k, const BootArgs* parts_dev, int nparts,
                                      uint32_t blocks, bool narrow, int sets, hipStream_t s,
                                      bool traced = false) {
  switch (k) {
    case MixedKernel::split:
      return launch_mixed_split(parts_dev, nparts, blocks, mixed_split_waves(narrow), traced, s);
    case MixedKernel::one_wave: return launch_mixed(parts_dev, nparts, blocks, s);
    case MixedKernel::pair:
      return launch_mixed_pair(parts_dev, nparts, blocks, mixed_class(sets), s);
    default: return launch_mixed_batch(parts_dev, nparts, blocks, mixed_class(sets), s);
  }
}

Engine& engine();
int ensure_engine(int device);

}  // namespace hb

struct hb_sk {
  hb::SecretKey sk;
};

struct hb_ek {
  hb::ParamSet p;
  std::vector<uint64_t> bsk_host, ksk_host;  // canonical layout (kept for export)
  void* bsk_dev = nullptr;
  void* ksk_dev = nullptr;
  void* lut_dev = nullptr;  // N words staging for hb_bootstrap
  std::mutex mu;            // guards the staging buffers below
  uint64_t* io_dev = nullptr;
  size_t io_cap = 0;  // gates
  void* acc_dev = nullptr;
  size_t acc_cap = 0;  // gates
  // Recorded after every launch that uses acc_dev; a launch on any stream first waits for
  // it, and ensure_acc synchronises with it before freeing, so kernels of calls on
  // different streams never share the accumulators.
  hipEvent_t acc_ev = nullptr;
  bool acc_ev_used = false;
  // Mux staging (6 rows of n+1 words per gate), guarded by acc_ev like acc_dev.
  uint64_t* mux_dev = nullptr;
  size_t mux_cap = 0;  // gates
  // Pair mode's flag words (kPairFlagWords per gate, zeroed once) and the epoch of its last
  // launch, guarded by acc_ev like acc_dev; the status word (epoch << 32 of the last launch
  // whose wait timed out) and the last epoch whose status the host has read.
  uint64_t* pair_flags = nullptr;
  size_t pair_cap = 0;  // gates
  uint32_t pair_epoch = 0;
  uint64_t* pair_status = nullptr;  // kPairStatusWords: timeout epoch, ticket counter
  uint32_t pair_checked = 0;
  uint64_t pair_tickets = 0;  // workgroups of the pair launches issued so far
};

namespace hb {
// Kernel arguments for `count` gates of `ek` (see BootArgs).
BootArgs make_args(const hb_ek* ek, int gate, size_t count, const uint64_t* c0,
                   const uint64_t* c1, const uint64_t* c2, const uint64_t* lut_dev, uint64_t* out,
                   void* acc);
}  // namespace hb
