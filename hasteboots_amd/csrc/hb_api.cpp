// C ABI of the engine (include/hb.h): handle management, key upload to HBM in the
// device layout, and launches of the HIP kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hb.h"
#include "hb_field.hpp"
#include "hb_internal.hpp"
#include "hb_kernels.hpp"
#include "hb_keys.hpp"
#include "hb_sample.hpp"

using namespace hb;

namespace {
thread_local std::string g_err;

// Per-gate scratch of hb_gate_batch_dev_noalloc (accumulator and byte digits, 4N words of
// the set's word size per gate), reserved on every GPU the library runs on for the life of
// the process. Host code uses the address hipGetSymbolAddress returns (Engine::
// noalloc_scratch), never this name.
static __device__ __align__(256) unsigned char scratch[HB_NOALLOC_SCRATCH_BYTES];
}  // namespace

namespace hb {

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

Engine& engine() {
  static Engine e;
  return e;
}

LaunchTimer& launch_timer() {
  thread_local LaunchTimer t;
  return t;
}

int timer_start(hipStream_t s) {
  LaunchTimer& t = launch_timer();
  if (!t.ev0) {
    HIP_TRY(hipEventCreate(&t.ev0));
    HIP_TRY(hipEventCreate(&t.ev1));
  }
  HIP_TRY(hipEventRecord(t.ev0, s));
  return HB_OK;
}

int timer_stop(hipStream_t s) {
  LaunchTimer& t = launch_timer();
  HIP_TRY(hipEventRecord(t.ev1, s));
  t.used = true;
  return HB_OK;
}

}  // namespace hb

namespace {

uint64_t to_device_word(const ParamSet& p, uint64_t v, uint64_t scale) {
  // scale: a canonical multiplier folded into the constant (e.g. N^-1).
  if (p.is_goldilocks) return Gl::to_mont(Gl::canon(Gl::mul(v, scale)));
  uint32_t x = (uint32_t)((v * scale) % p.q);
  return Bb::to_mont(x);
}

}  // namespace

namespace {

// Autotune cache: text lines "<set> <gates per workgroup> <waves per SIMD>", at the path
// named by the HB_AUTOTUNE_CACHE environment variable (no file, no cache).
const char* tune_cache_path() { return std::getenv("HB_AUTOTUNE_CACHE"); }

// The batched-kernel variants, in the order of hb_tune_result_t::ms.
const int kVariants[HB_TUNE_VARIANTS][2] = {{1, 1}, {2, 1}, {4, 1}, {2, 2}, {4, 2}, {4, 3}};

// Variant used for a batch of `count` gates of parameter set `set`; gpw = 0 selects the
// latency-mode kernel (split_waves wavefronts per gate).
void batch_variant(const Engine& E, int set, size_t count, int* gpw, int* wps) {
  if (count <= E.split_max[set]) {
    *gpw = 0, *wps = 1;
  } else if (count <= E.simd_slots) {
    *gpw = 1, *wps = 1;
  } else {
    *gpw = E.gpw[set], *wps = E.wps[set];
  }
}

bool valid_variant(int gpw, int wps) {
  for (const auto& v : kVariants)
    if (v[0] == gpw && v[1] == wps) return true;
  return false;
}

void load_tune_cache(Engine& E) {
  const char* path = tune_cache_path();
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int set, gpw, wps;
  while (std::fscanf(f, "%d %d %d", &set, &gpw, &wps) == 3) {
    if (set >= 0 && set < 3 && valid_variant(gpw, wps)) {
      E.gpw[set] = gpw;
      E.wps[set] = wps;
      E.tuned[set] = true;
    }
  }
  std::fclose(f);
}

void save_tune_cache(const Engine& E) {
  const char* path = tune_cache_path();
  if (!path) return;
  FILE* f = std::fopen(path, "w");
  if (!f) return;
  for (int set = 0; set < 3; set++)
    if (E.tuned[set]) std::fprintf(f, "%d %d %d\n", set, E.gpw[set], E.wps[set]);
  std::fclose(f);
}

}  // namespace

int hb::ensure_engine(int device) {
  Engine& E = engine();
  if (E.device >= 0) {
    if (device != E.device) return fail(HB_ERR_ARG, "engine already bound to device %d", E.device);
    HIP_TRY(hipSetDevice(E.device));
    return HB_OK;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return fail(HB_ERR_DEVICE, "no HIP device");
  if (device < 0 || device >= n) return fail(HB_ERR_ARG, "device %d out of range (%d)", device, n);
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&E.mixed_ev, hipEventDisableTiming));
  // Twiddle tables.
  for (int f = 0; f < 2; f++) {
    ParamSet p;
    param_set(f == 1 ? SET_GOLDILOCKS : SET_BABYBEAR, &p);
    for (int lg = 10; lg <= 11; lg++) {
      HostNtt ntt;
      ntt.init(p, lg);
      for (int inv = 0; inv < 2; inv++) {
        const std::vector<uint64_t>& src = inv ? ntt.psi_inv_rev : ntt.psi_rev;
        size_t bytes = src.size() * (f ? 8 : 4);
        void* d = nullptr;
        HIP_TRY(hipMalloc(&d, bytes));
        if (f) {
          std::vector<uint64_t> w(src.size());
          for (size_t k = 0; k < src.size(); k++) w[k] = Gl::to_mont(src[k]);
          HIP_TRY(hipMemcpy(d, w.data(), bytes, hipMemcpyHostToDevice));
        } else {
          std::vector<uint32_t> w(src.size());
          for (size_t k = 0; k < src.size(); k++) w[k] = Bb::to_mont((uint32_t)src[k]);
          HIP_TRY(hipMemcpy(d, w.data(), bytes, hipMemcpyHostToDevice));
        }
        E.tables.tw[f][lg - 10][inv] = d;
      }
      E.tables.ninv[f][lg - 10] = f ? Gl::to_mont(ntt.n_inv) : Bb::to_mont((uint32_t)ntt.n_inv);
    }
  }
  {  // twisted digit-transform table (ntt_fwd_digits): beta_g^p, then the pass-3 twiddles
    ParamSet p;
    param_set(SET_GOLDILOCKS, &p);
    HostNtt ntt;
    ntt.init(p, 10);
    auto psi_pow = [&](uint64_t e) {
      uint64_t r = 1, b = ntt.psi;
      for (e %= 2048; e; e >>= 1, b = ntt.mul(b, b))
        if (e & 1) r = ntt.mul(r, b);
      return r;
    };
    auto rev = [](uint32_t x, int bits) {
      uint32_t r = 0;
      for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
      return r;
    };
    std::vector<uint64_t> t(kTwxWords, 0);
    for (uint32_t pos = 0; pos < 1024; pos++)
      t[68 * (pos >> 6) + (pos & 63)] =
          Gl::to_mont(psi_pow((uint64_t)rev(16 + (pos >> 6), 5) * (pos & 63)));
    for (uint32_t ip = 0; ip < 16; ip++) {
      t[1088 + ip] = Gl::to_mont(psi_pow(rev(256 + ip, 10) + 2048 - 2));
      for (uint32_t h = 0; h < 2; h++)
        t[1104 + 2 * ip + h] = Gl::to_mont(psi_pow(rev(512 + 2 * ip + h, 10) + 2048 - 1));
    }
    HIP_TRY(hipMalloc(&E.tables.twx, t.size() * 8));
    HIP_TRY(hipMemcpy(E.tables.twx, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  }
  E.tables.ready = true;
  E.device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      E.simd_slots = (uint32_t)prop.multiProcessorCount * 4;
    for (int set = 0; set < SET_COUNT; set++) E.split_max[set] = default_split_max(set, E.simd_slots);
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
      E.wall_khz = (uint32_t)khz;
  }
  {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(HB_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    E.noalloc_scratch = p;
  }
  load_tune_cache(E);
  return HB_OK;
}

namespace {

uint32_t ilog2u(uint32_t x) {
  uint32_t r = 0;
  while ((1u << r) < x) r++;
  return r;
}

}  // namespace

namespace {

// Makes canonical keys resident in the device form the bootstrap kernels read: bsk_c and
// ksk_c are device arrays of canonical words, converted on `s` (× N^-1, Montgomery).
int install_keys(hb_ek* ek, const uint64_t* bsk_c, const uint64_t* ksk_c, hipStream_t s) {
  const ParamSet& p = ek->p;
  const size_t wb = p.is_goldilocks ? 8 : 4;
  const uint64_t nb = bsk_words(p), nk = ksk_words(p);
  HostNtt ntt;
  ntt.init(p, ilog2u(p.N));
  HIP_TRY(hipMalloc(&ek->bsk_dev, nb * wb));
  HIP_TRY(hipMalloc(&ek->ksk_dev, nk * wb));
  HIP_TRY(launch_keys_to_device(bsk_c, ek->bsk_dev, nb, ntt.n_inv, p.is_goldilocks, p.q, s));
  HIP_TRY(launch_keys_to_device(ksk_c, ek->ksk_dev, nk, ntt.n_inv, p.is_goldilocks, p.q, s));
  HIP_TRY(hipMalloc(&ek->lut_dev, (size_t)p.N * 8));
  HIP_TRY(hipEventCreateWithFlags(&ek->acc_ev, hipEventDisableTiming));
  return HB_OK;
}

// Device scratch freed when the scope ends (after the stream's work on it is done: the
// callers synchronise first).
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

// From the canonical host copies in ek->bsk_host / ksk_host.
int upload_keys(hb_ek* ek) {
  Engine& E = engine();
  DevBuf bsk, ksk;
  HIP_TRY(hipMalloc(&bsk.p, ek->bsk_host.size() * 8));
  HIP_TRY(hipMalloc(&ksk.p, ek->ksk_host.size() * 8));
  HIP_TRY(hipMemcpyAsync(bsk.p, ek->bsk_host.data(), ek->bsk_host.size() * 8,
                         hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipMemcpyAsync(ksk.p, ek->ksk_host.data(), ek->ksk_host.size() * 8,
                         hipMemcpyHostToDevice, E.stream));
  int rc = install_keys(ek, static_cast<uint64_t*>(bsk.p), static_cast<uint64_t*>(ksk.p), E.stream);
  HIP_TRY(hipStreamSynchronize(E.stream));
  return rc;
}

// BinaryBlindRotationKey::generate + KeySwitchingRLWEKey::generate on the GPU: the words
// of generate_bsk / generate_ksk (hb_keys.cpp), canonical layout, into device arrays.
int keygen_device(const SecretKey& sk, uint64_t* bsk_c, uint64_t* ksk_c, hipStream_t s) {
  const ParamSet& p = sk.p;
  Engine& E = engine();
  const uint32_t N = p.N;
  if (N != 1024) return fail(HB_ERR_ARG, "device key generation needs N = 1024");
  HostNtt ntt;
  ntt.init(p, ilog2u(N));
  std::vector<uint64_t> lwe_ntt(N, 0);
  for (uint32_t k = 0; k < p.n; k++) lwe_ntt[k] = sk.lwe_sk[k];
  ntt.fwd(lwe_ntt.data());
  std::vector<uint8_t> bits(p.n);
  for (uint32_t k = 0; k < p.n; k++) bits[k] = (uint8_t)sk.lwe_sk[k];
  const uint32_t bsk_rows = p.n * 2 * p.br_len;
  DevBuf err, ring, lwe, sk_bits;
  HIP_TRY(hipMalloc(&err.p, (size_t)bsk_rows * N * 8));
  HIP_TRY(hipMalloc(&ring.p, (size_t)N * 8));
  HIP_TRY(hipMalloc(&lwe.p, (size_t)N * 8));
  HIP_TRY(hipMalloc(&sk_bits.p, p.n));
  HIP_TRY(hipMemcpyAsync(ring.p, sk.ring_sk_ntt.data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(lwe.p, lwe_ntt.data(), (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(sk_bits.p, bits.data(), p.n, hipMemcpyHostToDevice, s));
  KeygenArgs a{};
  derive_key(sk.seed, a.key);
  a.N = N;
  a.is_goldilocks = p.is_goldilocks;
  a.q = p.q;
  a.lwe_sk = static_cast<const uint8_t*>(sk_bits.p);
  const void* tw = E.tables.tw[p.is_goldilocks ? 1 : 0][0][0];
  const uint32_t cus = E.simd_slots / 4;
  a.kind = STREAM_BSK, a.rows = bsk_rows, a.L = p.br_len, a.basis_bits = p.br_bits;
  a.is_bsk = 1, a.sigma = p.ring_sigma;
  const uint64_t* ring_ntt = static_cast<const uint64_t*>(ring.p);
  HIP_TRY(launch_keygen(a, bsk_c, static_cast<uint64_t*>(err.p), ring_ntt, ring_ntt, tw, cus, s));
  a.kind = STREAM_KSK, a.rows = p.ks_len, a.L = 1, a.basis_bits = p.ks_bits;
  a.is_bsk = 0, a.sigma = p.ks_sigma;
  HIP_TRY(launch_keygen(a, ksk_c, static_cast<uint64_t*>(err.p),
                        static_cast<const uint64_t*>(lwe.p), ring_ntt, tw, cus, s));
  HIP_TRY(hipStreamSynchronize(s));
  return HB_OK;
}

// The next launch on `s` that uses ek->acc_dev starts after every earlier one (on any
// stream) has finished with it.
int acc_acquire(hb_ek* ek, hipStream_t s) {
  if (ek->acc_ev_used) HIP_TRY(hipStreamWaitEvent(s, ek->acc_ev, 0));
  return HB_OK;
}
int acc_release(hb_ek* ek, hipStream_t s) {
  HIP_TRY(hipEventRecord(ek->acc_ev, s));
  ek->acc_ev_used = true;
  return HB_OK;
}

int ensure_acc(hb_ek* ek, size_t count) {
  if (count <= ek->acc_cap) return HB_OK;
  if (ek->acc_ev_used) HIP_TRY(hipEventSynchronize(ek->acc_ev));
  if (ek->acc_dev) HIP_TRY(hipFree(ek->acc_dev));
  ek->acc_dev = nullptr;
  ek->acc_cap = 0;
  const size_t wb = ek->p.is_goldilocks ? 8 : 4;
  HIP_TRY(hipMalloc(&ek->acc_dev, count * kScratchWordsPerGate * wb));
  ek->acc_cap = count;
  return HB_OK;
}

int ensure_io(hb_ek* ek, size_t count) {
  if (count <= ek->io_cap) return HB_OK;
  if (ek->io_dev) HIP_TRY(hipFree(ek->io_dev));
  ek->io_dev = nullptr;
  ek->io_cap = 0;
  HIP_TRY(hipMalloc((void**)&ek->io_dev, count * 4 * (ek->p.n + 1) * sizeof(uint64_t)));
  ek->io_cap = count;
  return HB_OK;
}

void lut_params(const ParamSet& p, int gate, BootArgs& a) {
  const uint32_t N = p.N;
  const uint64_t q8 = p.q >> 3, nq8 = p.q - q8;
  uint32_t mid;
  uint64_t lo, hi;
  switch (gate) {
    case GATE_NAND: mid = (N >> 1) + (N >> 2); lo = q8; hi = nq8; break;
    case GATE_AND:
    case GATE_MAJORITY: mid = (N >> 1) + (N >> 2); lo = nq8; hi = q8; break;
    case GATE_OR: mid = N >> 2; lo = nq8; hi = q8; break;
    case GATE_NOR: mid = N >> 2; lo = q8; hi = nq8; break;
    case GATE_XOR: mid = N >> 1; lo = nq8; hi = q8; break;
    case GATE_XNOR: mid = N >> 1; lo = q8; hi = nq8; break;
    default: mid = N; lo = hi = 0;
  }
  a.lut_mid = mid;
  a.lut_lo = lo;
  a.lut_hi = hi;
}

}  // namespace

BootArgs hb::make_args(const hb_ek* ek, int gate, size_t count, const uint64_t* c0,
                       const uint64_t* c1, const uint64_t* c2, const uint64_t* lut_dev,
                       uint64_t* out, void* acc) {
  const ParamSet& p = ek->p;
  Engine& E = engine();
  BootArgs a{};
  a.bsk = ek->bsk_dev;
  a.ksk = ek->ksk_dev;
  const int f = p.is_goldilocks ? 1 : 0;
  a.tw_fwd = f ? E.tables.twx : E.tables.tw[f][0][0];  // the digit transforms' table
  a.tw_inv = E.tables.tw[f][0][1];
  a.c0 = c0;
  a.c1 = c1;
  a.c2 = c2;
  a.lut = lut_dev;
  a.out = out;
  a.acc = acc;
  a.qf = (double)p.q;
  a.q8 = p.q >> 3;
  a.n = p.n;
  a.br_len = p.br_len;
  a.br_bits = p.br_bits;
  a.ks_len = p.ks_len;
  a.ks_bits = p.ks_bits;
  a.count = (uint32_t)count;
  a.gate = gate;
  a.is_goldilocks = p.is_goldilocks;
  a.first_block = 0;
  lut_params(p, gate, a);
  return a;
}

namespace {

// Launches the untraced batch `a` of parameter set p in the shape batch_variant picks. A
// batch of the tuned batched variant whose last round of resident gates (waves per SIMD x
// SIMDs) would be partly filled runs its whole rounds batched and the rest, when latency
// mode takes that many, in latency mode after them: Goldilocks 2560 gates 157 ms in one
// batched launch against 104 + 28 ms. The parts write disjoint rows; the latency-mode
// kernel keeps its accumulators in LDS.
hipError_t launch_shaped(const Engine& E, const ParamSet& p, const BootArgs& a, hipStream_t s) {
  const size_t count = a.count;
  int gpw, wps;
  batch_variant(E, p.set, count, &gpw, &wps);
  auto split = [&](const BootArgs& b) {
    return launch_bootstrap_split(
        b, p.is_goldilocks, split_waves(E, p.is_goldilocks, b.count, split8_fits(p.br_len, p.br_bits)),
        s);
  };
  if (gpw == 0) return split(a);
  const size_t round = (size_t)wps * E.simd_slots, rest = count % round;
  if (gpw > 1 && wps <= 2 && !a.trace && !a.idx && count > round && rest > 0 &&
      rest <= E.split_max[p.set]) {
    BootArgs head = a, tail = a;
    head.count = (uint32_t)(count - rest);
    tail.count = (uint32_t)rest;
    const size_t off = (size_t)head.count * (p.n + 1);
    tail.c0 = a.c0 + off;
    if (a.c1) tail.c1 = a.c1 + off;
    if (a.c2) tail.c2 = a.c2 + off;
    tail.out = a.out + off;
    hipError_t e = launch_bootstrap(head, p.is_goldilocks, gpw, wps, s);
    return e != hipSuccess ? e : split(tail);
  }
  return launch_bootstrap(a, p.is_goldilocks, gpw, wps, s);
}

// Pair mode for this batch: Goldilocks with a byte-digit gadget (8 x 8 bits, 13 x 5 bits)
// or BabyBear (at most 4 digits), within the latency-mode cutoff and one gate per two CUs
// (bootstrap_gate_pair).
bool pair_batch(const Engine& E, const ParamSet& p, size_t count) {
  const bool ok = p.is_goldilocks
                      ? (p.br_bits == 8 && p.br_len == 8) || (p.br_bits == 5 && p.br_len == 13)
                      : p.br_len <= 4;
  return E.pair_mode && ok && count <= E.split_max[p.set] && count <= E.simd_slots / 8;
}

int ensure_pair_flags(hb_ek* ek, size_t count) {
  if (!ek->pair_status) {
    HIP_TRY(hipMalloc((void**)&ek->pair_status, kPairStatusWords * sizeof(uint64_t)));
    HIP_TRY(hipMemset(ek->pair_status, 0, kPairStatusWords * sizeof(uint64_t)));
    ek->pair_tickets = 0;
  }
  if (count <= ek->pair_cap) return HB_OK;
  if (ek->acc_ev_used) HIP_TRY(hipEventSynchronize(ek->acc_ev));
  if (ek->pair_flags) HIP_TRY(hipFree(ek->pair_flags));
  ek->pair_flags = nullptr;
  ek->pair_cap = 0;
  const size_t bytes = count * kPairFlagWords * sizeof(uint64_t);
  HIP_TRY(hipMalloc((void**)&ek->pair_flags, bytes));
  HIP_TRY(hipMemset(ek->pair_flags, 0, bytes));  // epochs start at 1
  ek->pair_cap = count;
  return HB_OK;
}

int launch_gates(hb_ek* ek, int gate, size_t count, const uint64_t* c0, const uint64_t* c1,
                 const uint64_t* c2, const uint64_t* lut_dev, uint64_t* out, hipStream_t s) {
  const ParamSet& p = ek->p;
  Engine& E = engine();
  int rc = ensure_acc(ek, count);
  if (rc) return rc;
  const bool pair = pair_batch(E, p, count);
  if (pair) {
    rc = ensure_pair_flags(ek, count);
    if (rc) return rc;
  }
  BootArgs a = make_args(ek, gate, count, c0, c1, c2, lut_dev, out, ek->acc_dev);
  rc = acc_acquire(ek, s);
  if (rc) return rc;
  rc = timer_start(s);
  if (rc) return rc;
  bool done = false;
  if (pair) {
    a.pair_flags = ek->pair_flags;
    a.pair_epoch = ++ek->pair_epoch;
    a.pair_status = ek->pair_status;
    a.pair_ticket0 = ek->pair_tickets;
    a.pair_wait_ticks = E.pair_wait_ticks();
    HIP_TRY(launch_bootstrap_pair(a, s));
    ek->pair_tickets += 2 * (uint64_t)count;
    done = true;
  }
  if (!done) HIP_TRY(launch_shaped(E, p, a, s));
  rc = timer_stop(s);
  if (rc) return rc;
  return acc_release(ek, s);
}

// Pair-mode status of ek's launches since the last check, read on stream s once the
// launches are complete: HB_ERR_DEVICE if a wait of one of them timed out (its outputs are
// then wrong).
int pair_check(hb_ek* ek, hipStream_t s) {
  if (!ek->pair_status || ek->pair_checked == ek->pair_epoch) return HB_OK;
  uint64_t v = 0;
  HIP_TRY(hipMemcpyAsync(&v, ek->pair_status, sizeof v, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const uint32_t seen = ek->pair_checked;
  ek->pair_checked = ek->pair_epoch;
  if ((uint32_t)(v >> 32) > seen)
    return fail(HB_ERR_DEVICE, "pair mode: an exchange wait of launch %u timed out; the outputs "
                "of that launch are invalid", (unsigned)(v >> 32));
  return HB_OK;
}

int ensure_mux(hb_ek* ek, size_t count) {
  if (count <= ek->mux_cap) return HB_OK;
  if (ek->acc_ev_used) HIP_TRY(hipEventSynchronize(ek->acc_ev));
  if (ek->mux_dev) HIP_TRY(hipFree(ek->mux_dev));
  ek->mux_dev = nullptr;
  ek->mux_cap = 0;
  HIP_TRY(hipMalloc((void**)&ek->mux_dev, count * 6 * (ek->p.n + 1) * sizeof(uint64_t)));
  ek->mux_cap = count;
  return HB_OK;
}

// Evaluator::mux (tfhe/src/bfhe/evaluate.rs:245-263) for `count` gates on stream s, select
// c0, then c1, else c2. Rows of n+1 words in the key's mux staging W: [0, c) c0, [c, 2c)
// not(c0), [2c, 3c) c1, [3c, 4c) c2; one launch of 2c AND gates (c0 & c1, not(c0) & c2) into
// [4c, 6c); then one launch of c OR-LUT bootstraps of the pairwise sums into outs (the OR
// gate's combination is the same component-wise add).
int launch_mux(hb_ek* ek, size_t count, const uint64_t* c0, const uint64_t* c1,
               const uint64_t* c2, uint64_t* outs, hipStream_t s) {
  int rc = ensure_mux(ek, count);
  if (rc) return rc;
  rc = acc_acquire(ek, s);  // the previous user of W may run on another stream
  if (rc) return rc;
  const ParamSet& p = ek->p;
  const size_t n1 = p.n + 1, rows = count * n1;
  uint64_t* W = ek->mux_dev;
  HIP_TRY(hipMemcpyAsync(W, c0, rows * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(launch_lwe_not(c0, W + rows, (uint32_t)count, (uint32_t)n1, p.q, p.is_goldilocks, s));
  HIP_TRY(hipMemcpyAsync(W + 2 * rows, c1, rows * 8, hipMemcpyDeviceToDevice, s));
  HIP_TRY(hipMemcpyAsync(W + 3 * rows, c2, rows * 8, hipMemcpyDeviceToDevice, s));
  rc = launch_gates(ek, GATE_AND, 2 * count, W, W + 2 * rows, nullptr, nullptr, W + 4 * rows, s);
  if (rc) return rc;
  return launch_gates(ek, GATE_OR, count, W + 4 * rows, W + 5 * rows, nullptr, nullptr, outs, s);
}

// Gate kinds the batch entry points take: the seven single-bootstrap gates and mux.
bool batch_gate_ok(int gate) { return (gate >= 0 && gate <= GATE_MAJORITY) || gate == GATE_MUX; }
bool needs_c2(int gate) { return gate == GATE_MAJORITY || gate == GATE_MUX; }

// Waits for the calling thread's last timed launch (the host-buffer calls return only
// after their results are on the host).
int record_ms(hipStream_t s) {
  (void)s;
  LaunchTimer& t = launch_timer();
  HIP_TRY(hipEventSynchronize(t.ev1));
  return HB_OK;
}

}  // namespace

extern "C" {

int hb_params(int set, hb_params_t* out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  memset(out, 0, sizeof *out);
  out->set = p.set;
  out->is_goldilocks = p.is_goldilocks;
  out->n = p.n;
  out->N = p.N;
  out->t = p.t;
  out->br_bits = p.br_bits;
  out->br_len = p.br_len;
  out->ks_bits = p.ks_bits;
  out->ks_len = p.ks_len;
  out->q = p.q;
  out->lwe_sigma = p.lwe_sigma;
  out->ring_sigma = p.ring_sigma;
  out->ks_sigma = p.ks_sigma;
  return HB_OK;
}

int hb_key_words(int set, uint64_t* bw, uint64_t* kw) {
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  if (bw) *bw = bsk_words(p);
  if (kw) *kw = ksk_words(p);
  return HB_OK;
}

int hb_sk_create(int set, uint64_t seed, hb_sk_t** out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  auto* h = new (std::nothrow) hb_sk;
  if (!h) return fail(HB_ERR_NOMEM, "allocation failed");
  if (!secret_key_create(set, seed, &h->sk)) {
    delete h;
    return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  }
  *out = h;
  return HB_OK;
}

int hb_sk_destroy(hb_sk_t* sk) {
  delete sk;
  return HB_OK;
}

int hb_sk_export(const hb_sk_t* sk, uint64_t* lwe_sk, uint64_t* ring_sk) {
  if (!sk) return fail(HB_ERR_ARG, "null sk");
  if (lwe_sk) memcpy(lwe_sk, sk->sk.lwe_sk.data(), sk->sk.lwe_sk.size() * 8);
  if (ring_sk) memcpy(ring_sk, sk->sk.ring_sk.data(), sk->sk.ring_sk.size() * 8);
  return HB_OK;
}

int hb_encrypt(hb_sk_t* sk, uint32_t bit, uint64_t* out) {
  if (!sk || !out) return fail(HB_ERR_ARG, "null argument");
  if (bit >= sk->sk.p.t) return fail(HB_ERR_ARG, "message %u exceeds plaintext space", bit);
  encrypt(sk->sk, bit, out);
  return HB_OK;
}

int hb_decrypt(const hb_sk_t* sk, const uint64_t* c, uint32_t* bit, uint64_t* noise) {
  if (!sk || !c || !bit) return fail(HB_ERR_ARG, "null argument");
  decrypt(sk->sk, c, bit, noise);
  return HB_OK;
}

int hb_ek_import(int set, const uint64_t* bsk, const uint64_t* ksk, int device, hb_ek_t** out) {
  if (!bsk || !ksk || !out) return fail(HB_ERR_ARG, "null argument");
  ParamSet p;
  if (!param_set(set, &p)) return fail(HB_ERR_ARG, "unknown parameter set %d", set);
  int rc = ensure_engine(device);
  if (rc) return rc;
  auto* ek = new (std::nothrow) hb_ek;
  if (!ek) return fail(HB_ERR_NOMEM, "allocation failed");
  ek->p = p;
  ek->bsk_host.assign(bsk, bsk + bsk_words(p));
  ek->ksk_host.assign(ksk, ksk + ksk_words(p));
  rc = upload_keys(ek);
  if (rc) {
    hb_ek_destroy(ek);
    return rc;
  }
  *out = ek;
  return HB_OK;
}

int hb_ek_create(const hb_sk_t* sk, int device, hb_ek_t** out) {
  if (!sk || !out) return fail(HB_ERR_ARG, "null argument");
  int rc = ensure_engine(device);
  if (rc) return rc;
  std::unique_ptr<hb_ek> ek(new (std::nothrow) hb_ek);
  if (!ek) return fail(HB_ERR_NOMEM, "allocation failed");
  ek->p = sk->sk.p;
  ek->bsk_host.resize(bsk_words(ek->p));
  ek->ksk_host.resize(ksk_words(ek->p));
  Engine& E = engine();
  auto make = [&]() -> int {
    DevBuf bsk, ksk;
    HIP_TRY(hipMalloc(&bsk.p, ek->bsk_host.size() * 8));
    HIP_TRY(hipMalloc(&ksk.p, ek->ksk_host.size() * 8));
    uint64_t* b = static_cast<uint64_t*>(bsk.p);
    uint64_t* k = static_cast<uint64_t*>(ksk.p);
    int r = keygen_device(sk->sk, b, k, E.stream);
    if (r) return r;
    r = install_keys(ek.get(), b, k, E.stream);
    if (r) return r;
    HIP_TRY(hipMemcpyAsync(ek->bsk_host.data(), b, ek->bsk_host.size() * 8, hipMemcpyDeviceToHost,
                           E.stream));
    HIP_TRY(hipMemcpyAsync(ek->ksk_host.data(), k, ek->ksk_host.size() * 8, hipMemcpyDeviceToHost,
                           E.stream));
    HIP_TRY(hipStreamSynchronize(E.stream));
    return HB_OK;
  };
  rc = make();
  if (rc) {
    hb_ek_destroy(ek.release());
    return rc;
  }
  *out = ek.release();
  return HB_OK;
}

int hb_keygen_dev(const hb_sk_t* sk, uint64_t* bsk_dev, uint64_t* ksk_dev, void* stream) {
  if (!sk || !bsk_dev || !ksk_dev) return fail(HB_ERR_ARG, "null argument");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  hipStream_t s = stream ? (hipStream_t)stream : engine().stream;
  return keygen_device(sk->sk, bsk_dev, ksk_dev, s);
}

int hb_ek_export(const hb_ek_t* ek, uint64_t* bsk, uint64_t* ksk) {
  if (!ek) return fail(HB_ERR_ARG, "null ek");
  if (bsk) memcpy(bsk, ek->bsk_host.data(), ek->bsk_host.size() * 8);
  if (ksk) memcpy(ksk, ek->ksk_host.data(), ek->ksk_host.size() * 8);
  return HB_OK;
}

int hb_keygen_host(const hb_sk_t* sk, uint64_t* bsk, uint64_t* ksk) {
  if (!sk) return fail(HB_ERR_ARG, "null sk");
  if (bsk) generate_bsk(sk->sk, bsk);
  if (ksk) generate_ksk(sk->sk, ksk);
  return HB_OK;
}

int hb_ek_destroy(hb_ek_t* ek) {
  if (!ek) return HB_OK;
  if (ek->bsk_dev) (void)hipFree(ek->bsk_dev);
  if (ek->ksk_dev) (void)hipFree(ek->ksk_dev);
  if (ek->lut_dev) (void)hipFree(ek->lut_dev);
  if (ek->io_dev) (void)hipFree(ek->io_dev);
  if (ek->acc_ev_used) (void)hipEventSynchronize(ek->acc_ev);
  if (ek->acc_dev) (void)hipFree(ek->acc_dev);
  if (ek->mux_dev) (void)hipFree(ek->mux_dev);
  if (ek->pair_flags) (void)hipFree(ek->pair_flags);
  if (ek->pair_status) (void)hipFree(ek->pair_status);
  if (ek->acc_ev) (void)hipEventDestroy(ek->acc_ev);
  delete ek;
  return HB_OK;
}

int hb_ek_params(const hb_ek_t* ek, hb_params_t* out) {
  if (!ek) return fail(HB_ERR_ARG, "null ek");
  return hb_params(ek->p.set, out);
}

int hb_gate_batch_dev(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0, const uint64_t* c1,
                      const uint64_t* c2, uint64_t* outs, void* stream) {
  if (!ek || !c0 || !outs) return fail(HB_ERR_ARG, "null argument");
  if (!batch_gate_ok(gate)) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (!c1) return fail(HB_ERR_ARG, "gate needs two inputs");
  if (needs_c2(gate) && !c2) return fail(HB_ERR_ARG, "gate %d needs three inputs", gate);
  if (count == 0) return HB_OK;
  if (count > 0x3FFFFFFF) return fail(HB_ERR_ARG, "count too large");
  hipStream_t s = stream ? (hipStream_t)stream : engine().stream;
  std::lock_guard<std::mutex> g(ek->mu);
  if (gate == GATE_MUX) return launch_mux(ek, count, c0, c1, c2, outs, s);
  return launch_gates(ek, gate, count, c0, c1, c2, nullptr, outs, s);
}

int hb_gate_batch_dev_noalloc(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0,
                              const uint64_t* c1, const uint64_t* c2, uint64_t* outs,
                              void* stream) {
  if (!ek || !c0 || !c1 || !outs) return fail(HB_ERR_ARG, "null argument");
  if (gate < 0 || gate > GATE_MAJORITY) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (gate == GATE_MAJORITY && !c2) return fail(HB_ERR_ARG, "majority needs three inputs");
  if (count == 0) return HB_OK;
  const ParamSet& p = ek->p;
  const size_t per_gate = kScratchWordsPerGate * (p.is_goldilocks ? 8 : 4);
  const size_t max_gates = HB_NOALLOC_SCRATCH_BYTES / per_gate;
  if (count > max_gates)
    return fail(HB_ERR_CAPACITY, "%zu gates exceed the fixed scratch (%zu gates of this set)",
                count, max_gates);
  Engine& E = engine();
  if (!E.noalloc_scratch) return fail(HB_ERR_STATE, "engine not set up");
  hipStream_t s = stream ? (hipStream_t)stream : E.stream;
  // The kernels write every scratch word they read (the accumulator is initialised from
  // the LUT, the digits from the rotated accumulator), so nothing is cleared first.
  BootArgs a = make_args(ek, gate, count, c0, c1, c2, nullptr, outs, E.noalloc_scratch);
  HIP_TRY(launch_shaped(E, p, a, s));
  return HB_OK;
}

int hb_gate_batch(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s, const uint64_t* c1s,
                  const uint64_t* c2s, uint64_t* outs) {
  if (!ek || !c0s || !c1s || !outs) return fail(HB_ERR_ARG, "null argument");
  if (!batch_gate_ok(gate)) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (needs_c2(gate) && !c2s) return fail(HB_ERR_ARG, "gate %d needs three inputs", gate);
  if (count == 0) return HB_OK;
  if (count > 0x3FFFFFFF) return fail(HB_ERR_ARG, "count too large");
  Engine& E = engine();
  std::lock_guard<std::mutex> g(ek->mu);
  int rc = ensure_io(ek, count);
  if (rc) return rc;
  const size_t n1 = ek->p.n + 1, bytes = count * n1 * 8;
  uint64_t* d0 = ek->io_dev;
  uint64_t* d1 = d0 + count * n1;
  uint64_t* d2 = d1 + count * n1;
  uint64_t* dout = d2 + count * n1;
  HIP_TRY(hipMemcpyAsync(d0, c0s, bytes, hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipMemcpyAsync(d1, c1s, bytes, hipMemcpyHostToDevice, E.stream));
  if (c2s) HIP_TRY(hipMemcpyAsync(d2, c2s, bytes, hipMemcpyHostToDevice, E.stream));
  rc = gate == GATE_MUX ? launch_mux(ek, count, d0, d1, d2, dout, E.stream)
                        : launch_gates(ek, gate, count, d0, d1, c2s ? d2 : nullptr, nullptr, dout,
                                       E.stream);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(outs, dout, bytes, hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  rc = pair_check(ek, E.stream);
  if (rc) return rc;
  return record_ms(E.stream);
}

int hb_mixed_batch_dev(const hb_batch_part_t* parts, size_t nparts, void* stream) {
  if (!parts && nparts) return fail(HB_ERR_ARG, "null parts");
  Engine& E = engine();
  std::vector<size_t> order;
  uint64_t total = 0;
  for (size_t i = 0; i < nparts; i++) {
    const hb_batch_part_t& q = parts[i];
    if (!q.ek) return fail(HB_ERR_ARG, "part %zu: null ek", i);
    if (q.gate < 0 || q.gate > GATE_MAJORITY) return fail(HB_ERR_ARG, "part %zu: bad gate", i);
    if (q.count == 0) continue;
    if (!q.c0s || !q.c1s || !q.outs) return fail(HB_ERR_ARG, "part %zu: null buffer", i);
    if (q.gate == GATE_MAJORITY && !q.c2s) return fail(HB_ERR_ARG, "part %zu: majority needs c2", i);
    total += q.count;
    order.push_back(i);
  }
  if (total == 0) return HB_OK;
  if (total > 0x7FFFFFFF) return fail(HB_ERR_ARG, "mixed batch too large");
  // Heaviest parameter sets first so the light gates fill the tail of the grid.
  auto cost = [&](size_t i) {
    const ParamSet& p = parts[i].ek->p;
    return (double)(p.br_len + 2) * (p.is_goldilocks ? 2.0 : 1.0);
  };
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
  std::lock_guard<std::mutex> g(E.mu);
  const size_t per_gate = kScratchWordsPerGate * 8;
  const bool grow = E.mixed_acc_bytes < total * per_gate || E.mixed_parts_cap < order.size();
  if (grow && E.mixed_ev_used) HIP_TRY(hipEventSynchronize(E.mixed_ev));
  if (E.mixed_acc_bytes < total * per_gate) {
    if (E.mixed_acc) HIP_TRY(hipFree(E.mixed_acc));
    E.mixed_acc = nullptr;
    E.mixed_acc_bytes = 0;
    HIP_TRY(hipMalloc(&E.mixed_acc, total * per_gate));
    E.mixed_acc_bytes = total * per_gate;
  }
  if (E.mixed_parts_cap < order.size()) {
    if (E.mixed_parts) HIP_TRY(hipFree(E.mixed_parts));
    E.mixed_parts = nullptr;
    E.mixed_parts_cap = 0;
    HIP_TRY(hipMalloc((void**)&E.mixed_parts, order.size() * sizeof(BootArgs)));
    E.mixed_parts_cap = order.size();
  }
  std::vector<BootArgs> args;
  uint32_t first = 0;
  for (size_t i : order) {
    const hb_batch_part_t& q = parts[i];
    if (q.ek->p.N != 1024) return fail(HB_ERR_ARG, "part %zu: unsupported ring dimension", i);
    BootArgs a = make_args(q.ek, q.gate, q.count, q.c0s, q.c1s, q.c2s, nullptr, q.outs,
                           (char*)E.mixed_acc + (size_t)first * per_gate);
    first += (uint32_t)q.count;
    args.push_back(a);
  }
  bool narrow = true;
  int sets = 0;
  for (const BootArgs& a : args) narrow = narrow && split8_fits(a.br_len, a.br_bits);
  for (size_t i : order) sets |= 1 << parts[i].ek->p.set;
  // Measured, 348 BabyBear + 56 Goldilocks + 574 Zama NANDs: 104.5 ms in latency mode
  // against 113.4 ms with one wavefront per gate.
  const MixedKernel kind = mixed_kernel_for(E, sets, first, false);
  const uint32_t blocks = assign_blocks(args.data(), args.size(), mixed_gpw(kind));
  hipStream_t s = stream ? (hipStream_t)stream : E.stream;
  // the part table and accumulators are free once the previous mixed launch is done
  if (E.mixed_ev_used) HIP_TRY(hipStreamWaitEvent(s, E.mixed_ev, 0));
  HIP_TRY(hipMemcpyAsync(E.mixed_parts, args.data(), args.size() * sizeof(BootArgs),
                         hipMemcpyHostToDevice, s));
  HIP_TRY(hipStreamSynchronize(s));  // `args` is pageable host memory
  int rc = timer_start(s);
  if (rc) return rc;
  HIP_TRY(launch_mixed_kernel(kind, E.mixed_parts, (int)args.size(), blocks, narrow, sets, s));
  rc = timer_stop(s);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(E.mixed_ev, s));
  E.mixed_ev_used = true;
  return HB_OK;
}

int hb_mixed_batch(const hb_batch_part_t* parts, size_t nparts) {
  if (!parts && nparts) return fail(HB_ERR_ARG, "null parts");
  Engine& E = engine();
  uint64_t words = 0;
  for (size_t i = 0; i < nparts; i++) {
    if (!parts[i].ek) return fail(HB_ERR_ARG, "part %zu: null ek", i);
    words += parts[i].count * (uint64_t)(parts[i].ek->p.n + 1) * 4;
  }
  std::vector<hb_batch_part_t> dev(parts, parts + nparts);
  // one host-buffer mixed call at a time: the staging buffer holds this call's inputs
  // until its kernel has read them and its outputs until they are copied back
  static std::mutex host_mu;
  std::lock_guard<std::mutex> host_lock(host_mu);
  {
    std::lock_guard<std::mutex> g(E.mu);
    if (E.mixed_io_words < words) {
      if (E.mixed_ev_used) HIP_TRY(hipEventSynchronize(E.mixed_ev));
      if (E.mixed_io) HIP_TRY(hipFree(E.mixed_io));
      E.mixed_io = nullptr;
      E.mixed_io_words = 0;
      HIP_TRY(hipMalloc((void**)&E.mixed_io, words * 8));
      E.mixed_io_words = words;
    }
    uint64_t off = 0;
    for (size_t i = 0; i < nparts; i++) {
      const hb_batch_part_t& q = parts[i];
      const size_t bytes = q.count * (size_t)(q.ek->p.n + 1) * 8;
      if (q.count == 0) continue;
      if (!q.c0s || !q.c1s || !q.outs) return fail(HB_ERR_ARG, "part %zu: null buffer", i);
      uint64_t* base = E.mixed_io + off;
      const size_t w = bytes / 8;
      HIP_TRY(hipMemcpyAsync(base, q.c0s, bytes, hipMemcpyHostToDevice, E.stream));
      HIP_TRY(hipMemcpyAsync(base + w, q.c1s, bytes, hipMemcpyHostToDevice, E.stream));
      if (q.c2s) HIP_TRY(hipMemcpyAsync(base + 2 * w, q.c2s, bytes, hipMemcpyHostToDevice, E.stream));
      dev[i].c0s = base;
      dev[i].c1s = base + w;
      dev[i].c2s = q.c2s ? base + 2 * w : nullptr;
      dev[i].outs = base + 3 * w;
      off += 4 * w;
    }
  }
  int rc = hb_mixed_batch_dev(dev.data(), nparts, nullptr);
  if (rc) return rc;
  for (size_t i = 0; i < nparts; i++) {
    if (parts[i].count == 0) continue;
    const size_t bytes = parts[i].count * (size_t)(parts[i].ek->p.n + 1) * 8;
    HIP_TRY(hipMemcpyAsync(parts[i].outs, dev[i].outs, bytes, hipMemcpyDeviceToHost, E.stream));
  }
  HIP_TRY(hipStreamSynchronize(E.stream));
  return record_ms(E.stream);
}

namespace {

// Traced launch for dense inputs; gate GATE_BOOTSTRAP takes c0s as the ciphertexts and
// `lut` (N canonical host words) as the shared test polynomial.
int traced_batch(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s, const uint64_t* c1s,
                 const uint64_t* c2s, const uint64_t* lut, uint64_t* outs, uint64_t* core_trace) {
  Engine& E = engine();
  std::lock_guard<std::mutex> g(ek->mu);
  int rc = ensure_io(ek, count);
  if (rc) return rc;
  rc = ensure_acc(ek, count);  // digit scratch (the accumulators live in the trace)
  if (rc) return rc;
  const ParamSet& p = ek->p;
  if (lut) {
    for (uint32_t k = 0; k < p.N; k++)
      if (lut[k] >= p.q) return fail(HB_ERR_ARG, "LUT word %u not canonical", k);
    HIP_TRY(hipMemcpyAsync(ek->lut_dev, lut, p.N * 8, hipMemcpyHostToDevice, E.stream));
  }
  const size_t n1 = p.n + 1, bytes = count * n1 * 8;
  const size_t wb = p.is_goldilocks ? 8 : 4;
  const size_t words = count * (size_t)(p.n + 1) * 2 * p.N;
  void* tr = nullptr;
  HIP_TRY(hipMalloc(&tr, words * wb));
  uint64_t* d0 = ek->io_dev;
  uint64_t* d1 = d0 + count * n1;
  uint64_t* d2 = d1 + count * n1;
  uint64_t* dout = d2 + count * n1;
  auto cleanup = [&](int code) {
    (void)hipFree(tr);
    return code;
  };
  if (hipMemcpyAsync(d0, c0s, bytes, hipMemcpyHostToDevice, E.stream) != hipSuccess ||
      (c1s && hipMemcpyAsync(d1, c1s, bytes, hipMemcpyHostToDevice, E.stream) != hipSuccess) ||
      (c2s && hipMemcpyAsync(d2, c2s, bytes, hipMemcpyHostToDevice, E.stream) != hipSuccess))
    return cleanup(fail(HB_ERR_DEVICE, "input copy failed"));
  BootArgs a = make_args(ek, gate, count, d0, c1s ? d1 : nullptr, c2s ? d2 : nullptr,
                         lut ? (const uint64_t*)ek->lut_dev : nullptr, dout, ek->acc_dev);
  a.trace = tr;
  if (acc_acquire(ek, E.stream) || timer_start(E.stream))
    return cleanup(HB_ERR_DEVICE);
  // latency mode (8 waves, the accumulator history written from LDS) up to the set's
  // cutoff, else one wavefront per gate with the history as its accumulator
  const bool split = count <= E.split_max[p.set] && split8_fits(p.br_len, p.br_bits);
  const hipError_t le = split ? launch_bootstrap_split(a, p.is_goldilocks, 8, E.stream)
                              : launch_bootstrap(a, p.is_goldilocks, 1, 1, E.stream);
  if (le != hipSuccess) return cleanup(fail(HB_ERR_DEVICE, "traced launch failed"));
  if (timer_stop(E.stream) || acc_release(ek, E.stream)) return cleanup(HB_ERR_DEVICE);
  std::vector<uint8_t> host(words * wb);
  if (hipMemcpyAsync(outs, dout, bytes, hipMemcpyDeviceToHost, E.stream) != hipSuccess ||
      hipMemcpyAsync(host.data(), tr, words * wb, hipMemcpyDeviceToHost, E.stream) != hipSuccess ||
      hipStreamSynchronize(E.stream) != hipSuccess)
    return cleanup(fail(HB_ERR_DEVICE, "trace copy failed"));
  if (p.is_goldilocks) {
    std::memcpy(core_trace, host.data(), words * 8);
  } else {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(host.data());
    for (size_t k = 0; k < words; k++) core_trace[k] = w[k];
  }
  return cleanup(record_ms(E.stream));
}

}  // namespace

int hb_gate_batch_traced(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s,
                         const uint64_t* c1s, const uint64_t* c2s, uint64_t* outs,
                         uint64_t* core_trace) {
  if (!ek || !c0s || !c1s || !outs || !core_trace) return fail(HB_ERR_ARG, "null argument");
  if (gate < 0 || gate > GATE_MAJORITY) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (gate == GATE_MAJORITY && !c2s) return fail(HB_ERR_ARG, "majority needs three inputs");
  if (count == 0) return HB_OK;
  return traced_batch(ek, gate, count, c0s, c1s, c2s, nullptr, outs, core_trace);
}

int hb_bootstrap_traced(hb_ek_t* ek, size_t count, const uint64_t* cs, const uint64_t* lut,
                        uint64_t* outs, uint64_t* core_trace) {
  if (!ek || !cs || !lut || !outs || !core_trace) return fail(HB_ERR_ARG, "null argument");
  if (count == 0) return HB_OK;
  return traced_batch(ek, GATE_BOOTSTRAP, count, cs, nullptr, nullptr, lut, outs, core_trace);
}

int hb_nand(hb_ek_t* ek, const uint64_t* c0, const uint64_t* c1, uint64_t* out) {
  return hb_gate_batch(ek, GATE_NAND, 1, c0, c1, nullptr, out);
}

int hb_gate(hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1, const uint64_t* c2,
            uint64_t* out) {
  return hb_gate_batch(ek, gate, 1, c0, c1, c2, out);
}

int hb_bootstrap(hb_ek_t* ek, const uint64_t* c, const uint64_t* lut, uint64_t* out) {
  if (!ek || !c || !lut || !out) return fail(HB_ERR_ARG, "null argument");
  Engine& E = engine();
  std::lock_guard<std::mutex> g(ek->mu);
  int rc = ensure_io(ek, 1);
  if (rc) return rc;
  const size_t n1 = ek->p.n + 1;
  for (uint32_t k = 0; k < ek->p.N; k++)
    if (lut[k] >= ek->p.q) return fail(HB_ERR_ARG, "LUT word %u not canonical", k);
  uint64_t* d0 = ek->io_dev;
  uint64_t* dout = d0 + 3 * n1;
  HIP_TRY(hipMemcpyAsync(d0, c, n1 * 8, hipMemcpyHostToDevice, E.stream));
  HIP_TRY(hipMemcpyAsync(ek->lut_dev, lut, ek->p.N * 8, hipMemcpyHostToDevice, E.stream));
  rc = launch_gates(ek, GATE_BOOTSTRAP, 1, d0, nullptr, nullptr, (const uint64_t*)ek->lut_dev,
                    dout, E.stream);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(out, dout, n1 * 8, hipMemcpyDeviceToHost, E.stream));
  HIP_TRY(hipStreamSynchronize(E.stream));
  rc = pair_check(ek, E.stream);
  if (rc) return rc;
  return record_ms(E.stream);
}

// Evaluator::not (tfhe/src/bfhe/evaluate.rs:98-104): -c, then b += Q/4.
int hb_not(const hb_ek_t* ek, const uint64_t* c, uint64_t* out) {
  if (!ek || !c || !out) return fail(HB_ERR_ARG, "null argument");
  const ParamSet& p = ek->p;
  HostNtt ops;
  ops.p = p;
  for (uint32_t k = 0; k <= p.n; k++) out[k] = c[k] == 0 ? 0 : p.q - c[k];
  out[p.n] = ops.add(out[p.n], p.q >> 2);
  return HB_OK;
}

int hb_ntt_dev(int field, int inverse, uint32_t log_n, uint64_t* data, size_t count, void* stream) {
  if (!data || (field != 0 && field != 1) || (log_n != 10 && log_n != 11))
    return fail(HB_ERR_ARG, "bad ntt arguments");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  hipStream_t s = stream ? (hipStream_t)stream : E.stream;
  rc = timer_start(s);
  if (rc) return rc;
  HIP_TRY(launch_ntt(field, inverse, log_n, data, count, E.tables.tw[field][log_n - 10][inverse ? 1 : 0],
                     E.tables.ninv[field][log_n - 10], E.simd_slots / 4, s));
  return timer_stop(s);
}

int hb_external_product_dev(int field, uint32_t log_n, uint32_t bits, uint32_t len,
                            const uint64_t* rgsw_dev, uint64_t* rlwe_dev, size_t count,
                            void* stream) {
  if (!rgsw_dev || !rlwe_dev || (field != 0 && field != 1) || log_n != 10 || bits == 0 ||
      bits > 16 || len == 0 || (uint64_t)bits * len < (field ? 64u : 31u))
    return fail(HB_ERR_ARG, "bad external product arguments");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  hipStream_t s = stream ? (hipStream_t)stream : E.stream;
  // The kernel reads the RGSW in device word form with 1/N folded in: convert here
  // into a cached buffer (one RGSW).
  ParamSet p;
  param_set(field ? SET_GOLDILOCKS : SET_BABYBEAR, &p);
  const size_t words = (size_t)2 * len * 2 * 1024;
  std::vector<uint64_t> host(words);
  HIP_TRY(hipMemcpy(host.data(), rgsw_dev, words * 8, hipMemcpyDeviceToHost));
  HostNtt ntt;
  ntt.init(p, 10);
  // One converted RGSW shared by every call: the previous call's kernel must be done with
  // it before it is rewritten.
  static std::mutex conv_mu;
  static void* conv = nullptr;
  static size_t conv_bytes = 0;
  static hipEvent_t conv_ev = nullptr;
  std::lock_guard<std::mutex> conv_lock(conv_mu);
  if (!conv_ev) HIP_TRY(hipEventCreateWithFlags(&conv_ev, hipEventDisableTiming));
  HIP_TRY(hipEventSynchronize(conv_ev));
  const size_t wb = field ? 8 : 4;
  if (conv_bytes < words * wb) {
    if (conv) (void)hipFree(conv);
    HIP_TRY(hipMalloc(&conv, words * wb));
    conv_bytes = words * wb;
  }
  if (field) {
    for (size_t k = 0; k < words; k++) host[k] = to_device_word(p, host[k], ntt.n_inv);
    HIP_TRY(hipMemcpy(conv, host.data(), words * 8, hipMemcpyHostToDevice));
  } else {
    std::vector<uint32_t> w(words);
    for (size_t k = 0; k < words; k++) w[k] = (uint32_t)to_device_word(p, host[k], ntt.n_inv);
    HIP_TRY(hipMemcpy(conv, w.data(), words * 4, hipMemcpyHostToDevice));
  }
  rc = timer_start(s);
  if (rc) return rc;
  HIP_TRY(launch_extprod(field, rlwe_dev, count, conv, len, bits,
                         field ? E.tables.twx : E.tables.tw[field][0][0],
                         E.tables.tw[field][0][1], s));
  rc = timer_stop(s);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(conv_ev, s));
  return HB_OK;
}

int hb_dev_alloc(size_t bytes, void** out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  HIP_TRY(hipMalloc(out, bytes));
  return HB_OK;
}
int hb_dev_free(void* p) {
  if (p) HIP_TRY(hipFree(p));
  return HB_OK;
}
int hb_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return HB_OK;
}
int hb_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return HB_OK;
}
int hb_sync(void) {
  Engine& E = engine();
  if (E.stream) HIP_TRY(hipStreamSynchronize(E.stream));
  HIP_TRY(hipDeviceSynchronize());
  return HB_OK;
}
int hb_stream_create(void** out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  hipStream_t s = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *out = s;
  return HB_OK;
}
int hb_stream_destroy(void* stream) {
  if (stream) HIP_TRY(hipStreamDestroy((hipStream_t)stream));
  return HB_OK;
}
int hb_stream_sync(void* stream) {
  if (!stream) return fail(HB_ERR_ARG, "null stream");
  HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
  return HB_OK;
}
int hb_device_count(int* out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *out = n;
  return HB_OK;
}
int hb_last_kernel_ms(double* out) {
  if (!out) return fail(HB_ERR_ARG, "null out");
  LaunchTimer& t = launch_timer();
  if (!t.used) return fail(HB_ERR_STATE, "no kernel launched yet on this thread");
  HIP_TRY(hipEventSynchronize(t.ev1));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, t.ev0, t.ev1));
  *out = ms;
  return HB_OK;
}
const char* hb_last_error(void) { return g_err.c_str(); }

}  // extern "C"

// ----------------------------------------------------------------------------
// Captured batch launches and the autotuner.
// ----------------------------------------------------------------------------
struct hb_batch_graph {
  hb_ek* ek = nullptr;
  void* acc = nullptr;  // private per-gate scratch (count * kScratchWordsPerGate words)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  size_t count = 0;
};

int hb_batch_graph_create(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s,
                          const uint64_t* c1s, const uint64_t* c2s, uint64_t* outs,
                          hb_batch_graph_t** out) {
  if (!ek || !c0s || !c1s || !outs || !out) return fail(HB_ERR_ARG, "null argument");
  if (gate < 0 || gate > GATE_MAJORITY) return fail(HB_ERR_ARG, "bad gate %d", gate);
  if (gate == GATE_MAJORITY && !c2s) return fail(HB_ERR_ARG, "majority needs three inputs");
  if (count == 0 || count > 0x7FFFFFFF) return fail(HB_ERR_ARG, "bad count");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  std::unique_ptr<hb_batch_graph> g(new hb_batch_graph);
  g->ek = ek;
  g->count = count;
  const size_t wb = ek->p.is_goldilocks ? 8 : 4;
  HIP_TRY(hipMalloc(&g->acc, count * kScratchWordsPerGate * wb));
  BootArgs a = make_args(ek, gate, count, c0s, c1s, c2s, nullptr, outs, g->acc);
  hipStream_t cs;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) {
    (void)hipFree(g->acc);
    return fail(HB_ERR_DEVICE, "capture stream: %s", hipGetErrorString(e));
  }
  e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t le = launch_shaped(E, ek->p, a, cs);
    e = hipStreamEndCapture(cs, &g->graph);
    if (le != hipSuccess) e = le;
  }
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess) {
    if (g->graph) (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->acc);
    return fail(HB_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
  }
  e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->acc);
    return fail(HB_ERR_DEVICE, "graph instantiate failed: %s", hipGetErrorString(e));
  }
  *out = g.release();
  return HB_OK;
}

int hb_batch_graph_launch(hb_batch_graph_t* g, void* stream) {
  if (!g) return fail(HB_ERR_ARG, "null graph");
  Engine& E = engine();
  hipStream_t s = stream ? (hipStream_t)stream : E.stream;
  int rc = timer_start(s);
  if (rc) return rc;
  HIP_TRY(hipGraphLaunch(g->exec, s));
  return timer_stop(s);
}

int hb_batch_graph_destroy(hb_batch_graph_t* g) {
  if (!g) return HB_OK;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  if (g->acc) (void)hipFree(g->acc);
  delete g;
  return HB_OK;
}

int hb_autotune(hb_ek_t* ek, size_t probe_gates, hb_tune_result_t* out) {
  if (!ek || !out) return fail(HB_ERR_ARG, "null argument");
  if (probe_gates == 0 || probe_gates > 65536) return fail(HB_ERR_ARG, "probe_gates out of range");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  std::lock_guard<std::mutex> lk(ek->mu);
  rc = ensure_io(ek, probe_gates);
  if (rc) return rc;
  rc = ensure_acc(ek, probe_gates);
  if (rc) return rc;
  const size_t n1 = ek->p.n + 1;
  // Uniform canonical words as inputs: every blind-rotation round is active, as for
  // real ciphertexts (a zero input would skip all rounds).
  {
    std::vector<uint64_t> h(probe_gates * 2 * n1);
    uint64_t x = 0x9E3779B97F4A7C15ull;
    for (auto& w : h) {
      x ^= x << 13, x ^= x >> 7, x ^= x << 17;
      w = x % ek->p.q;
    }
    HIP_TRY(hipMemcpy(ek->io_dev, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  }
  uint64_t* c0 = ek->io_dev;
  uint64_t* c1 = c0 + probe_gates * n1;
  uint64_t* o = c1 + probe_gates * n1;
  BootArgs a = make_args(ek, GATE_NAND, probe_gates, c0, c1, nullptr, nullptr, o, ek->acc_dev);
  double best = 1e300;
  int best_v = 0;
  rc = acc_acquire(ek, E.stream);
  if (rc) return rc;
  LaunchTimer& t = launch_timer();
  for (int v = 0; v < HB_TUNE_VARIANTS; v++) {
    const int gpw = kVariants[v][0], wps = kVariants[v][1];
    HIP_TRY(launch_bootstrap(a, ek->p.is_goldilocks, gpw, wps, E.stream));  // warm-up
    rc = timer_start(E.stream);
    if (rc) return rc;
    HIP_TRY(launch_bootstrap(a, ek->p.is_goldilocks, gpw, wps, E.stream));
    rc = timer_stop(E.stream);
    if (rc) return rc;
    HIP_TRY(hipEventSynchronize(t.ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, t.ev0, t.ev1));
    out->ms[v] = ms;
    if (ms < best) {
      best = ms;
      best_v = v;
    }
  }
  rc = acc_release(ek, E.stream);
  if (rc) return rc;
  E.gpw[ek->p.set] = kVariants[best_v][0];
  E.wps[ek->p.set] = kVariants[best_v][1];
  E.tuned[ek->p.set] = true;
  save_tune_cache(E);
  out->gates_per_workgroup = kVariants[best_v][0];
  out->waves_per_simd = kVariants[best_v][1];
  out->probe_gates = probe_gates;
  return HB_OK;
}

int hb_tuned_config(int set, int* gates_per_workgroup, int* waves_per_simd, int* tuned) {
  if (set < 0 || set > 2 || !gates_per_workgroup || !waves_per_simd)
    return fail(HB_ERR_ARG, "bad argument");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  *gates_per_workgroup = engine().gpw[set];
  *waves_per_simd = engine().wps[set];
  if (tuned) *tuned = engine().tuned[set] ? 1 : 0;
  return HB_OK;
}

int hb_set_tuned_config(int set, int gates_per_workgroup, int waves_per_simd) {
  if (set < 0 || set > 2) return fail(HB_ERR_ARG, "bad parameter set %d", set);
  if (!valid_variant(gates_per_workgroup, waves_per_simd))
    return fail(HB_ERR_ARG, "no kernel variant with %d gates per workgroup, %d waves per SIMD",
                gates_per_workgroup, waves_per_simd);
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  engine().gpw[set] = gates_per_workgroup;
  engine().wps[set] = waves_per_simd;
  engine().tuned[set] = true;
  return HB_OK;
}
int hb_modmul_peak(uint32_t waves, uint32_t iters, double* modmul_per_s, double* ms_out) {
  if (!modmul_per_s || waves == 0 || iters == 0 || waves > (1u << 20))
    return fail(HB_ERR_ARG, "bad modmul peak arguments");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  uint64_t* out = nullptr;
  HIP_TRY(hipMalloc(&out, (size_t)waves * 64 * 8));
  hipError_t e = launch_modmul_peak(out, waves, iters, 7, E.stream);  // warm-up
  if (e == hipSuccess && timer_start(E.stream) != HB_OK) e = hipErrorUnknown;
  if (e == hipSuccess) e = launch_modmul_peak(out, waves, iters, 11, E.stream);
  if (e == hipSuccess && timer_stop(E.stream) != HB_OK) e = hipErrorUnknown;
  LaunchTimer& t = launch_timer();
  if (e == hipSuccess) e = hipEventSynchronize(t.ev1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, t.ev0, t.ev1);
  (void)hipFree(out);
  if (e != hipSuccess) return fail(HB_ERR_DEVICE, "modmul peak: %s", hipGetErrorString(e));
  // per lane and iteration: 4 gl_mac2 = 8 products
  *modmul_per_s = (double)waves * 64.0 * iters * 8.0 / (ms * 1e-3);
  if (ms_out) *ms_out = ms;
  return HB_OK;
}

int hb_set_latency_mode_for(int set, uint32_t max_gates) {
  if (set < 0 || set >= SET_COUNT) return fail(HB_ERR_ARG, "bad parameter set %d", set);
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  Engine& E = engine();
  E.split_max[set] = max_gates == HB_LATENCY_DEFAULT ? default_split_max(set, E.simd_slots)
                                                     : max_gates;
  return HB_OK;
}

int hb_set_pair_mode(int enable) {
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  engine().pair_mode = enable != 0;
  return HB_OK;
}

int hb_set_pair_wait_limit(uint32_t microseconds) {
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  engine().pair_wait_us = microseconds == HB_PAIR_WAIT_DEFAULT ? kPairWaitDefaultUs : microseconds;
  return HB_OK;
}

int hb_ek_pair_check(hb_ek_t* ek) {
  if (!ek) return fail(HB_ERR_ARG, "null ek");
  std::lock_guard<std::mutex> g(ek->mu);
  return pair_check(ek, engine().stream);
}

int hb_pair_mode(int* enabled) {
  if (!enabled) return fail(HB_ERR_ARG, "null argument");
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  *enabled = engine().pair_mode ? 1 : 0;
  return HB_OK;
}

int hb_set_latency_mode(uint32_t max_gates) {
  for (int set = 0; set < SET_COUNT; set++) {
    int rc = hb_set_latency_mode_for(set, max_gates);
    if (rc) return rc;
  }
  return HB_OK;
}

int hb_latency_mode_for(int set, uint32_t* max_gates) {
  if (!max_gates) return fail(HB_ERR_ARG, "null argument");
  if (set < 0 || set >= SET_COUNT) return fail(HB_ERR_ARG, "bad parameter set %d", set);
  int rc = ensure_engine(engine().device < 0 ? 0 : engine().device);
  if (rc) return rc;
  *max_gates = engine().split_max[set];
  return HB_OK;
}

int hb_latency_mode(uint32_t* max_gates) { return hb_latency_mode_for(SET_GOLDILOCKS, max_gates); }
