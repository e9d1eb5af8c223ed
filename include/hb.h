/*
 * hb.h — C ABI of the MI355X TFHE gate-bootstrapping engine (libhasteboots.so).
 *
 * This is the drop-in boundary for the reference `tfhe` crate's gate API
 * (tfhe/src/lib.rs:1-9). Each entry point names the reference interface it
 * replaces. Conventions:
 *   - Every function returns an hb_status (0 = HB_OK); the reference panics or
 *     returns Err where these return a negative code.
 *   - Words are little-endian uint64_t canonical field values (in [0, Q)), for all
 *     parameter sets. An LWE ciphertext is n+1 words: a[0..n) then b.
 *   - Handles are opaque and library-owned. hb_ek_t may be shared across threads
 *     and streams: its keys are read-only after creation, and the per-key scratch of
 *     the batched kernels is handed from one launch to the next by a device event, so
 *     calls on different streams are ordered on the device where they share it (they
 *     do not overlap) and never corrupt each other; the same holds for the engine-wide
 *     scratch of mixed-parameter launches. hb_sk_t carries an encryption PRG and is
 *     not thread-safe (the reference keeps its PRG in a RefCell,
 *     fhe_core/src/secret_key.rs:55).
 *   - Entry points with a `_dev` suffix take device pointers on the engine's GPU
 *     and an optional stream (NULL = the engine's own stream); all others take
 *     host pointers and return after the result is on the host.
 */
#ifndef HB_H_
#define HB_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  HB_OK = 0,
  HB_ERR_ARG = -1,       /* invalid argument (null pointer, bad set id, bad size) */
  HB_ERR_DEVICE = -2,    /* no usable GPU, or a HIP runtime call failed */
  HB_ERR_NOMEM = -3,     /* host or device allocation failed */
  HB_ERR_CAPACITY = -4,  /* request exceeds a configured limit (serving admission) */
  HB_ERR_STATE = -5      /* call not valid in the handle's current state */
} hb_status;

/* Parameter set identifiers: tfhe/src/bfhe/parameters.rs:45-99. */
typedef enum {
  HB_SET_BABYBEAR = 0,   /* BABYBEAR_BINARY_128_BITS_PARAMETERS */
  HB_SET_GOLDILOCKS = 1, /* GOLDILOCKS_BINARY_128_BITS_PARAMETERS */
  HB_SET_ZAMA = 2        /* ZAMA_BINARY_128_BITS_PARAMETERS (= vfhe ZAMA_GOLDILOCKS_PARAMETERS) */
} hb_set_id;

/* Gate kinds: tfhe::Evaluator::{nand, and, or, nor, xor, xnor, majority, mux}
 * (tfhe/src/bfhe/evaluate.rs:113-263). Value 7 is reserved (hb_trace_expand uses it for
 * hb_bootstrap_traced gates). Mux takes three inputs (select, then, else) and costs three
 * bootstraps: two AND gates in one launch, their sum, then an OR-LUT bootstrap
 * (evaluate.rs:245-263). */
typedef enum {
  HB_GATE_NAND = 0,
  HB_GATE_AND = 1,
  HB_GATE_OR = 2,
  HB_GATE_NOR = 3,
  HB_GATE_XOR = 4,
  HB_GATE_XNOR = 5,
  HB_GATE_MAJORITY = 6,
  HB_GATE_MUX = 8
} hb_gate_kind;

typedef struct {
  int32_t set;
  int32_t is_goldilocks;
  uint32_t n;       /* lwe_dimension */
  uint32_t N;       /* ring_dimension */
  uint32_t t;       /* lwe_plain_modulus */
  uint32_t br_bits; /* blind_rotation_basis_bits */
  uint32_t br_len;  /* decompose_len of the blind-rotation basis */
  uint32_t ks_bits; /* key_switching_basis_bits */
  uint32_t ks_len;  /* decompose_len of the key-switching basis */
  uint32_t pad_;
  uint64_t q;       /* ring_modulus = lwe cipher modulus */
  double lwe_sigma, ring_sigma, ks_sigma;
} hb_params_t;

typedef struct hb_sk hb_sk_t;
typedef struct hb_ek hb_ek_t;

/* Parameters::new via the static constants (fhe_core/src/parameter.rs:111-172). */
int hb_params(int set, hb_params_t* out);

/* Sizes in words of the canonical key layouts used by hb_ek_export / hb_ek_import:
 * BSK [i<n][c_neg_s_m, c_m][j<br_len][a, b][N], KSK [j<ks_len][a, b][N], both in the
 * NTT domain (slot i = evaluation at psi^(2*bitrev(i)+1)). */
int hb_key_words(int set, uint64_t* bsk_words, uint64_t* ksk_words);

/* KeyGen::generate_secret_key (tfhe/src/key_gen.rs:12) → SecretKeyPack::new
 * (fhe_core/src/secret_key.rs:67-89), deterministic from `seed` (the reference seeds
 * from OS entropy). */
int hb_sk_create(int set, uint64_t seed, hb_sk_t** out);
int hb_sk_destroy(hb_sk_t* sk);
/* lwe_secret_key (n words) and ring_secret_key (N words), canonical. */
int hb_sk_export(const hb_sk_t* sk, uint64_t* lwe_sk, uint64_t* ring_sk);

/* Encryptor::encrypt (tfhe/src/encrypt.rs:20) → SecretKeyPack::encrypt
 * (fhe_core/src/secret_key.rs:129-150). Draws from the key's encryption stream. */
int hb_encrypt(hb_sk_t* sk, uint32_t bit, uint64_t* out_lwe);
/* Decryptor::decrypt / decrypt_with_noise (tfhe/src/decrypt.rs:20,26). `noise` may be NULL. */
int hb_decrypt(const hb_sk_t* sk, const uint64_t* lwe, uint32_t* bit, uint64_t* noise);

/* Evaluator::new (tfhe/src/bfhe/evaluate.rs:72-76) → EvaluationKey::new (:26-39):
 * generates BSK + KSK on the GPU (hb_keygen_dev) and keeps them resident in HBM of
 * `device`, with a canonical host copy for hb_ek_export. */
int hb_ek_create(const hb_sk_t* sk, int device, hb_ek_t** out);
/* Same, from keys in the canonical layout (e.g. produced elsewhere). */
int hb_ek_import(int set, const uint64_t* bsk, const uint64_t* ksk, int device, hb_ek_t** out);
int hb_ek_export(const hb_ek_t* ek, uint64_t* bsk, uint64_t* ksk);
/* Host-only key generation into caller buffers (canonical layout), no device needed. */
int hb_keygen_host(const hb_sk_t* sk, uint64_t* bsk, uint64_t* ksk);
/* The same keys generated on the GPU (BinaryBlindRotationKey::generate,
 * fhe_core/src/blind_rotation/rlwe/binary.rs:207-238; KeySwitchingRLWEKey::generate,
 * fhe_core/src/key_switch/rlwe.rs:37-112) into device arrays of hb_key_words words,
 * canonical layout, word for word equal to hb_keygen_host's; returns when they are written.
 * hb_ek_create generates its keys this way. */
int hb_keygen_dev(const hb_sk_t* sk, uint64_t* bsk_dev, uint64_t* ksk_dev, void* stream);

/* Versioned serialized forms (no reference counterpart): a 72-byte header (magic "HBTF",
 * format version 1, kind, the parameter set and its parameters, count, payload length, a
 * checksum) then canonical words; the layout is in hb_format.cpp. With buf == NULL the
 * *_serialize calls only report the size. Reading refuses other versions, parameters that
 * differ from the set's, a wrong length or checksum, and non-canonical words (HB_ERR_ARG). */
int hb_keys_serialize(int set, const uint64_t* bsk, const uint64_t* ksk, void* buf, size_t cap,
                      size_t* size);
int hb_keys_deserialize(const void* buf, size_t size, int* set, uint64_t* bsk, uint64_t* ksk);
int hb_ek_serialize(const hb_ek_t* ek, void* buf, size_t cap, size_t* size);
/* hb_keys_deserialize + hb_ek_import. */
int hb_ek_deserialize(const void* buf, size_t size, int device, hb_ek_t** out);
/* `count` LWE ciphertexts of n+1 canonical words each. hb_lwe_deserialize with lwes == NULL
 * reports set and count only; cap_words is the room in lwes. */
int hb_lwe_serialize(int set, size_t count, const uint64_t* lwes, void* buf, size_t cap,
                     size_t* size);
int hb_lwe_deserialize(const void* buf, size_t size, int* set, size_t* count, uint64_t* lwes,
                       size_t cap_words);
int hb_ek_destroy(hb_ek_t* ek);
int hb_ek_params(const hb_ek_t* ek, hb_params_t* out);

/* Evaluator::nand (tfhe/src/bfhe/evaluate.rs:113-121): one gate, one launch. */
int hb_nand(hb_ek_t* ek, const uint64_t* c0, const uint64_t* c1, uint64_t* out);
/* Evaluator::{and, or, nor, xor, xnor, majority, mux} (evaluate.rs:129-263); c2 only for
 * majority and mux. */
int hb_gate(hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1, const uint64_t* c2,
            uint64_t* out);
/* Evaluator::bootstrap (evaluate.rs:86-88 → :42-60) with an arbitrary LUT of N words. */
int hb_bootstrap(hb_ek_t* ek, const uint64_t* c, const uint64_t* lut, uint64_t* out);
/* Evaluator::not (evaluate.rs:98-104): linear, no bootstrap. */
int hb_not(const hb_ek_t* ek, const uint64_t* c, uint64_t* out);

/* Batched gates: `count` independent gates of one kind in a single launch (no
 * reference counterpart; the reference runs gates one per thread, tfhe/examples/bfhe.rs:139-148).
 * Host arrays of count*(n+1) words; c2s only for majority and mux (mux: two launches,
 * 2*count AND gates then count OR-LUT bootstraps). */
int hb_gate_batch(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s, const uint64_t* c1s,
                  const uint64_t* c2s, uint64_t* outs);
/* Same on device-resident arrays; `stream` is a hipStream_t or NULL. Asynchronous. */
int hb_gate_batch_dev(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s_dev,
                      const uint64_t* c1s_dev, const uint64_t* c2s_dev, uint64_t* outs_dev,
                      void* stream);

/* No reference counterpart. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): hb_gate_batch_dev
 * for gate kinds 0-6 with the kernels' per-gate scratch (accumulator and byte digits, 4N
 * words of the set's word size per gate) taken from a fixed device array of
 * HB_NOALLOC_SCRATCH_BYTES that the library reserves on its GPU for the life of the
 * process; the call allocates nothing. At most 1024 Goldilocks / Zama or 2048 BabyBear gates
 * per call (HB_ERR_CAPACITY above). Every call on the GPU shares that one array, which keeps
 * the last call's contents: calls whose kernels can overlap (other streams, other threads
 * without ordering) race and corrupt each other's results, so the caller serialises them.
 * Not timed by hb_last_kernel_ms. */
#define HB_NOALLOC_SCRATCH_BYTES (32u << 20)
int hb_gate_batch_dev_noalloc(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s_dev,
                              const uint64_t* c1s_dev, const uint64_t* c2s_dev,
                              uint64_t* outs_dev, void* stream);

/* Captured batch launch: hb_gate_batch_dev for fixed device buffers recorded once as a
 * hipGraph (with its own accumulator scratch) and replayed by hb_batch_graph_launch
 * without per-call host work. Inputs may be rewritten between launches; replays on one
 * graph must not overlap. No reference counterpart (the reference has no device launch). */
typedef struct hb_batch_graph hb_batch_graph_t;
int hb_batch_graph_create(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s_dev,
                          const uint64_t* c1s_dev, const uint64_t* c2s_dev, uint64_t* outs_dev,
                          hb_batch_graph_t** out);
int hb_batch_graph_launch(hb_batch_graph_t* graph, void* stream);
int hb_batch_graph_destroy(hb_batch_graph_t* graph);

/* Autotuner for the batched kernel: times its variants on `probe_gates` gates with this
 * key's parameter set and keeps the fastest for that set (used by every later batch
 * launch of the set); when the environment variable HB_AUTOTUNE_CACHE names a file the
 * choice is stored there and a new process reads it back. Variants, in the order of
 * ms[]: (gates per workgroup, waves per SIMD) = (1,1), (2,1), (4,1), (2,2), (4,2), (4,3).
 * Gates per workgroup share one LDS copy of the twiddle tables; 2 (3) waves per SIMD cap
 * registers at 256 (168) per lane so the waves hide each other's latency. The tuned variant
 * applies to batches larger than one wave per SIMD of the device (4 x compute units);
 * smaller batches always run (1,1). Probe with at least twice that many gates. No
 * reference counterpart. */
#define HB_TUNE_VARIANTS 6
typedef struct {
  int gates_per_workgroup;
  int waves_per_simd;
  uint64_t probe_gates;
  double ms[HB_TUNE_VARIANTS]; /* kernel time per variant */
} hb_tune_result_t;
int hb_autotune(hb_ek_t* ek, size_t probe_gates, hb_tune_result_t* out);
/* Current variant for a parameter set; *tuned = 1 if it came from hb_autotune, the cache
 * or hb_set_tuned_config (otherwise it is the built-in default). */
int hb_tuned_config(int set, int* gates_per_workgroup, int* waves_per_simd, int* tuned);
/* Pins the variant for a parameter set. */
int hb_set_tuned_config(int set, int gates_per_workgroup, int waves_per_simd);
/* Latency mode: batches (hb_nand, hb_gate*, hb_bootstrap, hb_gate_batch*, captured
 * graphs) of at most max_gates gates run 8 wavefronts per gate (4 for BabyBear launches of
 * more than one gate per compute unit), each wave taking a share of the digit transforms
 * and of the inverse transforms, which cuts a gate's latency about 4x at lower peak
 * throughput. Each parameter set has its own cutoff. Defaults: seven gates per compute
 * unit (1792 on MI355X) for Goldilocks and BabyBear and six (1536) for Zama, the largest
 * batches at which latency mode still beats the batched kernels on MI355X.
 * hb_set_latency_mode sets the cutoff of every set, hb_set_latency_mode_for of
 * one; 0 disables latency mode, HB_LATENCY_DEFAULT restores the built-in cutoff.
 * hb_latency_mode reports the Goldilocks cutoff, hb_latency_mode_for any set's. */
#define HB_LATENCY_DEFAULT 0xFFFFFFFFu
/* Pair mode (no reference counterpart): batches of every set (untraced) within
 * the latency-mode cutoff and of at most one gate per two compute units (128 on MI355X) run
 * each gate on two workgroups of 8 wavefronts on two compute units, one accumulator
 * component each, exchanging half of the external product through global memory every
 * round. Workgroups pair up in the order they start, so a launch needs no compute units
 * reserved for it and runs beside kernels on other streams.
 * Calls through hb_gate_batch* and the single-gate calls take it, and so do the serving
 * loop's small launches while its other launch slot is idle; captured graphs, the
 * no-allocation call and hb_mixed_batch keep latency mode. On by default;
 * hb_set_pair_mode(0) turns it off. Same output words as latency mode.
 * The two workgroups of a gate wait for each other's flags for a bounded time
 * (HB_PAIR_WAIT_DEFAULT: 2 s; a round's wait takes microseconds, a workgroup waiting for its
 * partner to start waits for a compute unit to free up); a wait that runs out ends the
 * launch's waits and leaves that launch's outputs wrong, and the launch is then reported as
 * HB_ERR_DEVICE: by hb_gate_batch and the single-gate calls that ran it or that follow it
 * on the same key, by hb_ek_pair_check (for hb_gate_batch_dev launches, once the caller
 * has synchronised their streams), and by the serving loop, which stops with that error.
 * hb_set_pair_wait_limit sets the limit in microseconds (0: every wait times out at once,
 * which tests use to force the error). */
#define HB_PAIR_WAIT_DEFAULT 0xFFFFFFFFu
int hb_set_pair_mode(int enable);
int hb_pair_mode(int* enabled);
int hb_set_pair_wait_limit(uint32_t microseconds);
int hb_ek_pair_check(hb_ek_t* ek);
int hb_set_latency_mode(uint32_t max_gates);
int hb_set_latency_mode_for(int set, uint32_t max_gates);
int hb_latency_mode(uint32_t* max_gates);
int hb_latency_mode_for(int set, uint32_t* max_gates);

/* Mixed-parameter batch: several parts, each with its own evaluation key (any of the
 * three parameter sets) and gate kind, bootstrapped in ONE kernel launch. No reference
 * counterpart (the reference evaluates one gate per call). Heavier parameter sets are
 * scheduled first. All keys must live on the engine's device. */
typedef struct {
  hb_ek_t* ek;
  int32_t gate;
  uint32_t pad_;
  uint64_t count;
  const uint64_t* c0s;
  const uint64_t* c1s;
  const uint64_t* c2s; /* majority only */
  uint64_t* outs;
} hb_batch_part_t;
/* Host pointers; returns when outputs are on the host. */
int hb_mixed_batch(const hb_batch_part_t* parts, size_t nparts);
/* Device pointers; asynchronous on `stream` (NULL = the engine's stream). */
int hb_mixed_batch_dev(const hb_batch_part_t* parts, size_t nparts, void* stream);

/* ---- Traced bootstrapping: vfhe::Evaluator (vfhe/src/bfhe/evaluate.rs:47-139, 196-208),
 * whose bootstrap returns a PBSTrace for the prover (trace/src/pbs_trace.rs:8-13). ----
 * Same gates as hb_gate_batch with the traced path's semantics (monomial products through
 * the NTT, every round executed), plus a compact core trace per gate: the accumulator
 * RLWE entering each of the n rounds and the final one, (n+1)*2N words per gate (raw
 * Goldilocks words / BabyBear values), a then b per slot. */
int hb_gate_batch_traced(hb_ek_t* ek, int gate, size_t count, const uint64_t* c0s,
                         const uint64_t* c1s, const uint64_t* c2s, uint64_t* outs,
                         uint64_t* core_trace);
/* Traced EvaluationKey::bootstrap with one caller LUT (N canonical words) shared by the
 * `count` ciphertexts in cs. */
int hb_bootstrap_traced(hb_ek_t* ek, size_t count, const uint64_t* cs, const uint64_t* lut,
                        uint64_t* outs, uint64_t* core_trace);
/* Words of one expanded round record and of the tail record (layouts in hb_trace.cpp).
 * hb_trace_expand takes gate = 7 with c1 = c2 = NULL for hb_bootstrap_traced gates. */
int hb_trace_sizes(int set, uint64_t* round_words, uint64_t* tail_words);
/* Rebuilds the PBSTrace components of one gate from its core trace: the listed rounds
 * (increasing, < n) into rec, and the modulus-switch / key-switch / sample-extraction
 * values into tail (NULL skips). Checks the core trace against the keys on the way
 * (HB_ERR_STATE if a round's product does not match). */
int hb_trace_expand(const hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1,
                    const uint64_t* c2, const uint64_t* core_trace, const uint32_t* rounds,
                    uint32_t n_rounds, uint64_t* rec, uint64_t* tail);

/* Finalized PBSTrace containers: what the prover takes after
 * ModulusSwitchingTrace::finalize(n + 1) and BlindRotationTrace::finalize(n)
 * (vfhe/examples/zk_binary_pbs/zk_nand_goldilocks.rs:97-105). Rows are the n blind-rotation
 * rounds zero padded to R = 2^ceil(log2 n) (1024); words are canonical field values; each
 * plane is rows x N, row-major, planes in the struct's field order:
 *   RLWETrace [poly.0, poly.1, ntt.0, ntt.1]         (basic_ops/rlwe_trace.rs:55-60)
 *   PolynomialTrace [poly, ntt]                      (basic_ops/rlwe_trace.rs:10-15)
 *   MonomialTrace [degree, coefficient] (R each)     (basic_ops/rlwe_trace.rs:32-38)
 *   HadamardTrace [bit.poly, bit.ntt, rlwe.poly.0, rlwe.poly.1, rlwe.ntt.0, rlwe.ntt.1]
 *                                                    (basic_ops/hadamard_trace.rs:316-342)
 * AccTrace (trace/src/acc_trace.rs:15-50, finalize :157-202): */
#define HB_TP_INITIAL_ACC 0            /* RLWETrace, one row (4N) */
#define HB_TP_FINAL_ACC 1              /* RLWETrace, one row (4N) */
#define HB_TP_INPUT_ACC 2              /* RLWETrace (4RN) */
#define HB_TP_OUTPUT_ACC 3             /* RLWETrace (4RN) */
#define HB_TP_INPUT_ACC_PERMUTED 4     /* output_acc with its poly planes' last row
                                          += initial - final; ntt planes as output_acc's */
#define HB_TP_MONOMIAL 5               /* PolynomialTrace (2RN) */
#define HB_TP_MONOMIAL_REPR 6          /* MonomialTrace (2R) */
#define HB_TP_MONOMIAL_TIMES_ACC 7     /* RLWETrace (4RN) */
#define HB_TP_EXTERNAL_PRODUCT_INPUT 8 /* RLWETrace (4RN) */
/* SumHadamardTrace (basic_ops/hadamard_trace.rs:329-401): */
#define HB_TP_HADAMARD 9               /* HadamardTrace `index` = half * br_len + digit (6RN) */
#define HB_TP_SUM_PROD 10              /* RLWETrace (4RN) */
#define HB_TP_PERMUTATION 11           /* AccTrace::permutation_info.permutation_table (R):
                                          new_rotation_left(log R, n, 1), row_perm_trace.rs:41-52 */
#define HB_TP_MODULUS_SWITCH 12        /* ModulusSwitchingTrace (modulus_switching_trace.rs:
                                          25-49, 77-133): input, output, output_witness,
                                          helper (M = 2^ceil(log2(n+1)) each), then
                                          modulus_after, blk_param, helper_upper_bound */
/* Words of container `part` for parameter set `set`. */
int hb_trace_part_words(int set, int part, uint64_t* words);
/* Writes container `part` (hadamard `index`, else ignored) of one gate, built from its core
 * trace and inputs as in hb_trace_expand, into out (hb_trace_part_words words). The
 * per-round key consistency check is hb_trace_expand's; this call trusts the core trace. */
int hb_trace_part(const hb_ek_t* ek, int gate, const uint64_t* c0, const uint64_t* c1,
                  const uint64_t* c2, const uint64_t* core_trace, int part, uint32_t index,
                  uint64_t* out);

/* ---- Serving loop (no reference counterpart: the reference evaluates one gate per call).
 * A server admits boolean circuits from many clients (each with its own hb_ek_t) as they
 * arrive, keeps their wires in HBM between layers, admits by an HBM budget (a device
 * arena; requests that do not fit wait in FIFO order) and batches the ready gates of all
 * admitted circuits into shared kernel launches (continuous batching, two launches in
 * flight). ---- */
typedef struct hb_server hb_server_t;

/* Wires 0..n_inputs-1 are the inputs; gate g reads wires in0[g], in1[g] (each < n_inputs+g)
 * and writes wire n_inputs+g. kind[g] is an hb_gate_kind of two inputs (NULL = all NAND). */
typedef struct {
  uint32_t n_inputs;
  uint32_t n_gates;
  uint32_t n_outputs;
  uint32_t pad_;
  const uint32_t* in0;
  const uint32_t* in1;
  const uint8_t* kind;
  const uint32_t* outputs; /* wire indices returned by hb_server_wait */
} hb_circuit_t;

typedef struct {
  uint64_t submitted, admitted, completed;
  uint64_t gates_done, launches;
  uint64_t queued, active;
  uint64_t bytes_in_use, budget;
  double busy_ms; /* summed event-timed kernel duration */
  /* gate latency = from a gate's inputs being ready (admission or previous layer done) to
   * the completion of the launch that ran it; gate-weighted percentiles since creation */
  double gate_p50_ms, gate_p99_ms;
  /* launches run in the throughput shape; largest launch so far (gates) */
  uint64_t batched_launches, max_launch_gates;
  /* launches run in pair mode (each gate on two compute units) */
  uint64_t pair_launches;
} hb_server_stats_t;

/* max_gates_per_launch: gate slots per launch. A launch takes at most one gate per compute
 * unit (256 on MI355X) in latency mode; when max_gates_per_launch allows it and the ready
 * gates of the admitted circuits number at least two batched workgroups of 4 per compute
 * unit (2048 on MI355X), it takes up to max_gates_per_launch gates in the batched kernel's
 * (4 gates per workgroup, 2 waves per SIMD) shape. hbm_budget_bytes: arena size (0 = half
 * of the free device memory, at most 32 GiB). */
int hb_server_create(int device, uint32_t max_gates_per_launch, uint64_t hbm_budget_bytes,
                     hb_server_t** out);
int hb_server_destroy(hb_server_t* server);
/* inputs: n_inputs LWE ciphertexts under ek's key. HB_ERR_CAPACITY if the circuit can
 * never fit the budget. */
int hb_server_submit(hb_server_t* server, hb_ek_t* ek, const hb_circuit_t* circuit,
                     const uint64_t* inputs, uint64_t* request_id);
/* As hb_server_submit; traced != 0 runs the circuit with the traced (vfhe) semantics and
 * keeps each gate's core trace ((n+1)*2N words, see hb_gate_batch_traced) in HBM, counted
 * against the budget, until hb_server_wait_traced returns it. */
int hb_server_submit_traced(hb_server_t* server, hb_ek_t* ek, const hb_circuit_t* circuit,
                            const uint64_t* inputs, int traced, uint64_t* request_id);
int hb_server_poll(hb_server_t* server, uint64_t request_id, int* done);
/* Blocks until done; outputs: n_outputs ciphertexts; latency = submit → done (ms). */
int hb_server_wait(hb_server_t* server, uint64_t request_id, uint64_t* outputs,
                   double* latency_ms);
/* As hb_server_wait, plus core_traces: n_gates * (n+1) * 2N words in gate order (traced
 * requests only; NULL skips). */
int hb_server_wait_traced(hb_server_t* server, uint64_t request_id, uint64_t* outputs,
                          uint64_t* core_traces, double* latency_ms);
int hb_server_stats(hb_server_t* server, hb_server_stats_t* out);

/* Standalone negacyclic NTT over `count` polynomials of N = 2^log_n words (device
 * pointers, canonical in and out, in place). Forward output in the key layout's slot
 * order. Replaces AbstractNTT::{transform_slice, inverse_transform_slice}
 * (algebra/src/transformation/mod.rs:85,95). field: 0 BabyBear, 1 Goldilocks. */
int hb_ntt_dev(int field, int inverse, uint32_t log_n, uint64_t* data_dev, size_t count, void* stream);

/* External product of `count` RLWEs (2N words each, a then b, coefficient domain,
 * canonical) with one NTT-domain RGSW in the canonical layout [2][len][2][N]
 * (RLWE::mul_assign_ntt_rgsw, lattice/src/rlwe.rs:573-596). Device pointers. */
int hb_external_product_dev(int field, uint32_t log_n, uint32_t bits, uint32_t len,
                            const uint64_t* rgsw_dev, uint64_t* rlwe_dev, size_t count,
                            void* stream);

/* Peak Goldilocks multiply-accumulate rate of this build on the device: `waves`
 * wavefronts each run iters x 4 register-resident MAC pairs (the MAC stage's sequence:
 * Montgomery product + modular add). *modmul_per_s counts products. Measurement only;
 * no reference counterpart. */
int hb_modmul_peak(uint32_t waves, uint32_t iters, double* modmul_per_s, double* ms);

/* Device memory helpers so callers without a HIP runtime of their own can stage data. */
int hb_dev_alloc(size_t bytes, void** out);
int hb_dev_free(void* p);
int hb_memcpy_h2d(void* dst_dev, const void* src, size_t bytes);
int hb_memcpy_d2h(void* dst, const void* src_dev, size_t bytes);
int hb_sync(void);
int hb_device_count(int* out);
/* Streams on the engine's device (default priority, non-blocking), for callers without a
 * HIP runtime of their own. */
int hb_stream_create(void** out);
int hb_stream_destroy(void* stream);
int hb_stream_sync(void* stream);

/* Duration in ms of the calling thread's last timed launch (hb_*_dev, host-buffer gate
 * calls, hb_batch_graph_launch, hb_ntt_dev, hb_external_product_dev), measured with HIP
 * events recorded on the stream the kernel ran on; each thread has its own events. */
int hb_last_kernel_ms(double* out);

/* Human-readable message for the last error on the calling thread. */
const char* hb_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* HB_H_ */
