"""Host-side mirror of the reference `tfhe` crate's gate API over the C ABI.

Reference interface (tfhe/src/lib.rs:1-9):
  KeyGen::generate_secret_key(params) -> SecretKeyPack         tfhe/src/key_gen.rs:12
  Encryptor::new(sk).encrypt(m) -> LWECiphertext               tfhe/src/encrypt.rs:14,20
  Decryptor::new(sk).decrypt(c) / decrypt_with_noise(c)        tfhe/src/decrypt.rs:20,26
  Evaluator::new(&sk) ; nand/and/or/nor/xor/xnor/majority/mux/not/bootstrap
                                                               tfhe/src/bfhe/evaluate.rs:72-263
Parameter constants: tfhe/src/bfhe/parameters.rs:45-99.

Differences, by design: key generation takes a seed (the reference draws from OS
entropy), and every Evaluator method has a batched form that runs many independent
gates in one GPU launch. Errors raise HbError instead of panicking.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import LATENCY_DEFAULT, PAIR_WAIT_DEFAULT, TUNE_VARIANTS, HbParams, HbTuneResult, check, lib, ptr

SET_BABYBEAR, SET_GOLDILOCKS, SET_ZAMA = 0, 1, 2
GATES = {"nand": 0, "and": 1, "or": 2, "nor": 3, "xor": 4, "xnor": 5, "majority": 6,
         "mux": 8}


@dataclass(frozen=True)
class Parameters:
    """Parameters<Q> (fhe_core/src/parameter.rs:71-77) for one of the named sets."""
    set_id: int
    n: int
    N: int
    t: int
    q: int
    br_bits: int
    br_len: int
    ks_bits: int
    ks_len: int
    is_goldilocks: bool

    @staticmethod
    def of(set_id: int) -> "Parameters":
        p = HbParams()
        check(lib().hb_params(set_id, C.byref(p)))
        return Parameters(p.set, p.n, p.N, p.t, p.q, p.br_bits, p.br_len, p.ks_bits, p.ks_len,
                          bool(p.is_goldilocks))

    def lwe_dimension(self) -> int:
        return self.n

    def ring_dimension(self) -> int:
        return self.N

    def lwe_cipher_modulus_value(self) -> int:
        return self.q


def BABYBEAR_BINARY_128_BITS_PARAMETERS() -> Parameters:
    return Parameters.of(SET_BABYBEAR)


def GOLDILOCKS_BINARY_128_BITS_PARAMETERS() -> Parameters:
    return Parameters.of(SET_GOLDILOCKS)


def ZAMA_BINARY_128_BITS_PARAMETERS() -> Parameters:
    return Parameters.of(SET_ZAMA)


class LWECiphertext:
    """LWE<Q> { a: Vec<Q>, b: Q } (lattice/src/lwe.rs:20-23), stored as n+1 canonical words."""

    __slots__ = ("words",)

    def __init__(self, words: np.ndarray):
        self.words = np.ascontiguousarray(words, dtype=np.uint64)

    @property
    def a(self) -> np.ndarray:
        return self.words[:-1]

    @property
    def b(self) -> int:
        return int(self.words[-1])

    def __eq__(self, other) -> bool:
        return isinstance(other, LWECiphertext) and np.array_equal(self.words, other.words)


class SecretKeyPack:
    """SecretKeyPack<Q> (fhe_core/src/secret_key.rs:42-56). Not thread-safe (holds a PRG)."""

    def __init__(self, params: Parameters, seed: int = 0):
        self.params = params
        self.seed = seed
        h = C.c_void_p()
        check(lib().hb_sk_create(params.set_id, C.c_uint64(seed), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().hb_sk_destroy(h)
            self._h = None

    def lwe_secret_key(self) -> np.ndarray:
        out = np.zeros(self.params.n, dtype=np.uint64)
        check(lib().hb_sk_export(self._h, ptr(out), None))
        return out

    def ring_secret_key(self) -> np.ndarray:
        out = np.zeros(self.params.N, dtype=np.uint64)
        check(lib().hb_sk_export(self._h, None, ptr(out)))
        return out

    def parameters(self) -> Parameters:
        return self.params

    def generate_keys_host(self) -> tuple[np.ndarray, np.ndarray]:
        """BSK and KSK in the canonical layout, generated on the host (no GPU needed)."""
        bw, kw = C.c_uint64(), C.c_uint64()
        check(lib().hb_key_words(self.params.set_id, C.byref(bw), C.byref(kw)))
        bsk = np.zeros(bw.value, dtype=np.uint64)
        ksk = np.zeros(kw.value, dtype=np.uint64)
        check(lib().hb_keygen_host(self._h, ptr(bsk), ptr(ksk)))
        return bsk, ksk

    def generate_keys_device(self) -> tuple[np.ndarray, np.ndarray]:
        """The same keys generated on the GPU (hb_keygen_dev), copied to the host."""
        bw, kw = C.c_uint64(), C.c_uint64()
        check(lib().hb_key_words(self.params.set_id, C.byref(bw), C.byref(kw)))
        L = lib()
        bufs = [C.c_void_p(), C.c_void_p()]
        try:
            for b, w in zip(bufs, (bw.value, kw.value)):
                check(L.hb_dev_alloc(w * 8, C.byref(b)))
            check(L.hb_keygen_dev(self._h, bufs[0], bufs[1], None))
            out = []
            for b, w in zip(bufs, (bw.value, kw.value)):
                a = np.zeros(w, dtype=np.uint64)
                check(L.hb_memcpy_d2h(a.ctypes.data, b, a.nbytes))
                out.append(a)
        finally:
            for b in bufs:
                if b.value:
                    L.hb_dev_free(b)
        return out[0], out[1]


def _serialized(fn, *args) -> bytes:
    size = C.c_size_t()
    check(fn(*args, None, 0, C.byref(size)))
    buf = C.create_string_buffer(size.value)
    check(fn(*args, buf, size.value, C.byref(size)))
    return buf.raw[:size.value]


def serialize_keys(set_id: int, bsk, ksk) -> bytes:
    """Canonical BSK + KSK in the versioned format (hb_keys_serialize)."""
    bsk = np.ascontiguousarray(bsk, np.uint64)
    ksk = np.ascontiguousarray(ksk, np.uint64)
    return _serialized(lib().hb_keys_serialize, set_id, ptr(bsk), ptr(ksk))


def deserialize_keys(data: bytes) -> tuple[int, np.ndarray, np.ndarray]:
    """(set id, bsk, ksk) from hb_keys_serialize's format, checked."""
    set_id = C.c_int()
    check(lib().hb_keys_deserialize(data, len(data), C.byref(set_id), None, None))
    bw, kw = C.c_uint64(), C.c_uint64()
    check(lib().hb_key_words(set_id.value, C.byref(bw), C.byref(kw)))
    bsk = np.zeros(bw.value, dtype=np.uint64)
    ksk = np.zeros(kw.value, dtype=np.uint64)
    check(lib().hb_keys_deserialize(data, len(data), None, ptr(bsk), ptr(ksk)))
    return set_id.value, bsk, ksk


def serialize_ciphertexts(set_id: int, cts) -> bytes:
    """LWE ciphertexts (count, n+1) in the versioned format (hb_lwe_serialize)."""
    cts = np.ascontiguousarray(np.atleast_2d(cts), np.uint64)
    return _serialized(lib().hb_lwe_serialize, set_id, cts.shape[0] if cts.size else 0, ptr(cts))


def deserialize_ciphertexts(data: bytes) -> tuple[int, np.ndarray]:
    """(set id, ciphertexts of shape (count, n+1)) from hb_lwe_serialize's format."""
    set_id, count = C.c_int(), C.c_size_t()
    check(lib().hb_lwe_deserialize(data, len(data), C.byref(set_id), C.byref(count), None, 0))
    P = Parameters.of(set_id.value)
    out = np.zeros((count.value, P.n + 1), dtype=np.uint64)
    check(lib().hb_lwe_deserialize(data, len(data), None, None, ptr(out), out.size))
    return set_id.value, out


class KeyGen:
    @staticmethod
    def generate_secret_key(params: Parameters, seed: int = 0) -> SecretKeyPack:
        return SecretKeyPack(params, seed)


class Encryptor:
    def __init__(self, sk: SecretKeyPack):
        self.sk = sk

    def encrypt(self, m: int) -> LWECiphertext:
        out = np.zeros(self.sk.params.n + 1, dtype=np.uint64)
        check(lib().hb_encrypt(self.sk._h, int(m), ptr(out)))
        return LWECiphertext(out)

    def encrypt_many(self, bits) -> np.ndarray:
        bits = np.asarray(bits, dtype=np.uint32)
        out = np.zeros((len(bits), self.sk.params.n + 1), dtype=np.uint64)
        for i, m in enumerate(bits):
            check(lib().hb_encrypt(self.sk._h, int(m), ptr(out[i])))
        return out


class Decryptor:
    def __init__(self, sk: SecretKeyPack):
        self.sk = sk

    def decrypt(self, c: LWECiphertext) -> int:
        return self.decrypt_with_noise(c)[0]

    def decrypt_with_noise(self, c) -> tuple[int, int]:
        words = c.words if isinstance(c, LWECiphertext) else np.ascontiguousarray(c, np.uint64)
        bit = C.c_uint32()
        noise = C.c_uint64()
        check(lib().hb_decrypt(self.sk._h, ptr(words), C.byref(bit),
                               C.cast(C.byref(noise), C.POINTER(C.c_uint64))))
        return int(bit.value), int(noise.value)


# Finalized PBSTrace containers, in hb.h's HB_TP_* order.
TRACE_PARTS = ("initial_acc", "final_acc", "input_acc", "output_acc", "input_acc_permuted",
               "monomial", "monomial_representation", "monomial_times_acc",
               "external_product_input", "hadamard", "sum_prod", "permutation_info",
               "modulus_switching")


class Evaluator:
    """Evaluator<Q> (tfhe/src/bfhe/evaluate.rs:63-263) backed by keys resident in HBM."""

    def __init__(self, sk: SecretKeyPack | None = None, device: int = 0, *, keys=None,
                 params: Parameters | None = None):
        h = C.c_void_p()
        if sk is not None:
            self.params = sk.params
            check(lib().hb_ek_create(sk._h, device, C.byref(h)))
        else:
            bsk, ksk = keys
            self.params = params
            bsk = np.ascontiguousarray(bsk, np.uint64)
            ksk = np.ascontiguousarray(ksk, np.uint64)
            check(lib().hb_ek_import(params.set_id, ptr(bsk), ptr(ksk), device, C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().hb_ek_destroy(h)
            self._h = None

    def parameters(self) -> Parameters:
        return self.params

    def serialize(self) -> bytes:
        """The evaluation key in the versioned format (hb_ek_serialize)."""
        return _serialized(lib().hb_ek_serialize, self._h)

    @classmethod
    def deserialize(cls, data: bytes, device: int = 0) -> "Evaluator":
        """An evaluator over keys read from hb_ek_serialize's format (hb_ek_deserialize)."""
        ev = cls.__new__(cls)
        h = C.c_void_p()
        check(lib().hb_ek_deserialize(data, len(data), device, C.byref(h)))
        ev._h = h
        p = HbParams()
        check(lib().hb_ek_params(h, C.byref(p)))
        ev.params = Parameters.of(p.set)
        return ev

    def export_keys(self) -> tuple[np.ndarray, np.ndarray]:
        bw, kw = C.c_uint64(), C.c_uint64()
        check(lib().hb_key_words(self.params.set_id, C.byref(bw), C.byref(kw)))
        bsk = np.zeros(bw.value, dtype=np.uint64)
        ksk = np.zeros(kw.value, dtype=np.uint64)
        check(lib().hb_ek_export(self._h, ptr(bsk), ptr(ksk)))
        return bsk, ksk

    def pair_check(self) -> None:
        """Raises if a pair-mode wait of this key's launches since the last check timed out
        (hb_ek_pair_check); call once the streams of device-buffer calls are synchronised."""
        check(lib().hb_ek_pair_check(self._h))

    def _gate(self, gate: str, c0, c1, c2=None) -> LWECiphertext:
        out = np.zeros(self.params.n + 1, dtype=np.uint64)
        check(lib().hb_gate(self._h, GATES[gate], ptr(c0.words), ptr(c1.words),
                            ptr(c2.words) if c2 is not None else None, ptr(out)))
        return LWECiphertext(out)

    def nand(self, c0: LWECiphertext, c1: LWECiphertext) -> LWECiphertext:
        out = np.zeros(self.params.n + 1, dtype=np.uint64)
        check(lib().hb_nand(self._h, ptr(c0.words), ptr(c1.words), ptr(out)))
        return LWECiphertext(out)

    def and_(self, c0, c1):
        return self._gate("and", c0, c1)

    def or_(self, c0, c1):
        return self._gate("or", c0, c1)

    def nor(self, c0, c1):
        return self._gate("nor", c0, c1)

    def xor(self, c0, c1):
        return self._gate("xor", c0, c1)

    def xnor(self, c0, c1):
        return self._gate("xnor", c0, c1)

    def majority(self, c0, c1, c2):
        return self._gate("majority", c0, c1, c2)

    def not_(self, c: LWECiphertext) -> LWECiphertext:
        out = np.zeros(self.params.n + 1, dtype=np.uint64)
        check(lib().hb_not(self._h, ptr(c.words), ptr(out)))
        return LWECiphertext(out)

    def mux(self, c0, c1, c2):
        """evaluate.rs:245-263: if c0 then c1 else c2, as (c0 AND c1) + (NOT c0 AND c2)
        bootstrapped with the OR LUT (hb_gate with HB_GATE_MUX, all on the device)."""
        return self._gate("mux", c0, c1, c2)

    def lut(self, kind: str) -> np.ndarray:
        """Test polynomials of evaluate.rs:267-465 (lut_step = 1), canonical words."""
        N, q = self.params.N, self.params.q
        q8 = q >> 3
        pos, neg = q8, q - q8
        mid, lo, hi = {
            "nand": ((N >> 1) + (N >> 2), pos, neg), "and": ((N >> 1) + (N >> 2), neg, pos),
            "majority": ((N >> 1) + (N >> 2), neg, pos), "or": (N >> 2, neg, pos),
            "nor": (N >> 2, pos, neg), "xor": (N >> 1, neg, pos), "xnor": (N >> 1, pos, neg),
        }[kind]
        v = np.full(N, hi, dtype=np.uint64)
        v[:mid] = lo
        return v

    def bootstrap(self, c: LWECiphertext, lut: np.ndarray) -> LWECiphertext:
        out = np.zeros(self.params.n + 1, dtype=np.uint64)
        lut = np.ascontiguousarray(lut, np.uint64)
        check(lib().hb_bootstrap(self._h, ptr(c.words), ptr(lut), ptr(out)))
        return LWECiphertext(out)

    def gate_batch(self, gate: str, c0s: np.ndarray, c1s: np.ndarray,
                   c2s: np.ndarray | None = None) -> np.ndarray:
        """`count` independent gates in one launch; arrays of shape (count, n+1)."""
        c0s = np.ascontiguousarray(c0s, np.uint64)
        c1s = np.ascontiguousarray(c1s, np.uint64)
        c2s = None if c2s is None else np.ascontiguousarray(c2s, np.uint64)
        out = np.zeros_like(c0s)
        check(lib().hb_gate_batch(self._h, GATES[gate], c0s.shape[0], ptr(c0s), ptr(c1s),
                                  ptr(c2s), ptr(out)))
        return out

    def nand_batch(self, c0s: np.ndarray, c1s: np.ndarray) -> np.ndarray:
        return self.gate_batch("nand", c0s, c1s)

    def autotune(self, probe_gates: int = 1024) -> dict:
        """Times the batched kernel's variants for this parameter set and keeps the
        fastest (hb_autotune). ms is keyed by (gates per workgroup, waves per SIMD)."""
        r = HbTuneResult()
        check(lib().hb_autotune(self._h, probe_gates, C.byref(r)))
        return {"gates_per_workgroup": r.gates_per_workgroup,
                "waves_per_simd": r.waves_per_simd, "probe_gates": r.probe_gates,
                "ms": {v: r.ms[i] for i, v in enumerate(TUNE_VARIANTS)}}

    def batch_graph(self, gate: str, count: int):
        """A captured (hipGraph) batch launch over device buffers owned by the returned
        BatchGraph; see BatchGraph."""
        return BatchGraph(self, gate, count)

    # ---- traced gates (vfhe::Evaluator, vfhe/src/bfhe/evaluate.rs:47-139) ----
    def gate_batch_traced(self, gate: str, c0s, c1s, c2s=None):
        """Gates with the traced path's semantics; returns (outputs, core traces) where the
        core trace of gate g is the accumulator entering each round plus the final one,
        shape (count, n+1, 2, N)."""
        c0s = np.ascontiguousarray(c0s, np.uint64)
        c1s = np.ascontiguousarray(c1s, np.uint64)
        c2s = None if c2s is None else np.ascontiguousarray(c2s, np.uint64)
        P = self.params
        out = np.zeros_like(c0s)
        core = np.zeros((c0s.shape[0], P.n + 1, 2, P.N), dtype=np.uint64)
        check(lib().hb_gate_batch_traced(self._h, GATES[gate], c0s.shape[0], ptr(c0s), ptr(c1s),
                                         ptr(c2s), ptr(out), ptr(core)))
        return out, core

    def bootstrap_traced(self, cs, lut):
        """Traced bootstrap of ciphertexts cs (count, n+1) with one LUT; (outputs, cores)."""
        cs = np.ascontiguousarray(cs, np.uint64)
        lut = np.ascontiguousarray(lut, np.uint64)
        P = self.params
        out = np.zeros_like(cs)
        core = np.zeros((cs.shape[0], P.n + 1, 2, P.N), dtype=np.uint64)
        check(lib().hb_bootstrap_traced(self._h, cs.shape[0], ptr(cs), ptr(lut), ptr(out),
                                        ptr(core)))
        return out, core

    def expand_trace(self, gate: str, c0, c1, core, rounds, c2=None, with_tail: bool = True):
        """PBSTrace components of one gate rebuilt from its core trace: one record per listed
        round and the tail record (modulus switch, key switch, sample extraction)."""
        rw, tw = C.c_uint64(), C.c_uint64()
        check(lib().hb_trace_sizes(self.params.set_id, C.byref(rw), C.byref(tw)))
        rounds = np.ascontiguousarray(rounds, np.uint32)
        rec = np.zeros((rounds.size, rw.value), dtype=np.uint64)
        tail = np.zeros(tw.value, dtype=np.uint64) if with_tail else None
        core = np.ascontiguousarray(core, np.uint64)
        c0 = np.ascontiguousarray(c0, np.uint64)
        c1 = None if c1 is None else np.ascontiguousarray(c1, np.uint64)
        c2 = None if c2 is None else np.ascontiguousarray(c2, np.uint64)
        gate_id = 7 if gate == "bootstrap" else GATES[gate]
        check(lib().hb_trace_expand(self._h, gate_id, ptr(c0), ptr(c1), ptr(c2), ptr(core),
                                    rounds.ctypes.data_as(C.POINTER(C.c_uint32)), rounds.size,
                                    ptr(rec), ptr(tail)))
        return rec, tail

    def trace_part(self, gate: str, c0, c1, core, part, index: int = 0, c2=None) -> np.ndarray:
        """One finalized PBSTrace container (a name of TRACE_PARTS or its id) of one gate,
        built from its core trace; shape (planes, rows, N), or (planes, rows) for the
        monomial representation, flat for permutation_info and modulus_switching."""
        part_id = TRACE_PARTS.index(part) if isinstance(part, str) else int(part)
        w = C.c_uint64()
        check(lib().hb_trace_part_words(self.params.set_id, part_id, C.byref(w)))
        out = np.zeros(w.value, dtype=np.uint64)
        core = np.ascontiguousarray(core, np.uint64)
        c0 = np.ascontiguousarray(c0, np.uint64)
        c1 = None if c1 is None else np.ascontiguousarray(c1, np.uint64)
        c2 = None if c2 is None else np.ascontiguousarray(c2, np.uint64)
        gate_id = 7 if gate == "bootstrap" else GATES[gate]
        check(lib().hb_trace_part(self._h, gate_id, ptr(c0), ptr(c1), ptr(c2), ptr(core),
                                  part_id, index, ptr(out)))
        N = self.params.N
        name = TRACE_PARTS[part_id]
        if name in ("permutation_info", "modulus_switching"):
            return out
        if name == "monomial_representation":
            return out.reshape(2, -1)
        planes = {"monomial": 2, "hadamard": 6}.get(name, 4)
        return out.reshape(planes, -1, N)


class BatchGraph:
    """hb_batch_graph_*: `count` gates of one kind over fixed device buffers, captured once
    as a hipGraph and replayed. set_inputs() stages host ciphertexts, launch() replays,
    outputs() copies the results back."""

    def __init__(self, ev: Evaluator, gate: str, count: int):
        L = lib()
        self.ev, self.count = ev, count
        self.n1 = ev.params.n + 1
        nb = count * self.n1 * 8
        self._bufs = []
        for _ in range(4 if gate == "majority" else 3):
            p = C.c_void_p()
            check(L.hb_dev_alloc(nb, C.byref(p)))
            self._bufs.append(p)
        c2 = self._bufs[3] if gate == "majority" else None
        g = C.c_void_p()
        check(L.hb_batch_graph_create(ev._h, GATES[gate], count, self._bufs[0], self._bufs[1],
                                      c2, self._bufs[2], C.byref(g)))
        self._g = g

    def set_inputs(self, c0s, c1s, c2s=None):
        """Stages host ciphertexts; waits for replays in flight (which read the buffers)."""
        L = lib()
        check(L.hb_sync())
        for buf, a in ((self._bufs[0], c0s), (self._bufs[1], c1s),
                       (self._bufs[3] if c2s is not None else None, c2s)):
            if a is None:
                continue
            a = np.ascontiguousarray(a, np.uint64)
            assert a.shape == (self.count, self.n1)
            check(L.hb_memcpy_h2d(buf, a.ctypes.data, a.nbytes))

    def launch(self, stream=None):
        check(lib().hb_batch_graph_launch(self._g, stream))

    def outputs(self) -> np.ndarray:
        """Results of the last replay (waits for it: replays run on a non-blocking stream,
        which the copy's null stream does not wait for by itself)."""
        out = np.zeros((self.count, self.n1), dtype=np.uint64)
        check(lib().hb_sync())
        check(lib().hb_memcpy_d2h(out.ctypes.data, self._bufs[2], out.nbytes))
        return out

    @property
    def device_buffers(self):
        return self._bufs

    def close(self):
        L = lib()
        if getattr(self, "_g", None):
            L.hb_batch_graph_destroy(self._g)
            self._g = None
        for b in getattr(self, "_bufs", []):
            L.hb_dev_free(b)
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tuned_config(set_id: int) -> tuple[int, int, bool]:
    """(gates per workgroup, waves per SIMD, whether it was tuned) for a parameter set."""
    g, w, t = C.c_int(), C.c_int(), C.c_int()
    check(lib().hb_tuned_config(set_id, C.byref(g), C.byref(w), C.byref(t)))
    return g.value, w.value, bool(t.value)


def set_tuned_config(set_id: int, gates_per_workgroup: int, waves_per_simd: int) -> None:
    check(lib().hb_set_tuned_config(set_id, gates_per_workgroup, waves_per_simd))


def latency_mode(set_id: int = SET_GOLDILOCKS) -> int:
    """Largest batch of the parameter set that runs in latency mode (0: never)."""
    v = C.c_uint32()
    check(lib().hb_latency_mode_for(set_id, C.byref(v)))
    return v.value


def set_latency_mode(max_gates: int, set_id: int | None = None) -> None:
    """Latency-mode cutoff of one parameter set, or of all when set_id is None;
    max_gates = LATENCY_DEFAULT restores the built-in cutoff."""
    if set_id is None:
        check(lib().hb_set_latency_mode(max_gates))
    else:
        check(lib().hb_set_latency_mode_for(set_id, max_gates))


def latency_modes() -> tuple[int, int, int]:
    """The cutoffs of all three parameter sets (for restore_latency_modes)."""
    return tuple(latency_mode(s) for s in (SET_BABYBEAR, SET_GOLDILOCKS, SET_ZAMA))


def restore_latency_modes(modes) -> None:
    for s, m in zip((SET_BABYBEAR, SET_GOLDILOCKS, SET_ZAMA), modes):
        set_latency_mode(m, s)


def pair_mode() -> bool:
    """Whether small batches run one gate on two compute units (hb_pair_mode)."""
    v = C.c_int()
    check(lib().hb_pair_mode(C.byref(v)))
    return bool(v.value)


def set_pair_mode(enable: bool) -> None:
    check(lib().hb_set_pair_mode(1 if enable else 0))


def set_pair_wait_limit(microseconds: int = PAIR_WAIT_DEFAULT) -> None:
    """How long a pair-mode wait lasts before it times out (hb_set_pair_wait_limit); a
    timed-out launch is reported as an error. PAIR_WAIT_DEFAULT restores the default (2 s),
    0 makes every wait time out at once."""
    check(lib().hb_set_pair_wait_limit(microseconds))


def mixed_batch(parts) -> list[np.ndarray]:
    """Bootstrap gates of several evaluation keys (any parameter sets) in ONE launch.

    parts: iterable of (evaluator, gate_name, c0s, c1s[, c2s]) with arrays of shape
    (count, n+1). Returns the output arrays in the same order."""
    from ._lib import HbBatchPart
    parts = list(parts)
    descs = (HbBatchPart * max(1, len(parts)))()
    keep, outs = [], []
    for i, part in enumerate(parts):
        ev, gate, c0s, c1s = part[:4]
        c2s = part[4] if len(part) > 4 else None
        c0s = np.ascontiguousarray(c0s, np.uint64)
        c1s = np.ascontiguousarray(c1s, np.uint64)
        c2s = None if c2s is None else np.ascontiguousarray(c2s, np.uint64)
        out = np.zeros_like(c0s)
        keep += [c0s, c1s, c2s]
        outs.append(out)
        d = descs[i]
        d.ek = ev._h
        d.gate = GATES[gate]
        d.count = c0s.shape[0]
        d.c0s, d.c1s = c0s.ctypes.data, c1s.ctypes.data
        d.c2s = c2s.ctypes.data if c2s is not None else None
        d.outs = out.ctypes.data
    check(lib().hb_mixed_batch(descs, len(parts)))
    return outs


def last_kernel_ms() -> float:
    v = C.c_double()
    check(lib().hb_last_kernel_ms(C.byref(v)))
    return v.value


def device_count() -> int:
    n = C.c_int()
    check(lib().hb_device_count(C.byref(n)))
    return n.value
