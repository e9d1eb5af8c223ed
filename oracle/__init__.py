"""CPU oracle (test infrastructure only): ctypes wrapper of oracle/liboracle.so.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module.
It checks the HIP product; the product never calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class Params(C.Structure):
    _fields_ = [("set", C.c_int), ("is_goldilocks", C.c_int), ("n", C.c_uint32),
                ("N", C.c_uint32), ("br_bits", C.c_uint32), ("ks_bits", C.c_uint32),
                ("br_len", C.c_uint32), ("ks_len", C.c_uint32), ("q", C.c_uint64),
                ("t", C.c_uint32)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB


_lib = None

# Finalized PBSTrace container ids (the order of orc_pbs_trace_part's part argument).
TRACE_PARTS = ("initial_acc", "final_acc", "input_acc", "output_acc", "input_acc_permuted",
               "monomial", "monomial_representation", "monomial_times_acc",
               "external_product_input", "hadamard", "sum_prod", "permutation_info",
               "modulus_switching")


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.orc_get_params.argtypes = [C.c_int, C.POINTER(Params)]
        l.orc_gl_op.argtypes = [C.c_int, u64p, u64p, u64p, C.c_size_t]
        l.orc_bb_op.argtypes = [C.c_int, u64p, u64p, u64p, C.c_size_t]
        l.orc_ntt.argtypes = [C.c_int, C.c_int, C.c_uint32, u64p, C.c_size_t]
        l.orc_ntt_psi.argtypes = [C.c_int, C.c_uint32]
        l.orc_ntt_psi.restype = C.c_uint64
        l.orc_negacyclic_schoolbook.argtypes = [C.c_int, C.c_uint32, u64p, u64p, u64p]
        l.orc_mod_switch.argtypes = [C.POINTER(Params), u64p, u32p]
        l.orc_monomial_sub_one.argtypes = [C.c_int, C.c_uint32, C.c_uint32, u64p, u64p]
        l.orc_external_product.argtypes = [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, u64p,
                                           u64p, u64p]
        l.orc_ek_create.argtypes = [C.POINTER(Params), u64p, u64p]
        l.orc_ek_create.restype = C.c_void_p
        l.orc_ek_destroy.argtypes = [C.c_void_p]
        l.orc_blind_rotate.argtypes = [C.c_void_p, u32p, u64p, u64p]
        l.orc_key_switch.argtypes = [C.c_void_p, u64p, u64p]
        l.orc_bootstrap.argtypes = [C.c_void_p, u64p, u64p, u64p]
        l.orc_gate_batch.argtypes = [C.c_void_p, C.c_int, C.c_size_t, u64p, u64p, u64p, u64p,
                                     C.c_int]
        l.orc_build_lut.argtypes = [C.POINTER(Params), C.c_int, u64p]
        l.orc_decrypt_with_noise.argtypes = [C.POINTER(Params), u64p, u64p, u64p, u64p]
        l.orc_trace_sizes.argtypes = [C.POINTER(Params), u64p, u64p]
        l.orc_bootstrap_traced.argtypes = [C.c_void_p, u64p, u64p, u64p, u32p, C.c_uint32, u64p,
                                           u64p]
        l.orc_pbs_trace_part_words.argtypes = [C.POINTER(Params), C.c_int, u64p]
        l.orc_pbs_trace_part.argtypes = [C.c_void_p, u64p, u64p, C.c_int, C.c_uint32, u64p]
        l.orc_raw_neg_events.argtypes = [C.c_int]
        l.orc_raw_neg_events.restype = C.c_uint64
        _lib = l
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def params(set_id: int) -> Params:
    p = Params()
    assert lib().orc_get_params(set_id, C.byref(p)) == 0
    return p


def gl_op(op: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.uint64)
    out = np.zeros_like(x)
    assert lib().orc_gl_op(op, _p(x), _p(y), _p(out), x.size) == 0
    return out


def bb_op(op: int, x, y=None) -> np.ndarray:
    x = np.ascontiguousarray(x, np.uint64)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.uint64)
    out = np.zeros_like(x)
    assert lib().orc_bb_op(op, _p(x), _p(y), _p(out), x.size) == 0
    return out


def ntt(is_goldilocks: bool, data: np.ndarray, inverse: bool = False) -> np.ndarray:
    d = np.ascontiguousarray(data, np.uint64).copy()
    n = d.shape[-1]
    log_n = n.bit_length() - 1
    count = d.size // n
    assert lib().orc_ntt(int(is_goldilocks), int(inverse), log_n, _p(d), count) == 0
    return d


def ntt_psi(is_goldilocks: bool, log_n: int) -> int:
    return int(lib().orc_ntt_psi(int(is_goldilocks), log_n))


def schoolbook(is_goldilocks: bool, a, b) -> np.ndarray:
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    c = np.zeros_like(a)
    assert lib().orc_negacyclic_schoolbook(int(is_goldilocks), a.size, _p(a), _p(b), _p(c)) == 0
    return c


def mod_switch(p: Params, c) -> np.ndarray:
    c = np.ascontiguousarray(c, np.uint64)
    out = np.zeros(p.n + 1, dtype=np.uint32)
    assert lib().orc_mod_switch(C.byref(p), _p(c), out.ctypes.data_as(u32p)) == 0
    return out


def monomial_sub_one(is_goldilocks: bool, y, r: int) -> np.ndarray:
    y = np.ascontiguousarray(y, np.uint64)
    x = np.zeros_like(y)
    assert lib().orc_monomial_sub_one(int(is_goldilocks), y.size, r, _p(y), _p(x)) == 0
    return x


def external_product(is_goldilocks: bool, bits: int, length: int, rgsw, a, b):
    rgsw = np.ascontiguousarray(rgsw, np.uint64)
    a = np.ascontiguousarray(a, np.uint64).copy()
    b = np.ascontiguousarray(b, np.uint64).copy()
    assert lib().orc_external_product(int(is_goldilocks), a.size, bits, length, _p(rgsw), _p(a),
                                      _p(b)) == 0
    return a, b


def raw_neg_events(reset: bool = True) -> int:
    """Goldilocks negations of a raw word above P on this thread since the last reset."""
    return int(lib().orc_raw_neg_events(int(reset)))


def build_lut(p: Params, kind: int) -> np.ndarray:
    out = np.zeros(p.N, dtype=np.uint64)
    assert lib().orc_build_lut(C.byref(p), kind, _p(out)) == 0
    return out


def decrypt_with_noise(p: Params, lwe_sk, c) -> tuple[int, int]:
    lwe_sk = np.ascontiguousarray(lwe_sk, np.uint64)
    c = np.ascontiguousarray(c, np.uint64)
    m, nz = C.c_uint64(), C.c_uint64()
    assert lib().orc_decrypt_with_noise(C.byref(p), _p(lwe_sk), _p(c),
                                        C.cast(C.byref(m), u64p), C.cast(C.byref(nz), u64p)) == 0
    return int(m.value), int(nz.value)


class OracleKeys:
    """Evaluation keys (canonical layout) held by the oracle."""

    def __init__(self, set_id: int, bsk: np.ndarray, ksk: np.ndarray):
        self.p = params(set_id)
        self.bsk = np.ascontiguousarray(bsk, np.uint64)
        self.ksk = np.ascontiguousarray(ksk, np.uint64)
        self.h = lib().orc_ek_create(C.byref(self.p), _p(self.bsk), _p(self.ksk))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ek_destroy(self.h)
            self.h = None

    def gate_batch(self, gate: int, c0s, c1s, c2s=None, threads: int = 1) -> np.ndarray:
        c0s = np.ascontiguousarray(c0s, np.uint64)
        c1s = np.ascontiguousarray(c1s, np.uint64)
        c2s = None if c2s is None else np.ascontiguousarray(c2s, np.uint64)
        out = np.zeros_like(c0s)
        assert lib().orc_gate_batch(self.h, gate, c0s.shape[0], _p(c0s), _p(c1s), _p(c2s),
                                    _p(out), threads) == 0
        return out

    def bootstrap(self, c, lut) -> np.ndarray:
        c = np.ascontiguousarray(c, np.uint64)
        lut = np.ascontiguousarray(lut, np.uint64)
        out = np.zeros(self.p.n + 1, dtype=np.uint64)
        assert lib().orc_bootstrap(self.h, _p(c), _p(lut), _p(out)) == 0
        return out

    def trace_sizes(self) -> tuple[int, int]:
        rw, tw = C.c_uint64(), C.c_uint64()
        lib().orc_trace_sizes(C.byref(self.p), C.cast(C.byref(rw), u64p), C.cast(C.byref(tw), u64p))
        return int(rw.value), int(tw.value)

    def bootstrap_traced(self, c, lut, rounds):
        """The traced (vfhe) bootstrap: (output, round records, tail record)."""
        c = np.ascontiguousarray(c, np.uint64)
        lut = np.ascontiguousarray(lut, np.uint64)
        rounds = np.ascontiguousarray(rounds, np.uint32)
        rw, tw = self.trace_sizes()
        rec = np.zeros((rounds.size, rw), dtype=np.uint64)
        tail = np.zeros(tw, dtype=np.uint64)
        out = np.zeros(self.p.n + 1, dtype=np.uint64)
        assert lib().orc_bootstrap_traced(self.h, _p(c), _p(lut), _p(out),
                                          rounds.ctypes.data_as(u32p), rounds.size, _p(rec),
                                          _p(tail)) == 0
        return out, rec, tail

    def pbs_trace_part(self, c, lut, part: int, index: int = 0) -> np.ndarray:
        """One finalized PBSTrace container (TRACE_PARTS) of the traced bootstrap of c."""
        c = np.ascontiguousarray(c, np.uint64)
        lut = np.ascontiguousarray(lut, np.uint64)
        w = C.c_uint64()
        assert lib().orc_pbs_trace_part_words(C.byref(self.p), part, C.cast(C.byref(w), u64p)) == 0
        out = np.zeros(int(w.value), dtype=np.uint64)
        assert lib().orc_pbs_trace_part(self.h, _p(c), _p(lut), part, index, _p(out)) == 0
        return out

    def blind_rotate(self, ms, lut) -> np.ndarray:
        ms = np.ascontiguousarray(ms, np.uint32)
        lut = np.ascontiguousarray(lut, np.uint64)
        out = np.zeros(2 * self.p.N, dtype=np.uint64)
        assert lib().orc_blind_rotate(self.h, ms.ctypes.data_as(u32p), _p(lut), _p(out)) == 0
        return out

    def key_switch(self, acc) -> np.ndarray:
        acc = np.ascontiguousarray(acc, np.uint64)
        out = np.zeros(self.p.n + 1, dtype=np.uint64)
        assert lib().orc_key_switch(self.h, _p(acc), _p(out)) == 0
        return out
