"""ctypes binding of libhasteboots.so (the C ABI declared in include/hb.h).

The library is the product: every bootstrap runs in its HIP kernels. Loading fails
loudly if the shared object is missing; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
# HB_LIB_PATH selects another build of the same library (e.g. for A/B timing).
LIB_PATH = Path(os.environ.get("HB_LIB_PATH", PKG / "libhasteboots.so"))
HEADER = PKG.parent / "include" / "hb.h"

u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)


class HbParams(C.Structure):
    _fields_ = [
        ("set", C.c_int32), ("is_goldilocks", C.c_int32),
        ("n", C.c_uint32), ("N", C.c_uint32), ("t", C.c_uint32),
        ("br_bits", C.c_uint32), ("br_len", C.c_uint32),
        ("ks_bits", C.c_uint32), ("ks_len", C.c_uint32), ("pad_", C.c_uint32),
        ("q", C.c_uint64),
        ("lwe_sigma", C.c_double), ("ring_sigma", C.c_double), ("ks_sigma", C.c_double),
    ]


class HbBatchPart(C.Structure):
    _fields_ = [("ek", C.c_void_p), ("gate", C.c_int32), ("pad_", C.c_uint32),
                ("count", C.c_uint64), ("c0s", C.c_void_p), ("c1s", C.c_void_p),
                ("c2s", C.c_void_p), ("outs", C.c_void_p)]


class HbCircuit(C.Structure):
    _fields_ = [("n_inputs", C.c_uint32), ("n_gates", C.c_uint32), ("n_outputs", C.c_uint32),
                ("pad_", C.c_uint32), ("in0", C.c_void_p), ("in1", C.c_void_p),
                ("kind", C.c_void_p), ("outputs", C.c_void_p)]


class HbServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("gates_done", C.c_uint64), ("launches", C.c_uint64), ("queued", C.c_uint64),
                ("active", C.c_uint64), ("bytes_in_use", C.c_uint64), ("budget", C.c_uint64),
                ("busy_ms", C.c_double), ("gate_p50_ms", C.c_double), ("gate_p99_ms", C.c_double),
                ("batched_launches", C.c_uint64), ("max_launch_gates", C.c_uint64),
                ("pair_launches", C.c_uint64)]


class HbTuneResult(C.Structure):
    _fields_ = [("gates_per_workgroup", C.c_int32), ("waves_per_simd", C.c_int32),
                ("probe_gates", C.c_uint64), ("ms", C.c_double * 6)]


TUNE_VARIANTS = [(1, 1), (2, 1), (4, 1), (2, 2), (4, 2), (4, 3)]
LATENCY_DEFAULT = 0xFFFFFFFF  # HB_LATENCY_DEFAULT
PAIR_WAIT_DEFAULT = 0xFFFFFFFF  # HB_PAIR_WAIT_DEFAULT


class HbError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hb error {code}: {msg}")
        self.code = code


_PROTOS = {
    "hb_params": (C.c_int, [C.c_int, C.POINTER(HbParams)]),
    "hb_key_words": (C.c_int, [C.c_int, u64p, u64p]),
    "hb_sk_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "hb_sk_destroy": (C.c_int, [C.c_void_p]),
    "hb_sk_export": (C.c_int, [C.c_void_p, u64p, u64p]),
    "hb_encrypt": (C.c_int, [C.c_void_p, C.c_uint32, u64p]),
    "hb_decrypt": (C.c_int, [C.c_void_p, u64p, u32p, u64p]),
    "hb_ek_create": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "hb_ek_import": (C.c_int, [C.c_int, u64p, u64p, C.c_int, C.POINTER(C.c_void_p)]),
    "hb_ek_export": (C.c_int, [C.c_void_p, u64p, u64p]),
    "hb_ek_destroy": (C.c_int, [C.c_void_p]),
    "hb_keygen_host": (C.c_int, [C.c_void_p, u64p, u64p]),
    "hb_ek_params": (C.c_int, [C.c_void_p, C.POINTER(HbParams)]),
    "hb_nand": (C.c_int, [C.c_void_p, u64p, u64p, u64p]),
    "hb_gate": (C.c_int, [C.c_void_p, C.c_int, u64p, u64p, u64p, u64p]),
    "hb_bootstrap": (C.c_int, [C.c_void_p, u64p, u64p, u64p]),
    "hb_not": (C.c_int, [C.c_void_p, u64p, u64p]),
    "hb_gate_batch": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, u64p, u64p, u64p, u64p]),
    "hb_gate_batch_dev": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    "hb_gate_batch_dev_noalloc": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hb_batch_graph_create": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "hb_batch_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hb_batch_graph_destroy": (C.c_int, [C.c_void_p]),
    "hb_autotune": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(HbTuneResult)]),
    "hb_tuned_config": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                  C.POINTER(C.c_int)]),
    "hb_set_tuned_config": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "hb_server_create": (C.c_int, [C.c_int, C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "hb_server_destroy": (C.c_int, [C.c_void_p]),
    "hb_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(HbCircuit), u64p,
                                   C.POINTER(C.c_uint64)]),
    "hb_server_submit_traced": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(HbCircuit), u64p,
                                          C.c_int, C.POINTER(C.c_uint64)]),
    "hb_server_wait_traced": (C.c_int, [C.c_void_p, C.c_uint64, u64p, u64p,
                                        C.POINTER(C.c_double)]),
    "hb_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int)]),
    "hb_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, u64p, C.POINTER(C.c_double)]),
    "hb_server_stats": (C.c_int, [C.c_void_p, C.POINTER(HbServerStats)]),
    "hb_gate_batch_traced": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, u64p, u64p, u64p, u64p,
                                       u64p]),
    "hb_bootstrap_traced": (C.c_int, [C.c_void_p, C.c_size_t, u64p, u64p, u64p, u64p]),
    "hb_trace_sizes": (C.c_int, [C.c_int, u64p, u64p]),
    "hb_trace_expand": (C.c_int, [C.c_void_p, C.c_int, u64p, u64p, u64p, u64p, u32p, C.c_uint32,
                                  u64p, u64p]),
    "hb_keygen_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "hb_keys_serialize": (C.c_int, [C.c_int, u64p, u64p, C.c_void_p, C.c_size_t,
                                    C.POINTER(C.c_size_t)]),
    "hb_keys_deserialize": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_int), u64p, u64p]),
    "hb_ek_serialize": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "hb_ek_deserialize": (C.c_int, [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]),
    "hb_lwe_serialize": (C.c_int, [C.c_int, C.c_size_t, u64p, C.c_void_p, C.c_size_t,
                                   C.POINTER(C.c_size_t)]),
    "hb_lwe_deserialize": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_int),
                                     C.POINTER(C.c_size_t), u64p, C.c_size_t]),
    "hb_trace_part_words": (C.c_int, [C.c_int, C.c_int, u64p]),
    "hb_trace_part": (C.c_int, [C.c_void_p, C.c_int, u64p, u64p, u64p, u64p, C.c_int,
                                C.c_uint32, u64p]),
    "hb_mixed_batch": (C.c_int, [C.POINTER(HbBatchPart), C.c_size_t]),
    "hb_mixed_batch_dev": (C.c_int, [C.POINTER(HbBatchPart), C.c_size_t, C.c_void_p]),
    "hb_ntt_dev": (C.c_int, [C.c_int, C.c_int, C.c_uint32, C.c_void_p, C.c_size_t, C.c_void_p]),
    "hb_external_product_dev": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
    "hb_set_latency_mode": (C.c_int, [C.c_uint32]),
    "hb_set_latency_mode_for": (C.c_int, [C.c_int, C.c_uint32]),
    "hb_set_pair_mode": (C.c_int, [C.c_int]),
    "hb_pair_mode": (C.c_int, [C.POINTER(C.c_int)]),
    "hb_set_pair_wait_limit": (C.c_int, [C.c_uint32]),
    "hb_ek_pair_check": (C.c_int, [C.c_void_p]),
    "hb_latency_mode": (C.c_int, [C.POINTER(C.c_uint32)]),
    "hb_latency_mode_for": (C.c_int, [C.c_int, C.POINTER(C.c_uint32)]),
    "hb_modmul_peak": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]),
    "hb_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
    "hb_dev_free": (C.c_int, [C.c_void_p]),
    "hb_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "hb_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    "hb_sync": (C.c_int, []),
    "hb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "hb_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "hb_stream_destroy": (C.c_int, [C.c_void_p]),
    "hb_stream_sync": (C.c_int, [C.c_void_p]),
    "hb_last_kernel_ms": (C.c_int, [C.POINTER(C.c_double)]),
    "hb_last_error": (C.c_char_p, []),
}

_lib = None


def header_symbols() -> list[str]:
    """Function names declared in include/hb.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(hb_\w+)\s*\(", text, re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m hasteboots_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    l = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, (res, args) in _PROTOS.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args
    _lib = l
    return l


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().hb_last_error()
        raise HbError(rc, msg.decode() if msg else "")


def ptr(a) -> u64p:
    """uint64 numpy array -> POINTER(c_uint64) (None passes through)."""
    if a is None:
        return None
    return a.ctypes.data_as(u64p)
