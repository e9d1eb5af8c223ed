"""Serving loop front end: circuits, the server handle and a Poisson load generator.

The scheduling itself runs natively (hb_server_* in the C ABI): requests are admitted by
an HBM budget, their wires stay on the device between layers, and the ready gates of all
admitted circuits share kernel launches (continuous batching).
"""
from __future__ import annotations

import ctypes as C
import threading
import time
from dataclasses import dataclass

import numpy as np

from ._lib import HbCircuit, HbServerStats, check, lib, ptr

NAND = 0
KINDS = {"nand": 0, "and": 1, "or": 2, "nor": 3, "xor": 4, "xnor": 5}
_PLAIN = {0: lambda a, b: 1 - (a & b), 1: lambda a, b: a & b, 2: lambda a, b: a | b,
          3: lambda a, b: 1 - (a | b), 4: lambda a, b: a ^ b, 5: lambda a, b: 1 - (a ^ b)}


@dataclass
class Circuit:
    """Gate DAG over wires: inputs are wires 0..n_inputs-1, gate g writes n_inputs+g."""
    n_inputs: int
    in0: np.ndarray
    in1: np.ndarray
    kind: np.ndarray
    outputs: np.ndarray

    @property
    def n_gates(self) -> int:
        return int(self.in0.size)

    def depth(self) -> int:
        lv = np.zeros(self.n_inputs + self.n_gates, dtype=np.int64)
        for g in range(self.n_gates):
            lv[self.n_inputs + g] = 1 + max(lv[self.in0[g]], lv[self.in1[g]])
        return int(lv.max())

    def eval_plain(self, bits) -> np.ndarray:
        w = np.zeros(self.n_inputs + self.n_gates, dtype=np.int64)
        w[:self.n_inputs] = np.asarray(bits, dtype=np.int64)
        for g in range(self.n_gates):
            w[self.n_inputs + g] = _PLAIN[int(self.kind[g])](int(w[self.in0[g]]), int(w[self.in1[g]]))
        return w[self.outputs]


def random_nand_circuit(rng: np.random.Generator, depth: int, width: int) -> Circuit:
    """`depth` layers of `width` NAND gates; each gate reads two distinct wires of the
    previous layer (the inputs for layer 1); the outputs are the last layer."""
    in0, in1 = [], []
    prev = np.arange(width)
    base = width
    for _ in range(depth):
        a = rng.integers(0, width, size=width)
        b = (a + rng.integers(1, width, size=width)) % width
        in0.append(prev[a])
        in1.append(prev[b])
        prev = base + np.arange(width)
        base += width
    in0 = np.concatenate(in0).astype(np.uint32)
    in1 = np.concatenate(in1).astype(np.uint32)
    return Circuit(width, in0, in1, np.zeros(in0.size, np.uint8), prev.astype(np.uint32))


class Server:
    """hb_server_t wrapper. Thread-safe: submit/wait may be called from many threads."""

    def __init__(self, device: int = 0, max_gates_per_launch: int = 2048, hbm_budget: int = 0):
        h = C.c_void_p()
        check(lib().hb_server_create(device, max_gates_per_launch, hbm_budget, C.byref(h)))
        self._h = h
        self._keep = {}
        self._lock = threading.Lock()

    def close(self):
        if getattr(self, "_h", None):
            check(lib().hb_server_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, evaluator, circuit: Circuit, inputs: np.ndarray, traced: bool = False) -> int:
        """traced: run with the vfhe semantics and keep every gate's core trace."""
        inputs = np.ascontiguousarray(inputs, np.uint64)
        arrays = [np.ascontiguousarray(a) for a in
                  (circuit.in0.astype(np.uint32), circuit.in1.astype(np.uint32),
                   circuit.kind.astype(np.uint8), circuit.outputs.astype(np.uint32))]
        c = HbCircuit(circuit.n_inputs, circuit.n_gates, int(circuit.outputs.size), 0,
                      arrays[0].ctypes.data, arrays[1].ctypes.data, arrays[2].ctypes.data,
                      arrays[3].ctypes.data)
        rid = C.c_uint64()
        check(lib().hb_server_submit_traced(self._h, evaluator._h, C.byref(c), ptr(inputs),
                                            int(traced), C.byref(rid)))
        P = evaluator.params
        with self._lock:
            self._keep[rid.value] = (int(circuit.outputs.size), P.n + 1, circuit.n_gates, P.N,
                                     traced)
        return rid.value

    def wait(self, rid: int):
        """(outputs, latency_ms), plus the core traces (n_gates, n+1, 2, N) for traced requests."""
        with self._lock:
            n_out, n1, n_gates, N, traced = self._keep.pop(rid)
        out = np.zeros((n_out, n1), dtype=np.uint64)
        lat = C.c_double()
        core = np.zeros((n_gates, n1, 2, N), dtype=np.uint64) if traced else None
        check(lib().hb_server_wait_traced(self._h, rid, ptr(out), ptr(core), C.byref(lat)))
        if traced:
            return out, lat.value, core
        return out, lat.value

    def done(self, rid: int) -> bool:
        d = C.c_int()
        check(lib().hb_server_poll(self._h, rid, C.byref(d)))
        return bool(d.value)

    def stats(self) -> dict:
        s = HbServerStats()
        check(lib().hb_server_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in HbServerStats._fields_}


def run_load(server: Server, clients, rate: float, duration: float, rng: np.random.Generator,
             depth=(1, 64), width=(2, 64), verify: bool = True, log=None,
             circuits: int | None = None) -> dict:
    """Poisson arrivals at `rate` circuits/s for `duration` seconds, or of exactly
    `circuits` circuits when that is given. Each arrival picks a client (evaluator,
    encryptor, decryptor) uniformly, a NAND circuit with depth and width uniform in the
    given inclusive ranges, and uniform input bits. Returns latency percentiles (with the
    sample count), throughput and the number of outputs that decrypt wrongly."""
    t0 = time.perf_counter()
    pending = []
    t_next = t0
    total_gates = 0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if circuits is None and t_next - t0 > duration:
            break
        if circuits is not None and len(pending) >= circuits:
            break
        while time.perf_counter() < t_next:
            time.sleep(min(0.002, max(0.0, t_next - time.perf_counter())))
        ci = int(rng.integers(0, len(clients)))
        ev, enc, dec = clients[ci]
        d = int(rng.integers(depth[0], depth[1] + 1))
        w = int(rng.integers(width[0], width[1] + 1))
        circ = random_nand_circuit(rng, d, w)
        bits = rng.integers(0, 2, size=w)
        rid = server.submit(ev, circ, enc.encrypt_many(bits))
        pending.append((rid, circ, bits, dec))
        total_gates += circ.n_gates
    t_arrivals = time.perf_counter() - t0
    lat, wrong = [], 0
    for rid, circ, bits, dec in pending:
        out, ms = server.wait(rid)
        lat.append(ms)
        if verify:
            expect = circ.eval_plain(bits)
            got = [dec.decrypt_with_noise(out[k])[0] for k in range(out.shape[0])]
            wrong += int(np.sum(np.asarray(got) != expect))
    elapsed = time.perf_counter() - t0
    res = {"rate": rate, "circuits": len(pending), "latency_samples": len(lat),
           "gates": total_gates,
           "elapsed_s": round(elapsed, 3), "arrival_window_s": round(t_arrivals, 3),
           "gates_per_s": round(total_gates / elapsed, 1) if elapsed > 0 else 0.0,
           "p50_ms": round(float(np.percentile(lat, 50)), 1) if lat else None,
           "p99_ms": round(float(np.percentile(lat, 99)), 1) if lat else None,
           "wrong_outputs": wrong}
    if log:
        log(f"serving rate {rate}/s: {res}")
    return res
