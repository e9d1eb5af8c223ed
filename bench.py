"""Benchmark: batched NAND gate bootstrapping on one MI355X.

Workload (default): 4096 independent NAND gates with GOLDILOCKS_BINARY_128_BITS_PARAMETERS
in one kernel launch, keys (BSK + KSK) resident in HBM, ciphertexts already in HBM when
the timed region starts. One step = one such launch. Also reported in the same run:
  plain   — the same kernel with one gate per launch (bootstraps/s and p50/p99 latency),
  roofline — key-stream + I/O bytes per launch over the kernel's event-timed duration vs
             8 TB/s: every gate's wavefront streams the whole BSK + KSK for itself (tiles of
             one gate per key load), so a launch of G gates makes G key-stream passes,
  cpu_baseline — the CPU oracle (a restatement of the reference path) on the host cores,
  single_nand — one NAND per call for each parameter set on the GPU and on one CPU thread,
             beside the reference's published single-NAND latencies.

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

SETS = {"babybear": 0, "goldilocks": 1, "zama": 2}
SET_NAMES = {0: "BABYBEAR_BINARY_128_BITS", 1: "GOLDILOCKS_BINARY_128_BITS",
             2: "ZAMA_BINARY_128_BITS"}
METRIC = "NAND bootstraps/sec, batched vs plain on 1 MI355X; p50/p99 gate latency"
HBM_PEAK = 8.0e12
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def modmuls_per_gate(P) -> int:
    """Modular multiplications one bootstrap performs in this engine (all rounds active):
    per round 2*br_len forward + 2 inverse NTTs of (N/2)log2(N) butterflies and 4*br_len*N
    MAC products; key switching ks_len forward + 1 inverse NTT and 2*ks_len*N products."""
    N, logn = P.N, P.N.bit_length() - 1
    ntt = (N // 2) * logn
    per_round = (2 * P.br_len + 2) * ntt + 4 * P.br_len * N
    ks = (P.ks_len + 1) * ntt + 2 * P.ks_len * N
    return P.n * per_round + ks


def profiled_traffic(set_name: str, gates: int):
    """HBM bytes per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE passes (scaled by a
    copy kernel of known size), recorded in profiles/ for this workload; None if absent."""
    p = ROOT / "profiles" / f"bootstrap_kernel_{set_name}_{gates}_hbm_traffic.md"
    if not p.exists():
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", p.read_text())
    return (int(m.group(1)), str(p.relative_to(ROOT))) if m else (None, None)


def microbench(L, check, hb, reps: int = 3) -> dict:
    """Kernel microbenchmarks on Goldilocks: the MAC peak of this build (modmul/s), the
    negacyclic NTT over 4096 polynomials for N = 1024 and 2048 (forward and inverse,
    bytes = read + write of every word), and the external product of 4096 RLWEs with one
    RGSW of the Zama shape (B = 2^5, 13 digits). Times are HIP events on the launch stream."""
    P = (1 << 64) - (1 << 32) + 1
    out = {}
    mps, ms = C.c_double(), C.c_double()
    check(L.hb_modmul_peak(8192, 1024, C.byref(mps), C.byref(ms)))
    out["modmul_peak_per_s"] = mps.value
    out["modmul_peak_ms"] = round(ms.value, 3)
    rng = np.random.default_rng([2, 2])
    count = 4096

    def dev(a):
        p = C.c_void_p()
        check(L.hb_dev_alloc(a.nbytes, C.byref(p)))
        check(L.hb_memcpy_h2d(p, a.ctypes.data, a.nbytes))
        return p

    for log_n in (10, 11):
        N = 1 << log_n
        d = dev(rng.integers(0, P, size=(count, N), dtype=np.uint64))
        for inv in (0, 1):
            check(L.hb_ntt_dev(1, inv, log_n, d, count, None))
            check(L.hb_sync())
            t = []
            for _ in range(reps):
                check(L.hb_ntt_dev(1, inv, log_n, d, count, None))
                t.append(hb.last_kernel_ms())
            kms = float(np.median(t))
            out[f"ntt_{'inv' if inv else 'fwd'}_N{N}"] = {
                "polys": count, "ms": round(kms, 4),
                "GB_per_s": round(2 * 8 * N * count / (kms * 1e-3) / 1e9, 1)}
        L.hb_dev_free(d)
    N, bits, length = 1024, 5, 13
    rgsw = dev(rng.integers(0, P, size=2 * length * 2 * N, dtype=np.uint64))
    rl = dev(rng.integers(0, P, size=(count, 2 * N), dtype=np.uint64))
    check(L.hb_external_product_dev(1, 10, bits, length, rgsw, rl, count, None))
    check(L.hb_sync())
    t = []
    for _ in range(reps):
        check(L.hb_external_product_dev(1, 10, bits, length, rgsw, rl, count, None))
        t.append(hb.last_kernel_ms())
    kms = float(np.median(t))
    alg = count * 4 * N * 8 + 2 * length * 2 * N * 8
    mm = count * ((2 * length + 2) * (N // 2) * 10 + 4 * length * N)
    out["external_product_zama_shape"] = {
        "rlwes": count, "ms": round(kms, 4),
        "GB_per_s": round(alg / (kms * 1e-3) / 1e9, 1),
        "modmul_per_s": round(mm / (kms * 1e-3), 1)}
    L.hb_dev_free(rgsw)
    L.hb_dev_free(rl)
    return out


def profiled_valu(set_name: str, gates: int):
    """VALU instructions per gate (wave64 instructions) counted by rocprofv3 SQ_INSTS_VALU
    for this workload, recorded in profiles/; (None, None) if absent."""
    p = ROOT / "profiles" / f"bootstrap_kernel_{set_name}_{gates}_valu.md"
    if not p.exists():
        return None, None
    m = re.search(r"valu_instructions_per_gate:\s*(\d+)", p.read_text())
    return (int(m.group(1)), str(p.relative_to(ROOT))) if m else (None, None)


def mixed_leg(hb, seed: int = 41) -> dict:
    """Config 4: one launch bootstrapping NAND gates of BabyBear, Goldilocks and Zama keys
    together; per-set counts ~ Geometric(mean 300) truncated at 2048, uniform bits.
    Host buffers in and out (hb_mixed_batch); the kernel time is the launch's HIP events."""
    rng = np.random.default_rng([seed, 4])
    counts = [int(min(2048, rng.geometric(1 / 300))) for _ in range(3)]
    parts, checks = [], []
    for set_id, count in zip((0, 1, 2), counts):
        P = hb.Parameters.of(set_id)
        sk = hb.KeyGen.generate_secret_key(P, seed=100 + set_id)
        ev = hb.Evaluator(sk)
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        parts.append((ev, "nand", enc.encrypt_many(b0), enc.encrypt_many(b1)))
        checks.append((hb.Decryptor(sk), b0, b1))
    hb.mixed_batch(parts)  # warm-up (kernel load, scratch)
    t = time.perf_counter()
    outs = hb.mixed_batch(parts)
    wall = time.perf_counter() - t
    kms = hb.last_kernel_ms()
    wrong = 0
    for (dec, b0, b1), out in zip(checks, outs):
        wrong += sum(dec.decrypt(hb.LWECiphertext(out[i])) != 1 - (int(b0[i]) & int(b1[i]))
                     for i in range(out.shape[0]))
    total = sum(counts)
    return {"counts": {"babybear": counts[0], "goldilocks": counts[1], "zama": counts[2]},
            "gates": total, "kernel_ms": round(kms, 3),
            "gates_per_s_kernel": round(total / (kms * 1e-3), 1),
            "gates_per_s_end_to_end": round(total / wall, 1), "decrypt_mismatches": int(wrong)}


# Table 1 of the reference's performance workbook (statistics/HasteBoots_Performance.xlsx,
# "Table 2"!B4:H6): one NAND, single thread, Apple M4 Pro (inferred).
PUBLISHED_SINGLE_NAND_MS = {"zama": 160.18, "babybear": 28.497, "goldilocks": 107.96}


def single_nand_leg(hb, cpu: bool, reps: int = 5, seed: int = 61) -> dict:
    """One NAND per call for every parameter set, the published benchmark's shape: GPU
    wall time per gate (host ciphertexts in and out, latency-mode launch), and the CPU
    oracle's single-thread time on the same keys, beside the published figure."""
    out = {}
    for name, set_id in SETS.items():
        P = hb.Parameters.of(set_id)
        sk = hb.KeyGen.generate_secret_key(P, seed=seed + set_id)
        bsk, ksk = sk.generate_keys_host()
        ev = hb.Evaluator(keys=(bsk, ksk), params=P, device=0)
        enc, dec = hb.Encryptor(sk), hb.Decryptor(sk)
        rng = np.random.default_rng([seed, set_id])
        b0, b1 = rng.integers(0, 2, size=(2, reps + 1))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        ev.gate_batch("nand", c0[:1], c1[:1])  # warm-up
        lat, wrong = [], 0
        for i in range(1, reps + 1):
            s = time.perf_counter()
            o = ev.gate_batch("nand", c0[i:i + 1], c1[i:i + 1])
            lat.append((time.perf_counter() - s) * 1e3)
            wrong += dec.decrypt(hb.LWECiphertext(o[0])) != 1 - (int(b0[i]) & int(b1[i]))
        row = {"gpu_p50_ms": round(float(np.percentile(lat, 50)), 3), "wrong": int(wrong),
               "published_ms": PUBLISHED_SINGLE_NAND_MS[name]}
        if cpu:
            import oracle
            ok = oracle.OracleKeys(set_id, bsk, ksk)
            s = time.perf_counter()
            ok.gate_batch(0, c0[:1], c1[:1], threads=1)
            row["cpu_port_single_thread_ms"] = round((time.perf_counter() - s) * 1e3, 2)
        out[name] = row
        log(f"single NAND {name}: {row}")
    out["published_hardware"] = "Apple M4 Pro (inferred), single thread"
    return out


def keygen_leg(hb, L, check) -> dict:
    """Evaluation-key generation per parameter set: on the host (hb_keygen_host, up to 16
    threads) and on the GPU (hb_keygen_dev into device buffers, returns when written;
    median of 3 after one warm-up), and hb_ek_create end to end (device keygen, device-form
    conversion, canonical host copy)."""
    out = {}
    for set_id, name in SET_NAMES.items():
        P = hb.Parameters.of(set_id)
        sk = hb.KeyGen.generate_secret_key(P, seed=0)
        t = time.perf_counter()
        bsk, ksk = sk.generate_keys_host()
        host_s = time.perf_counter() - t
        bufs = [C.c_void_p(), C.c_void_p()]
        for b, a in zip(bufs, (bsk, ksk)):
            check(L.hb_dev_alloc(a.nbytes, C.byref(b)))
        ts = []
        for _ in range(4):
            t = time.perf_counter()
            check(L.hb_keygen_dev(sk._h, bufs[0], bufs[1], None))
            ts.append(time.perf_counter() - t)
        dev = np.zeros_like(ksk)
        check(L.hb_memcpy_d2h(dev.ctypes.data, bufs[1], dev.nbytes))
        same_ksk = bool(np.array_equal(dev, ksk))
        for b in bufs:
            L.hb_dev_free(b)
        t = time.perf_counter()
        ev = hb.Evaluator(sk)
        create_s = time.perf_counter() - t
        del ev
        out[name] = {"host_s": round(host_s, 3), "device_ms": round(1e3 * float(np.median(ts[1:])), 2),
                     "ek_create_s": round(create_s, 3), "key_bytes": int((bsk.size + ksk.size) * 8),
                     "ksk_equal": same_ksk}
        log(f"keygen {name}: {out[name]}")
    return out


LEGS = ("end_to_end", "single_nand", "mixed_batch", "microbench", "serving", "keygen",
        "cpu_baseline")


def compose_line(line: dict, legs: dict) -> dict:
    """The JSON line: the headline fields, then each leg's own result under its own key
    (None for a skipped leg)."""
    unknown = set(legs) - set(LEGS)
    if unknown:
        raise KeyError(f"unknown legs {sorted(unknown)}")
    return {**line, **{k: legs.get(k) for k in LEGS}}


def serving_leg(hb, P, set_id, ev, enc, dec, args) -> dict | None:
    """Poisson arrivals of NAND circuits from several clients (own keys each), a fixed
    number of circuits per arrival rate, swept past the saturation knee."""
    rates = [float(r) for r in args.serving_rates.split(",") if r.strip()]
    if not rates:
        return None
    from hasteboots_amd.serving import Server, run_load
    t = time.time()
    clients = [(ev, enc, dec)]
    for seed in range(1, args.serving_clients):
        sk_c = hb.KeyGen.generate_secret_key(P, seed=seed)
        bsk_c, ksk_c = sk_c.generate_keys_host()
        clients.append((hb.Evaluator(keys=(bsk_c, ksk_c), params=P, device=0),
                        hb.Encryptor(sk_c), hb.Decryptor(sk_c)))
    log(f"serving: {len(clients)} clients with their own keys ready in {time.time() - t:.1f}s")
    sweep = []
    for rate in rates:
        srv = Server(device=0, max_gates_per_launch=args.serving_max_gates)
        res = run_load(srv, clients, rate, 0.0, np.random.default_rng([7, int(rate * 1000)]),
                       circuits=args.serving_circuits)
        st = srv.stats()
        srv.close()
        res.update({"gate_p50_ms": round(st["gate_p50_ms"], 1),
                    "gate_p99_ms": round(st["gate_p99_ms"], 1),
                    "launches": st["launches"], "batched_launches": st["batched_launches"],
                    "max_launch_gates": st["max_launch_gates"],
                    "kernel_busy_ms": round(st["busy_ms"], 1)})
        log(f"serving rate {rate}/s: {res}")
        sweep.append(res)
    best = max(sweep, key=lambda r: r["gates_per_s"])
    return {"clients": len(clients), "parameter_set": SET_NAMES[set_id],
            "max_gates_per_launch": args.serving_max_gates,
            "circuits_per_rate": args.serving_circuits,
            "circuit_depth": [1, 64], "circuit_width": [2, 64], "arrivals": "poisson",
            "latency": "circuit = submit to outputs on host; gate = inputs ready to "
                       "launch done",
            "peak_gates_per_s": best["gates_per_s"], "peak_at_rate": best["rate"],
            "sweep": sweep}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--set", choices=list(SETS), default="goldilocks")
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--plain", type=int, default=8, help="gates timed on the plain path")
    ap.add_argument("--check", type=int, default=256, help="gates compared with the oracle")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--serving-rates", default="4,8,12,16,20,24,32",
                    help="comma-separated circuit arrival rates (1/s) for the serving leg; "
                         "empty to skip")
    ap.add_argument("--serving-circuits", type=int, default=100,
                    help="circuits drawn at every serving rate")
    ap.add_argument("--serving-max-gates", type=int, default=4096,
                    help="the serving loop's max_gates_per_launch")
    ap.add_argument("--serving-clients", type=int, default=8)
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the gates-per-workgroup autotuner before the batched leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager launches instead of replays of the captured hipGraph")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-parameter leg")
    ap.add_argument("--no-single", action="store_true",
                    help="skip the one-NAND-per-call leg of every parameter set")
    ap.add_argument("--no-keygen", action="store_true",
                    help="skip the host / device key-generation timings")
    ap.add_argument("--no-micro", action="store_true",
                    help="skip the NTT / external-product / MAC-peak microbenchmarks")
    ap.add_argument("--only", choices=["micro", "launches", "serving"], default=None,
                    help="run one leg alone and print its JSON: 'micro' the microbenchmarks, "
                         "'launches' --steps eager launches of --batch gates with the default "
                         "launch rules (the kernel a rocprofv3 pass of that size measures), "
                         "'serving' the serving sweep")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the batched leg, copy a device buffer of this size once "
                         "(a reference kernel for scaling HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path "
              "(gates are independent and would shard across GPUs with no exchange). Run with "
              "--gpus 1.",
              file=sys.stderr)
        return 2

    import hasteboots_amd as hb
    from hasteboots_amd._lib import check, lib

    if hb.device_count() < 1:
        log("no GPU visible")
        return 1
    L = lib()
    if args.copy_probe_mib > 0:
        import torch
        nb = args.copy_probe_mib << 20
        src = torch.ones(nb, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nb} bytes device to device")
        del src, dst
    if args.only == "micro":
        micro = microbench(L, check, hb)
        print(json.dumps({"leg": "micro", "microbench": micro}), flush=True)
        return 0
    set_id = SETS[args.set]
    P = hb.Parameters.of(set_id)
    G, n1 = args.batch, P.n + 1

    t = time.time()
    sk = hb.KeyGen.generate_secret_key(P, seed=0)
    bsk, ksk = sk.generate_keys_host()
    ev = hb.Evaluator(keys=(bsk, ksk), params=P, device=0)
    log(f"keygen+upload {time.time() - t:.2f}s ({args.set})")

    if args.only == "serving":
        serving = serving_leg(hb, P, set_id, ev, hb.Encryptor(sk), hb.Decryptor(sk), args)
        print(json.dumps({"leg": "serving", "serving": serving}), flush=True)
        return 0

    rng = np.random.default_rng([0, 3])
    b0 = rng.integers(0, 2, size=G)
    b1 = rng.integers(0, 2, size=G)
    enc = hb.Encryptor(sk)
    c0 = enc.encrypt_many(b0)
    c1 = enc.encrypt_many(b1)

    def dev(arr):
        p = C.c_void_p()
        check(L.hb_dev_alloc(arr.nbytes, C.byref(p)))
        check(L.hb_memcpy_h2d(p, arr.ctypes.data, arr.nbytes))
        return p

    d0, d1 = dev(c0), dev(c1)
    dout = dev(np.zeros_like(c0))

    def launch(count, a=d0, b=d1, o=dout):
        check(L.hb_gate_batch_dev(ev._h, 0, count, a, b, None, o, None))

    if args.only == "launches":
        launch(G)  # warm-up
        check(L.hb_sync())
        kms = []
        for _ in range(args.steps):
            launch(G)
            kms.append(hb.last_kernel_ms())
        out = np.zeros_like(c0)
        check(L.hb_memcpy_d2h(out.ctypes.data, dout, out.nbytes))
        dec = hb.Decryptor(sk)
        wrong = sum(dec.decrypt(hb.LWECiphertext(out[i])) != 1 - (int(b0[i]) & int(b1[i]))
                    for i in range(G))
        print(json.dumps({"leg": "launches", "parameter_set": SET_NAMES[set_id], "gates": G,
                          "latency_mode_max_gates": hb.latency_mode(),
                          "kernel_ms": [round(x, 3) for x in kms],
                          "decrypt_mismatches": int(wrong)}), flush=True)
        for p in (d0, d1, dout):
            L.hb_dev_free(p)
        return 0

    tune = None
    if not args.no_autotune:
        tune = ev.autotune(probe_gates=G)
        log(f"autotune: {tune['gates_per_workgroup']} gates per workgroup, "
            f"{tune['waves_per_simd']} waves per SIMD (probe ms "
            f"{', '.join(f'{k}:{v:.1f}' for k, v in tune['ms'].items())})")
        tune["ms"] = {f"{k[0]}x{k[1]}": round(v, 3) for k, v in tune["ms"].items()}

    graph = None
    if not args.no_graph:
        graph = C.c_void_p()
        check(L.hb_batch_graph_create(ev._h, 0, G, d0, d1, None, dout, C.byref(graph)))

    def launch_batch():
        if graph is not None:
            check(L.hb_batch_graph_launch(graph, None))
        else:
            launch(G)

    for _ in range(args.warmup):
        launch_batch()
        check(L.hb_sync())
    log(f"warmup done ({args.warmup} launches)")

    kernel_ms = []
    check(L.hb_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        launch_batch()
        kernel_ms.append(hb.last_kernel_ms())
    check(L.hb_sync())
    t1 = time.perf_counter()
    wall = t1 - t0
    value = G * args.steps / wall
    log(f"batched: {value:.1f} bootstraps/s, kernel {np.mean(kernel_ms):.1f} ms/launch")

    # correctness of the measured launch
    out = np.zeros_like(c0)
    check(L.hb_memcpy_d2h(out.ctypes.data, dout, out.nbytes))
    dec = hb.Decryptor(sk)
    wrong = sum(dec.decrypt(hb.LWECiphertext(out[i])) != 1 - (int(b0[i]) & int(b1[i]))
                for i in range(G))
    oracle_ok = None
    if args.check > 0:
        import oracle
        k = min(args.check, G)
        ref = oracle.OracleKeys(set_id, bsk, ksk).gate_batch(
            0, c0[:k], c1[:k], threads=min(16, os.cpu_count() or 1))
        oracle_ok = bool(np.array_equal(ref, out[:k]))
    log(f"decrypt mismatches {wrong}/{G}; oracle bit-exact on {args.check}: {oracle_ok}")

    # end to end: host ciphertexts in, host results out (PCIe both ways), one call
    e2e = None
    if args.steps > 0:
        out_h = np.zeros_like(c0)
        ev.gate_batch("nand", c0, c1)  # scratch sized
        s = time.perf_counter()
        reps = 2
        for _ in range(reps):
            out_h = ev.gate_batch("nand", c0, c1)
        el = time.perf_counter() - s
        e2e = {"value": round(G * reps / el, 2), "unit": "bootstraps/s",
               "includes": "H2D of both inputs, kernel, D2H of outputs (hb_gate_batch)",
               "identical_to_device_path": bool(np.array_equal(out_h, out))}
        log(f"end to end (host buffers): {e2e['value']} bootstraps/s")

    # plain path: one gate per launch, after one untimed launch (the first call of a
    # path allocates its flag rows and loads its kernel)
    lat = []
    if args.plain:
        launch(1, d0, d1, dout)
    for i in range(args.plain):
        check(L.hb_sync())
        s = time.perf_counter()
        launch(1, C.c_void_p(d0.value + i * n1 * 8), C.c_void_p(d1.value + i * n1 * 8),
               C.c_void_p(dout.value + i * n1 * 8))
        check(L.hb_sync())
        lat.append((time.perf_counter() - s) * 1e3)
    plain = None
    if lat:
        plain_value = 1e3 / np.mean(lat)
        plain = {"value": round(plain_value, 3), "unit": "bootstraps/s",
                 "p50_ms": round(float(np.percentile(lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat, 99)), 3), "gates": args.plain}
        log(f"plain: {plain_value:.2f} bootstraps/s, p50 {np.percentile(lat, 50):.1f} ms")

    serving = serving_leg(hb, P, set_id, ev, enc, dec, args)

    mixed = None
    if not args.no_mixed:
        mixed = mixed_leg(hb)
        log(f"mixed batch: {mixed}")

    single = None
    if not args.no_single:
        single = single_nand_leg(hb, cpu=not args.no_cpu)

    micro = None
    if not args.no_micro:
        micro = microbench(L, check, hb)
        log(f"microbench: {micro}")

    keygen = None if args.no_keygen else keygen_leg(hb, L, check)

    wb = 8 if P.is_goldilocks else 4
    key_bytes = (bsk.size + ksk.size) * wb
    io_bytes = G * 3 * n1 * 8
    # Tiles of T = 1 gate per key load: each wavefront streams the keys for its own gate
    # (the gates of a workgroup share them only through the caches).
    gates_per_key_load = 1
    key_passes = -(-G // gates_per_key_load)
    traffic_alg = key_passes * key_bytes + io_bytes
    kms = float(np.mean(kernel_ms))
    achieved = traffic_alg / (kms * 1e-3)
    one_pass = (key_bytes + io_bytes) / (kms * 1e-3)
    mm = modmuls_per_gate(P)
    traffic, traffic_src = profiled_traffic(args.set, G)
    valu_per_gate, valu_src = profiled_valu(args.set, G)
    valu_rate = valu_per_gate * G / (kms * 1e-3) if valu_per_gate else None

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(16, os.cpu_count() or 1)
        ok = oracle.OracleKeys(set_id, bsk, ksk)
        s = time.perf_counter()
        ok.gate_batch(0, c0[:1], c1[:1], threads=1)
        cpu_single_s = time.perf_counter() - s
        per_thread = max(1, int(round(args.cpu_seconds / max(cpu_single_s, 1e-3))))
        per_thread = min(per_thread, max(1, G // threads))
        S = threads * per_thread
        s = time.perf_counter()
        ok.gate_batch(0, c0[:S], c1[:S], threads=threads)
        el = time.perf_counter() - s
        cpu = {"value": round(S / el, 3), "unit": "bootstraps/s", "cores": threads,
               "kind": "port",
               "sample": f"{S} NAND gates ({args.set}) on {threads} threads, one gate per "
                         f"thread task; single-thread latency {cpu_single_s * 1e3:.1f} ms",
               "single_thread_ms": round(cpu_single_s * 1e3, 2), "cpu_model": cpu_model(),
               "nproc": os.cpu_count()}
        log(f"cpu baseline: {cpu['value']} bootstraps/s on {threads} threads, "
            f"single {cpu_single_s * 1e3:.0f} ms")

    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "bootstraps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64" if P.is_goldilocks else "u32",
        "data": "synthetic (seeded keys and uniform random input bits)",
        "config": {"workload": f"batched NAND bootstrap, {SET_NAMES[set_id]}, {G} gates per "
                               f"{'eager' if graph is None else 'hipGraph-replayed'} launch, "
                               "keys resident in HBM",
                   "parameter_set": SET_NAMES[set_id], "gates_per_launch": G,
                   "n": P.n, "N": P.N, "br_bits": P.br_bits, "br_len": P.br_len,
                   "ks_bits": P.ks_bits, "ks_len": P.ks_len},
        "launch": "hipGraph replay" if graph is not None else "eager",
        "autotune": tune,
        "batched_vs_plain": round(value / plain["value"], 1) if plain else None,
        "plain": plain,
        "batch_latency_ms": {"p50": round(float(np.percentile(kernel_ms, 50)), 3),
                             "p99": round(float(np.percentile(kernel_ms, 99)), 3)},
        "kernel_ms_per_launch": round(kms, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 3),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 5), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": traffic_alg,
                     "key_bytes_per_pass": key_bytes, "key_stream_passes": key_passes,
                     "gates_per_key_load": gates_per_key_load, "io_bytes": io_bytes,
                     "one_pass_achieved": round(one_pass / 1e9, 3),
                     "one_pass_frac": round(one_pass / HBM_PEAK, 6),
                     "counter_achieved": round(traffic / (kms * 1e-3) / 1e9, 3) if traffic
                     else None,
                     "counter_frac": round(traffic / (kms * 1e-3) / HBM_PEAK, 4) if traffic
                     else None},
        "valu": {"bound": "valu_issue",
                 "achieved": round(valu_rate, 1) if valu_rate else None,
                 "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
                 "frac": round(valu_rate / VALU_PEAK, 4) if valu_rate else None,
                 "instructions_per_gate": valu_per_gate, "instructions_source": valu_src,
                 "peak_basis": "256 CUs x 4 SIMDs x one wave64 VALU instruction per 4 clocks "
                               "at 2.4 GHz",
                 "modmul_per_gate": mm,
                 "achieved_modmul_per_s": round(mm * G / (kms * 1e-3), 1),
                 "mac_peak_modmul_per_s": round(micro["modmul_peak_per_s"], 1) if micro
                 else None,
                 "mac_peak_source": "hb_modmul_peak: register-resident Montgomery MAC pairs, "
                                    "8192 wavefronts"},
        "correct": {"decrypt_mismatches": wrong, "oracle_bit_exact": oracle_ok,
                    "oracle_gates": args.check},
    }
    line = compose_line(line, {"end_to_end": e2e, "single_nand": single, "mixed_batch": mixed,
                               "microbench": micro, "serving": serving, "keygen": keygen,
                               "cpu_baseline": cpu})
    print(json.dumps(line), flush=True)
    if graph is not None:
        L.hb_batch_graph_destroy(graph)
    for p in (d0, d1, dout):
        L.hb_dev_free(p)
    return 0


if __name__ == "__main__":
    sys.exit(main())
