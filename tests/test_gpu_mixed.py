"""Mixed-parameter batches: gates of BabyBear, Goldilocks and Zama-Goldilocks keys in one
kernel launch, each output bit-exact with the oracle for its own parameter set."""
import sys

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu

TRUTH = {"nand": lambda a, b, c: 1 - (a & b), "and": lambda a, b, c: a & b,
         "or": lambda a, b, c: a | b, "xor": lambda a, b, c: a ^ b,
         "majority": lambda a, b, c: int(a + b + c >= 2)}
GATE_ID = {"nand": 0, "and": 1, "or": 2, "nor": 3, "xor": 4, "xnor": 5, "majority": 6}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def evaluators(hb, keyset):
    evs = {}
    for s in (0, 1, 2):
        P, sk, bsk, ksk = keyset(s)
        evs[s] = (P, sk, bsk, ksk, hb.Evaluator(keys=(bsk, ksk), params=P))
    return evs


def test_mixed_small_all_gates_bit_exact(hb, orc, evaluators):
    say("mixed small")
    rng = stream(40, "mixed-small")
    parts, meta = [], []
    for s, gate, count in [(0, "nand", 20), (1, "xor", 10), (2, "and", 6), (0, "majority", 8),
                           (1, "nand", 0), (2, "or", 5)]:
        P, sk, bsk, ksk, ev = evaluators[s]
        enc = hb.Encryptor(sk)
        bits = rng.integers(0, 2, size=(3, count))
        cs = [enc.encrypt_many(b) if count else np.zeros((0, P.n + 1), np.uint64) for b in bits]
        parts.append((ev, gate, cs[0], cs[1], cs[2] if gate == "majority" else None))
        meta.append((s, gate, bits, cs))
    outs = hb.mixed_batch(parts)
    for (s, gate, bits, cs), out in zip(meta, outs):
        if out.shape[0] == 0:
            continue
        P, sk, bsk, ksk, ev = evaluators[s]
        ok = orc.OracleKeys(s, bsk, ksk)
        ref = ok.gate_batch(GATE_ID[gate], cs[0], cs[1], cs[2] if gate == "majority" else None,
                            threads=16)
        assert np.array_equal(out, ref), (s, gate)
        dec = hb.Decryptor(sk)
        for i in range(out.shape[0]):
            expect = TRUTH[gate](*(int(bits[k][i]) for k in range(3)))
            assert dec.decrypt(hb.LWECiphertext(out[i])) == expect


def test_mixed_latency_mode_and_wave_per_gate_agree(hb, orc, evaluators):
    """The same mixed launch through the latency-mode kernel (8 wavefronts per gate, the
    default at this size) and the one-wavefront-per-gate kernel (latency mode off) gives the
    same words, which the oracle pins on a sample."""
    say("mixed: latency mode against one wave per gate")
    rng = stream(17, "mixed-modes")
    parts, meta = [], []
    for s, gate, count in [(2, "nand", 40), (1, "or", 25), (0, "nand", 60)]:
        P, sk, bsk, ksk, ev = evaluators[s]
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        parts.append((ev, gate, c0, c1, None))
        meta.append((s, gate, c0, c1))
    old = hb.latency_modes()
    try:
        hb.set_latency_mode(1 << 20)
        split = hb.mixed_batch(parts)
        hb.set_latency_mode(0)
        single = hb.mixed_batch(parts)
    finally:
        hb.restore_latency_modes(old)
    for (s, gate, c0, c1), a, b in zip(meta, split, single):
        assert np.array_equal(a, b), (s, gate)
        P, sk, bsk, ksk, ev = evaluators[s]
        ref = orc.OracleKeys(s, bsk, ksk).gate_batch(GATE_ID[gate], c0[:2], c1[:2], threads=2)
        assert np.array_equal(a[:2], ref), (s, gate)


def test_mixed_geometric_counts_config(hb, orc, evaluators):
    """Per-set counts ~ Geometric(mean 300) truncated at 2048, uniform bits, one launch.
    Every output decrypts to NAND; a seeded sample per set is compared word for word with
    the oracle, and a few gates evaluated alone equal their results inside the mix."""
    rng = stream(41, "mixed-config")
    counts = [int(min(2048, rng.geometric(1 / 300))) for _ in range(3)]
    say(f"mixed geometric counts {counts}")
    parts, data = [], []
    for s, count in zip((0, 1, 2), counts):
        P, sk, bsk, ksk, ev = evaluators[s]
        enc = hb.Encryptor(sk)
        b0 = rng.integers(0, 2, size=count)
        b1 = rng.integers(0, 2, size=count)
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        parts.append((ev, "nand", c0, c1))
        data.append((b0, b1, c0, c1))
    outs = hb.mixed_batch(parts)
    for s, out, (b0, b1, c0, c1) in zip((0, 1, 2), outs, data):
        P, sk, bsk, ksk, ev = evaluators[s]
        dec = hb.Decryptor(sk)
        bad = sum(dec.decrypt(hb.LWECiphertext(out[i])) != 1 - (int(b0[i]) & int(b1[i]))
                  for i in range(out.shape[0]))
        assert bad == 0, (s, bad)
        pick = np.unique(rng.integers(0, out.shape[0], size=12))
        ref = orc.OracleKeys(s, bsk, ksk).gate_batch(0, c0[pick], c1[pick], threads=16)
        assert np.array_equal(out[pick], ref), s
        alone = ev.nand(hb.LWECiphertext(c0[pick[0]]), hb.LWECiphertext(c1[pick[0]]))
        assert np.array_equal(alone.words, out[pick[0]])


def test_mixed_throughput_shape_every_output(hb, orc, keyset):
    """701 BabyBear + 903 Goldilocks + 301 Zama NANDs (above one gate per SIMD, so the
    launch takes the batched 4-gates-per-workgroup shape, with parts ending inside a
    workgroup): every output equals the oracle's."""
    import os
    rng = stream(43, "mixed-batch-shape")
    counts = {hb.SET_BABYBEAR: 701, hb.SET_GOLDILOCKS: 903, hb.SET_ZAMA: 301}
    parts, data = [], []
    for s, count in counts.items():
        P, sk, bsk, ksk = keyset(s)
        ev = hb.Evaluator(keys=(bsk, ksk), params=P)
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        parts.append((ev, "nand", c0, c1))
        data.append((s, bsk, ksk, c0, c1))
    outs = hb.mixed_batch(parts)
    for (s, bsk, ksk, c0, c1), out in zip(data, outs):
        print(f"oracle set {s}: {out.shape[0]} gates", flush=True)
        ref = orc.OracleKeys(s, bsk, ksk).gate_batch(0, c0, c1, threads=min(16, os.cpu_count() or 1))
        bad = np.flatnonzero(~np.all(out == ref, axis=1))
        assert bad.size == 0, (s, bad[:8])


@pytest.mark.parametrize("set_id", [0, 1, 2], ids=["babybear", "goldilocks", "zama"])
def test_single_set_throughput_shape_matches_batched_kernel(hb, orc, keyset, set_id):
    """A mixed launch of one parameter set above one gate per SIMD runs that set's own
    kernel body (one class); two clients' parts of 550 gates each, latency mode off so that
    the launch takes the throughput shape: every output equals the dedicated launch's, and
    a sample the oracle's."""
    modes = hb.latency_modes()
    hb.set_latency_mode(0)
    try:
        _single_set_throughput_shape(hb, orc, keyset, set_id)
    finally:
        hb.restore_latency_modes(modes)


def _single_set_throughput_shape(hb, orc, keyset, set_id):
    rng = stream(44 + set_id, "single-class")
    parts, data = [], []
    for seed in (0, 1):
        P, sk, bsk, ksk = keyset(set_id, seed)
        ev = hb.Evaluator(keys=(bsk, ksk), params=P)
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, 550))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        parts.append((ev, "nand", c0, c1))
        data.append((ev, bsk, ksk, c0, c1))
    outs = hb.mixed_batch(parts)
    for (ev, bsk, ksk, c0, c1), out in zip(data, outs):
        assert np.array_equal(out, ev.gate_batch("nand", c0, c1))
        pick = np.unique(rng.integers(0, 550, size=8))
        ref = orc.OracleKeys(set_id, bsk, ksk).gate_batch(0, c0[pick], c1[pick], threads=8)
        assert np.array_equal(out[pick], ref)
