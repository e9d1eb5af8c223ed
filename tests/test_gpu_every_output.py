"""Every output of the batched configurations against the oracle, word for word:
  - 4096 Goldilocks NAND gates through the captured (hipGraph) launch of the tuned
    4-gates-per-workgroup, 2-waves-per-SIMD kernel, replayed, all 4096 outputs;
  - the mixed launch with per-set counts ~ Geometric(mean 300) (at most 2048), every gate;
  - about 1000 BabyBear gates under the default latency-mode cutoff (one gate per SIMD),
    every output against the one-wavefront-per-gate kernel, a sample against the oracle;
plus the gates built from several bootstraps (mux) and the linear not, and batched calls
on two streams sharing one evaluation key."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu
THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def evs(hb, keyset):
    out = {}
    for s in (0, 1, 2):
        P, sk, bsk, ksk = keyset(s)
        out[s] = (P, sk, bsk, ksk, hb.Evaluator(keys=(bsk, ksk), params=P))
    return out


def test_graph_4096_tuned_variant_every_output(hb, orc, evs):
    P, sk, bsk, ksk, ev = evs[1]
    G = 4096
    rng = stream(3, "graph-4096")
    b0, b1 = rng.integers(0, 2, size=(2, G))
    enc = hb.Encryptor(sk)
    c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
    old = hb.tuned_config(hb.SET_GOLDILOCKS)
    hb.set_tuned_config(hb.SET_GOLDILOCKS, 4, 2)
    say("capturing a 4096-gate graph with variant (4, 2)")
    try:
        g = ev.batch_graph("nand", G)
    finally:
        hb.set_tuned_config(hb.SET_GOLDILOCKS, old[0], old[1])
    try:
        g.set_inputs(c0, c1)
        outs = []
        for rep in range(3):
            say(f"replay {rep}")
            g.launch()
            outs.append(g.outputs())
    finally:
        g.close()
    assert all(np.array_equal(o, outs[0]) for o in outs[1:])
    say(f"oracle on all {G} gates, {THREADS} threads")
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0, c1, threads=THREADS)
    bad = np.flatnonzero(~np.all(outs[0] == ref, axis=1))
    assert bad.size == 0, f"{bad.size} gates differ, first {bad[:8]}"


def test_mixed_geometric_every_output(hb, orc, evs):
    rng = stream(41, "mixed-every")
    counts = [int(min(2048, rng.geometric(1 / 300))) for _ in range(3)]
    say(f"mixed counts {counts}")
    parts, data = [], []
    for s, count in zip((0, 1, 2), counts):
        P, sk, bsk, ksk, ev = evs[s]
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        parts.append((ev, "nand", c0, c1))
        data.append((c0, c1))
    outs = hb.mixed_batch(parts)
    for s, out, (c0, c1) in zip((0, 1, 2), outs, data):
        P, sk, bsk, ksk, ev = evs[s]
        say(f"oracle set {s}: {out.shape[0]} gates")
        ref = orc.OracleKeys(s, bsk, ksk).gate_batch(0, c0, c1, threads=THREADS)
        bad = np.flatnonzero(~np.all(out == ref, axis=1))
        assert bad.size == 0, (s, bad[:8])


def test_babybear_default_cutoff_every_output(hb, orc, evs):
    """1000 BabyBear gates run in latency mode under the default policy (the cutoff is one
    gate per SIMD); every output equals the one-wavefront-per-gate kernel's."""
    P, sk, bsk, ksk, ev = evs[0]
    count = 1000
    modes = hb.latency_modes()
    assert hb.latency_mode(hb.SET_BABYBEAR) >= count, modes
    rng = stream(8, "bb-default")
    b0, b1 = rng.integers(0, 2, size=(2, count))
    enc = hb.Encryptor(sk)
    c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
    split = ev.gate_batch("nand", c0, c1)
    try:
        hb.set_latency_mode(0, hb.SET_BABYBEAR)
        single = ev.gate_batch("nand", c0, c1)
    finally:
        hb.restore_latency_modes(modes)
    assert hb.latency_modes() == modes
    assert np.array_equal(split, single)
    pick = np.unique(rng.integers(0, count, size=32))
    ref = orc.OracleKeys(0, bsk, ksk).gate_batch(0, c0[pick], c1[pick], threads=THREADS)
    assert np.array_equal(split[pick], ref)


def test_latency_cutoffs_are_per_set_and_restorable(hb):
    modes = hb.latency_modes()
    try:
        hb.set_latency_mode(5, hb.SET_ZAMA)
        assert hb.latency_modes() == (modes[0], modes[1], 5)
        hb.set_latency_mode(hb.LATENCY_DEFAULT)
        defaults = hb.latency_modes()
        # BabyBear, Goldilocks, Zama: Zama's heavier gates leave latency mode first
        assert defaults[0] == defaults[1] >= defaults[2] > 0
    finally:
        hb.restore_latency_modes(modes)
    assert hb.latency_modes() == modes


MUX = {(a, b, c): b if a else c for a in (0, 1) for b in (0, 1) for c in (0, 1)}


@pytest.mark.parametrize("set_id", [0, 1])
def test_mux_all_inputs_match_oracle(hb, orc, evs, set_id):
    """Evaluator::mux (evaluate.rs:245-263) for all 8 input combinations, batched in one
    call and one at a time, against the oracle's restatement and the truth table."""
    P, sk, bsk, ksk, ev = evs[set_id]
    enc, dec = hb.Encryptor(sk), hb.Decryptor(sk)
    combos = sorted(MUX)
    c = [enc.encrypt_many([k[i] for k in combos]) for i in range(3)]
    out = ev.gate_batch("mux", c[0], c[1], c[2])
    ref = orc.OracleKeys(set_id, bsk, ksk).gate_batch(8, c[0], c[1], c[2], threads=8)
    assert np.array_equal(out, ref)
    for i, k in enumerate(combos):
        m, noise = dec.decrypt_with_noise(out[i])
        assert m == MUX[k] and noise < P.q // 16, (k, m)
    one = ev.mux(hb.LWECiphertext(c[0][5]), hb.LWECiphertext(c[1][5]), hb.LWECiphertext(c[2][5]))
    assert np.array_equal(one.words, out[5])


@pytest.mark.parametrize("set_id", [0, 1, 2])
def test_not_is_linear_negation(hb, evs, set_id):
    """Evaluator::not (evaluate.rs:98-104): -c with Q/4 added to b; decrypts to the
    complement, twice gives back the input."""
    P, sk, bsk, ksk, ev = evs[set_id]
    enc, dec = hb.Encryptor(sk), hb.Decryptor(sk)
    q = P.q
    for m in (0, 1):
        c = enc.encrypt(m)
        nc = ev.not_(c)
        assert dec.decrypt(nc) == 1 - m
        expect = [(q - int(w)) % q for w in c.words]
        expect[-1] = (expect[-1] + (q >> 2)) % q
        assert [int(w) for w in nc.words] == expect
        back = ev.not_(nc)
        assert dec.decrypt(back) == m


def test_two_streams_share_one_key(hb, evs):
    """hb_gate_batch_dev on two streams of the same evaluation key, issued back to back
    with no host synchronisation: both results equal the single-stream ones (the shared
    accumulator scratch is handed over between the streams' kernels)."""
    from hasteboots_amd._lib import check, lib
    L = lib()
    P, sk, bsk, ksk, ev = evs[1]
    enc = hb.Encryptor(sk)
    rng = stream(9, "streams")
    n1 = P.n + 1
    sizes = (300, 40)  # latency mode, different grids
    ins = []
    for k, cnt in enumerate(sizes):
        b0, b1 = rng.integers(0, 2, size=(2, cnt))
        ins.append((enc.encrypt_many(b0), enc.encrypt_many(b1)))
    expect = [ev.gate_batch("nand", a, b) for a, b in ins]

    def dev(a):
        p = C.c_void_p()
        check(L.hb_dev_alloc(a.nbytes, C.byref(p)))
        check(L.hb_memcpy_h2d(p, a.ctypes.data, a.nbytes))
        return p

    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        check(L.hb_stream_create(C.byref(s)))
    bufs = [(dev(a), dev(b), dev(np.zeros_like(a))) for a, b in ins]
    try:
        for rep in range(2):
            for (d0, d1, do), s, cnt in zip(bufs, streams, sizes):
                check(L.hb_gate_batch_dev(ev._h, 0, cnt, d0, d1, None, do, s))
            for s in streams:
                check(L.hb_stream_sync(s))
            for (d0, d1, do), e in zip(bufs, expect):
                got = np.zeros_like(e)
                check(L.hb_memcpy_d2h(got.ctypes.data, do, got.nbytes))
                assert np.array_equal(got, e), rep
    finally:
        for trio in bufs:
            for p in trio:
                L.hb_dev_free(p)
        for s in streams:
            L.hb_stream_destroy(s)
