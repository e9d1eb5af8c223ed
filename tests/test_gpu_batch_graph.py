"""Config 3 pieces: the captured (hipGraph) batch launch and the autotuned kernel variants.

Every variant (1, 2, 4 gates per workgroup) and every graph replay must give the outputs
of the eager launch, which the oracle pins bit-exactly."""
import sys

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gold(hb, keyset):
    P, sk, bsk, ksk = keyset(1)
    return P, sk, bsk, ksk, hb.Evaluator(keys=(bsk, ksk), params=P)


_INPUTS = {}


def _inputs(hb, sk, seed, count):
    """Bits and ciphertexts, encrypted once per (seed, count): encryption draws fresh
    randomness on every call, so the cache is what makes the batch identical."""
    if (seed, count) not in _INPUTS:
        rng = stream(seed, "graph-bits")
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        _INPUTS[(seed, count)] = (b0, b1, enc.encrypt_many(b0), enc.encrypt_many(b1))
    return _INPUTS[(seed, count)]


@pytest.mark.parametrize("variant", [(1, 1), (2, 1), (4, 1), (2, 2), (4, 2), (4, 3)])
def test_variants_match_oracle(hb, orc, gold, variant):
    P, sk, bsk, ksk, ev = gold
    say(f"variant gates/workgroup, waves/SIMD = {variant}")
    # more gates than SIMDs and latency mode off, so the pinned variant is the one
    # launched; odd, so the last workgroup of the 2- and 4-gate variants is partly empty
    count = 1024 + 9
    b0, b1, c0, c1 = _inputs(hb, sk, 50, count)  # same batch for every variant
    old, modes = hb.tuned_config(hb.SET_GOLDILOCKS), hb.latency_modes()
    hb.set_tuned_config(hb.SET_GOLDILOCKS, *variant)
    hb.set_latency_mode(0)
    try:
        out = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_tuned_config(hb.SET_GOLDILOCKS, old[0], old[1])
        hb.restore_latency_modes(modes)
    sample = np.r_[0:6, count - 6:count]  # both ends of the grid, against the oracle
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=12)
    assert np.array_equal(out[sample], ref)
    dec = hb.Decryptor(sk)
    assert all(dec.decrypt(hb.LWECiphertext(out[i])) == 1 - (int(b0[i]) & int(b1[i]))
               for i in range(count))
    _VARIANT_OUT[variant] = out


_VARIANT_OUT = {}


def test_variants_agree():
    """All six variants give the same words on the same batch (runs after them)."""
    outs = list(_VARIANT_OUT.values())
    if len(outs) < 6:
        pytest.skip("needs the six variant runs of this module")
    assert all(np.array_equal(o, outs[0]) for o in outs[1:])


def test_graph_replays_match_eager(hb, gold):
    """Several replays with eager launches and host synchronisation in between; new inputs
    are staged between replays, and every replay is compared with the eager result."""
    P, sk, bsk, ksk, ev = gold
    count = 40
    g = ev.batch_graph("nand", count)
    dec = hb.Decryptor(sk)
    try:
        for rep in range(4):
            say(f"graph replay {rep}")
            b0, b1, c0, c1 = _inputs(hb, sk, 60 + rep, count)
            g.set_inputs(c0, c1)
            g.launch()
            eager = ev.gate_batch("nand", c0, c1)  # eager work between replays
            got = g.outputs()
            assert np.array_equal(got, eager), rep
            assert all(dec.decrypt(hb.LWECiphertext(got[i])) == 1 - (int(b0[i]) & int(b1[i]))
                       for i in range(count))
        # replaying the same inputs twice gives the same words
        g.launch()
        again = g.outputs()
        g.launch()
        assert np.array_equal(g.outputs(), again)
    finally:
        g.close()


def test_autotune_picks_a_variant_and_caches(hb, gold, tmp_path, monkeypatch):
    P, sk, bsk, ksk, ev = gold
    say("autotune")
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("HB_AUTOTUNE_CACHE", str(cache))
    r = ev.autotune(probe_gates=256)
    best = (r["gates_per_workgroup"], r["waves_per_simd"])
    assert best in r["ms"]
    assert all(ms > 0 for ms in r["ms"].values())
    assert min(r["ms"], key=r["ms"].get) == best
    assert hb.tuned_config(hb.SET_GOLDILOCKS) == (*best, True)
    assert f"1 {best[0]} {best[1]}" in cache.read_text().splitlines()


_EVALUATORS = {}


@pytest.mark.parametrize("set_id,count", [(0, 70), (0, 260), (2, 70), (1, 260)])
def test_latency_mode_wave_counts(hb, orc, keyset, set_id, count):
    """Latency mode runs 8 wavefronts per gate for the Goldilocks sets and for launches of
    at most one gate per CU, 4 for larger BabyBear launches (260 gates is more than one per
    CU on MI355X's 256). Both kernels give the words of the one-wave kernel and the oracle."""
    P, sk, bsk, ksk = keyset(set_id)
    say(f"latency mode, set {set_id}, {count} gates")
    if set_id not in _EVALUATORS:
        _EVALUATORS[set_id] = hb.Evaluator(keys=(bsk, ksk), params=P)
    ev = _EVALUATORS[set_id]
    b0, b1, c0, c1 = _inputs_for(hb, sk, set_id, 91, count)
    old = hb.latency_modes()
    try:
        hb.set_latency_mode(1 << 20)
        split = ev.gate_batch("nand", c0, c1)
        hb.set_latency_mode(0)
        single = ev.gate_batch("nand", c0, c1)
    finally:
        hb.restore_latency_modes(old)
    assert np.array_equal(split, single)
    sample = np.r_[0:3, count - 3:count]
    ref = orc.OracleKeys(set_id, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=6)
    assert np.array_equal(split[sample], ref)
    dec = hb.Decryptor(sk)
    assert all(dec.decrypt(hb.LWECiphertext(split[i])) == 1 - (int(b0[i]) & int(b1[i]))
               for i in range(count))


def _inputs_for(hb, sk, set_id, seed, count):
    key = (set_id, seed, count)
    if key not in _INPUTS:
        rng = stream(seed + 1000 * set_id, "latency-bits")
        enc = hb.Encryptor(sk)
        b0, b1 = rng.integers(0, 2, size=(2, count))
        _INPUTS[key] = (b0, b1, enc.encrypt_many(b0), enc.encrypt_many(b1))
    return _INPUTS[key]


@pytest.mark.parametrize("variant", [(2, 2), (4, 3)])
def test_babybear_variants_match_oracle(hb, orc, keyset, variant):
    """BabyBear through the 2- and 3-waves-per-SIMD variants (more gates than SIMDs, the
    last workgroup partly empty) against the oracle and decryption."""
    P, sk, bsk, ksk = keyset(0)
    say(f"BabyBear variant {variant}")
    if 0 not in _EVALUATORS:
        _EVALUATORS[0] = hb.Evaluator(keys=(bsk, ksk), params=P)
    ev = _EVALUATORS[0]
    count = 1024 + 9
    b0, b1, c0, c1 = _inputs_for(hb, sk, 0, 93, count)
    old, modes = hb.tuned_config(hb.SET_BABYBEAR), hb.latency_modes()
    hb.set_tuned_config(hb.SET_BABYBEAR, *variant)
    hb.set_latency_mode(0)
    try:
        out = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_tuned_config(hb.SET_BABYBEAR, old[0], old[1])
        hb.restore_latency_modes(modes)
    sample = np.r_[0:4, count - 4:count]
    ref = orc.OracleKeys(0, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=8)
    assert np.array_equal(out[sample], ref)
    dec = hb.Decryptor(sk)
    assert all(dec.decrypt(hb.LWECiphertext(out[i])) == 1 - (int(b0[i]) & int(b1[i]))
               for i in range(count))


def test_latency_mode_matches_wave_per_gate(hb, orc, gold):
    """The latency-mode kernel and the 1-wave kernel give the same words."""
    P, sk, bsk, ksk, ev = gold
    count = 70
    b0, b1, c0, c1 = _inputs(hb, sk, 77, count)
    old = hb.latency_modes()
    try:
        hb.set_latency_mode(1 << 20)
        split = ev.gate_batch("nand", c0, c1)
        hb.set_latency_mode(0)
        single = ev.gate_batch("nand", c0, c1)
    finally:
        hb.restore_latency_modes(old)
    assert np.array_equal(split, single)
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0[:8], c1[:8], threads=8)
    assert np.array_equal(split[:8], ref)


def test_partly_filled_round_runs_its_rest_in_latency_mode(hb, orc, gold):
    """2048 + 300 Goldilocks gates: with the (4, 2) variant the whole round of 2048 resident
    gates runs batched and the 300 left in latency mode after it (two launches); every
    output equals the one batched launch's (latency mode off), a sample the oracle's."""
    P, sk, bsk, ksk, ev = gold
    count = 2048 + 300
    b0, b1, c0, c1 = _inputs(hb, sk, 51, count)
    old, modes = hb.tuned_config(hb.SET_GOLDILOCKS), hb.latency_modes()
    hb.set_tuned_config(hb.SET_GOLDILOCKS, 4, 2)
    try:
        say(f"{count} gates, default launch shape")
        shaped = ev.gate_batch("nand", c0, c1)
        hb.set_latency_mode(0)
        say(f"{count} gates, one batched launch")
        whole = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_tuned_config(hb.SET_GOLDILOCKS, old[0], old[1])
        hb.restore_latency_modes(modes)
    assert np.array_equal(shaped, whole)
    sample = np.r_[2040:2056, count - 4:count]  # both sides of the split
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=12)
    assert np.array_equal(shaped[sample], ref)


@pytest.mark.parametrize("count", [1, 7, 128])
def test_pair_mode_matches_latency_mode(hb, orc, gold, count):
    """Goldilocks gates on two CUs each (pair mode: the external product's halves cross
    through global memory every round) give the latency-mode words, and the oracle's."""
    P, sk, bsk, ksk, ev = gold
    say(f"pair mode, {count} gates")
    b0, b1, c0, c1 = _inputs(hb, sk, 80, count)
    assert hb.pair_mode()
    paired = ev.gate_batch("nand", c0, c1)
    hb.set_pair_mode(False)
    try:
        split = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_pair_mode(True)
    assert np.array_equal(paired, split)
    sample = np.unique(np.r_[0:min(2, count), max(0, count - 2):count])
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=4)
    assert np.array_equal(paired[sample], ref)
    dec = hb.Decryptor(sk)
    assert all(dec.decrypt(hb.LWECiphertext(paired[i])) == 1 - (int(b0[i]) & int(b1[i]))
               for i in range(count))


@pytest.mark.parametrize("count", [1, 96])
def test_pair_mode_zama_matches_latency_mode(hb, orc, keyset, count):
    """Zama (13 five-bit digits, one or two per wave) in pair mode: the latency-mode words
    and the oracle's."""
    P, sk, bsk, ksk = keyset(2)
    say(f"pair mode zama, {count} gates")
    if 2 not in _EVALUATORS:
        _EVALUATORS[2] = hb.Evaluator(keys=(bsk, ksk), params=P)
    ev = _EVALUATORS[2]
    b0, b1, c0, c1 = _inputs_for(hb, sk, 2, 81, count)
    paired = ev.gate_batch("nand", c0, c1)
    hb.set_pair_mode(False)
    try:
        split = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_pair_mode(True)
    assert np.array_equal(paired, split)
    sample = np.unique(np.r_[0:1, count - 1:count])
    ref = orc.OracleKeys(2, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=2)
    assert np.array_equal(paired[sample], ref)


@pytest.mark.parametrize("count", [1, 33])
def test_pair_mode_babybear_matches_latency_mode(hb, orc, keyset, count):
    """BabyBear in pair mode (one digit row per wave 0-3 of each workgroup): the latency-mode
    words and the oracle's (config 1's single NAND runs this path)."""
    P, sk, bsk, ksk = keyset(0)
    say(f"pair mode babybear, {count} gates")
    if 0 not in _EVALUATORS:
        _EVALUATORS[0] = hb.Evaluator(keys=(bsk, ksk), params=P)
    ev = _EVALUATORS[0]
    b0, b1, c0, c1 = _inputs_for(hb, sk, 0, 82, count)
    paired = ev.gate_batch("nand", c0, c1)
    hb.set_pair_mode(False)
    try:
        split = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_pair_mode(True)
    assert np.array_equal(paired, split)
    sample = np.unique(np.r_[0:1, count - 1:count])
    ref = orc.OracleKeys(0, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=2)
    assert np.array_equal(paired[sample], ref)


def test_pair_wait_timeout_is_an_error(hb, keyset):
    """A pair-mode wait that runs out is reported as an error by the call that ran it.
    Forced from the host: with 0 polls per wait every wait of the launch fails at once (every
    wave still reaches the end). With the default limit restored the same call gives the
    latency-mode words, and a device-buffer launch is checked through pair_check."""
    from hasteboots_amd._lib import HbError
    P, sk, bsk, ksk = keyset(1)
    say("pair wait limit 0: expecting an error")
    if 1 not in _EVALUATORS:
        _EVALUATORS[1] = hb.Evaluator(keys=(bsk, ksk), params=P)
    ev = _EVALUATORS[1]
    b0, b1, c0, c1 = _inputs_for(hb, sk, 1, 83, 4)
    assert hb.pair_mode()
    hb.set_pair_wait_limit(0)
    try:
        with pytest.raises(HbError) as ei:
            ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_pair_wait_limit()
    assert "timed out" in str(ei.value)
    good = ev.gate_batch("nand", c0, c1)
    ev.pair_check()
    hb.set_pair_mode(False)
    try:
        split = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_pair_mode(True)
    assert np.array_equal(good, split)


def test_pair_launch_beside_another_keys_launch_on_another_stream(hb, keyset):
    """A pair-mode launch (8 gates of one key) issued on one stream while a 512-gate
    latency-mode launch of another key runs on a second stream, with no host synchronisation
    between them: the pair workgroups start as compute units free up and pair in the order
    they start, so both launches give the words they give alone and no wait times out."""
    import ctypes as C
    from hasteboots_amd._lib import check, lib
    L = lib()
    P, sk_a, bsk_a, ksk_a = keyset(1)
    _, sk_b, bsk_b, ksk_b = keyset(1, 1)
    ev_a = hb.Evaluator(keys=(bsk_a, ksk_a), params=P)
    ev_b = hb.Evaluator(keys=(bsk_b, ksk_b), params=P)
    say("pair launch beside another stream's latency-mode launch")
    _, _, a0, a1 = _inputs_for(hb, sk_a, 1, 84, 512)
    _, _, p0, p1 = _inputs_for(hb, sk_b, 1, 85, 8)
    expect_a = ev_a.gate_batch("nand", a0, a1)
    expect_b = ev_b.gate_batch("nand", p0, p1)

    def dev(x):
        d = C.c_void_p()
        check(L.hb_dev_alloc(x.nbytes, C.byref(d)))
        check(L.hb_memcpy_h2d(d, x.ctypes.data, x.nbytes))
        return d

    streams = [C.c_void_p(), C.c_void_p()]
    for s in streams:
        check(L.hb_stream_create(C.byref(s)))
    bufs = [(dev(a0), dev(a1), dev(np.zeros_like(a0))), (dev(p0), dev(p1), dev(np.zeros_like(p0)))]
    try:
        for rep in range(2):
            (d0, d1, do), (e0, e1, eo) = bufs
            check(L.hb_gate_batch_dev(ev_a._h, 0, 512, d0, d1, None, do, streams[0]))
            check(L.hb_gate_batch_dev(ev_b._h, 0, 8, e0, e1, None, eo, streams[1]))
            for s in streams:
                check(L.hb_stream_sync(s))
            ev_b.pair_check()
            for (_, _, o), e in zip(bufs, (expect_a, expect_b)):
                got = np.zeros_like(e)
                check(L.hb_memcpy_d2h(got.ctypes.data, o, got.nbytes))
                assert np.array_equal(got, e), rep
    finally:
        for trio in bufs:
            for d in trio:
                L.hb_dev_free(d)
        for s in streams:
            L.hb_stream_destroy(s)
        del ev_a, ev_b
