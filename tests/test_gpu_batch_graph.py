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


def _inputs(hb, sk, seed, count):
    rng = stream(seed, "graph-bits")
    enc = hb.Encryptor(sk)
    b0, b1 = rng.integers(0, 2, size=(2, count))
    return b0, b1, enc.encrypt_many(b0), enc.encrypt_many(b1)


@pytest.mark.parametrize("gpw", [1, 2, 4])
def test_variants_match_oracle(hb, orc, gold, gpw):
    P, sk, bsk, ksk, ev = gold
    say(f"variant gpw={gpw}")
    count = 9  # odd: the last workgroup of the 2- and 4-gate variants is partly empty
    b0, b1, c0, c1 = _inputs(hb, sk, 50 + gpw, count)
    old, _ = hb.tuned_config(hb.SET_GOLDILOCKS)
    hb.set_tuned_config(hb.SET_GOLDILOCKS, gpw)
    try:
        out = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_tuned_config(hb.SET_GOLDILOCKS, old)
    ref = orc.OracleKeys(1, bsk, ksk).gate_batch(0, c0, c1, threads=16)
    assert np.array_equal(out, ref)


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
    assert r["gates_per_workgroup"] in (1, 2, 4)
    assert all(ms > 0 for ms in r["ms"].values())
    assert min(r["ms"], key=r["ms"].get) == r["gates_per_workgroup"]
    assert hb.tuned_config(hb.SET_GOLDILOCKS) == (r["gates_per_workgroup"], True)
    assert f"1 {r['gates_per_workgroup']}" in cache.read_text().splitlines()
