"""Traced bootstrapping on the GPU: the core trace emitted by the kernel, expanded by the
library, equals the oracle's trace records (restating the reference's trace crate) word for
word; outputs follow the traced path's own semantics."""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import stream, uniform_field

pytestmark = pytest.mark.gpu
GP = 2**64 - 2**32 + 1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("set_id,rounds",
                         [(0, [0, 1, 2, 300, 726, 727]), (1, [0, 1, 727]), (2, [0, 363, 727])],
                         ids=["babybear", "goldilocks", "zama"])
def test_expanded_trace_matches_oracle(hb, orc, keyset, set_id, rounds):
    say(f"trace set={set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    enc = hb.Encryptor(sk)
    rng = stream(60 + set_id, "trace")
    b = rng.integers(0, 2, size=(2, 2))
    c0 = enc.encrypt_many(b[0])
    c1 = enc.encrypt_many(b[1])
    outs, core = ev.gate_batch_traced("nand", c0, c1)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    lut = orc.build_lut(orc.params(set_id), 0)
    plain = ev.nand_batch(c0, c1)
    for g in range(2):
        c = ((c0[g].astype(object) + c1[g].astype(object)) % P.q).astype(np.uint64)
        ref_out, ref_rec, ref_tail = ok.bootstrap_traced(c, lut, rounds)
        assert np.array_equal(outs[g], ref_out)
        # real keys: no raw words above P reach a negation, both paths agree
        assert np.array_equal(outs[g], plain[g])
        rec, tail = ev.expand_trace("nand", c0[g], c1[g], core[g], rounds)
        assert np.array_equal(rec, ref_rec), g
        assert np.array_equal(tail, ref_tail), g
        m = hb.Decryptor(sk).decrypt(hb.LWECiphertext(outs[g]))
        assert m == 1 - (int(b[0, g]) & int(b[1, g]))


def test_traced_semantics_on_raw_word_keys(hb, orc):
    """With keys that push the accumulator into raw words above P, the traced path (exact
    monomial products) and the untraced path (raw-word negation) give different outputs;
    the GPU follows each path's oracle exactly."""
    say("trace adversarial")
    P = hb.Parameters.of(hb.SET_GOLDILOCKS)
    n, N, L = P.n, P.N, P.br_len
    bsk = np.zeros((n, 2, L, 2, N), np.uint64)
    bsk[:, 1, 0, 1, :] = GP - 1
    bsk = bsk.reshape(-1)
    ksk = uniform_field(stream(6, "ksk"), GP, P.ks_len * 2 * N)
    rng = stream(6, "adv")
    lut = rng.integers(1, 1000, size=N, dtype=np.uint64)  # small words: sums land in [P, 2^64)
    cs = uniform_field(rng, GP, (3, n + 1))
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    ok = orc.OracleKeys(hb.SET_GOLDILOCKS, bsk, ksk)
    t_out, core = ev.bootstrap_traced(cs, lut)
    differ = 0
    for g in range(3):
        ref_t, _, _ = ok.bootstrap_traced(cs[g], lut, [])
        ref_u = ok.bootstrap(cs[g], lut)
        u_out = ev.bootstrap(hb.LWECiphertext(cs[g]), lut).words
        assert np.array_equal(t_out[g], ref_t)
        assert np.array_equal(u_out, ref_u)
        differ += int(not np.array_equal(ref_t, ref_u))
        rec, tail = ev.expand_trace("bootstrap", cs[g], None, core[g], [0, 5, 727])
        _, ref_rec, ref_tail = ok.bootstrap_traced(cs[g], lut, [0, 5, 727])
        assert np.array_equal(rec, ref_rec)
        assert np.array_equal(tail, ref_tail)
    assert differ == 3


@pytest.mark.parametrize("set_id", [1, 0], ids=["goldilocks", "babybear"])
def test_finalized_containers_match_oracle(hb, orc, keyset, set_id):
    """Every finalized PBSTrace container (AccTrace, SumHadamardTrace with all 2*br_len
    hadamards, permutation_info, ModulusSwitchingTrace), built by hb_trace_part from the
    GPU's core trace, equals the oracle's restatement of the trace crate's append + finalize
    word for word (trace/src/acc_trace.rs:157-202, modulus_switching_trace.rs:112-133)."""
    say(f"finalized containers set={set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    enc = hb.Encryptor(sk)
    c0, c1 = enc.encrypt_many([1]), enc.encrypt_many([0])
    outs, core = ev.gate_batch_traced("nand", c0, c1)
    c = ((c0[0].astype(object) + c1[0].astype(object)) % P.q).astype(np.uint64)
    lut = orc.build_lut(orc.params(set_id), 0)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    jobs = [(k, 0) for k, name in enumerate(hb.TRACE_PARTS) if name != "hadamard"]
    jobs += [(hb.TRACE_PARTS.index("hadamard"), j) for j in range(2 * P.br_len)]

    def run(job):
        part, j = job
        got = ev.trace_part("nand", c0[0], c1[0], core[0], part, j).reshape(-1)
        ref = ok.pbs_trace_part(c, lut, part, j)
        return job, bool(np.array_equal(got, ref))

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as pool:
        res = dict(pool.map(run, jobs))
    bad = [(hb.TRACE_PARTS[p], j) for (p, j), good in res.items() if not good]
    assert not bad, bad
    say(f"{len(res)} containers equal")


@pytest.mark.parametrize("set_id", [1, 2, 0], ids=["goldilocks", "zama", "babybear"])
def test_traced_latency_mode_equals_one_wave(hb, keyset, set_id):
    """Traced gates in latency mode (the 8-wave kernel writing the history from registers)
    and with latency mode off (one wavefront per gate, the history as its accumulator) give
    the same outputs and the same core trace, word for word."""
    say(f"traced latency mode against one wave, set={set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    enc = hb.Encryptor(sk)
    rng = stream(70 + set_id, "trace-modes")
    b = rng.integers(0, 2, size=(2, 5))
    c0, c1 = enc.encrypt_many(b[0]), enc.encrypt_many(b[1])
    modes = hb.latency_modes()
    try:
        split_out, split_core = ev.gate_batch_traced("nand", c0, c1)
        hb.set_latency_mode(0)
        one_out, one_core = ev.gate_batch_traced("nand", c0, c1)
    finally:
        hb.restore_latency_modes(modes)
    assert np.array_equal(split_out, one_out)
    assert np.array_equal(split_core, one_core)
