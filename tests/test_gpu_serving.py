"""Serving loop: continuous batching of circuits from several clients.

Each circuit's result must be identical (word for word) whether it runs alone or mixed
with other clients' circuits, decrypt to the plaintext evaluation, and match the oracle
gate by gate."""
import sys

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def clients(hb, keyset):
    out = []
    for seed in (0, 1, 2):
        P, sk, bsk, ksk = keyset(hb.SET_GOLDILOCKS, seed)
        ev = hb.Evaluator(keys=(bsk, ksk), params=P)
        out.append((ev, hb.Encryptor(sk), hb.Decryptor(sk)))
    return out


def test_circuit_alone_equals_in_mix(hb, clients):
    from hasteboots_amd.serving import Server, random_nand_circuit
    rng = stream(50, "serve")
    jobs = []
    for i in range(6):
        ev, enc, dec = clients[i % 3]
        circ = random_nand_circuit(rng, int(rng.integers(1, 7)), int(rng.integers(2, 13)))
        bits = rng.integers(0, 2, size=circ.n_inputs)
        jobs.append((ev, dec, circ, bits, enc.encrypt_many(bits)))
    srv = Server(max_gates_per_launch=512)
    alone = []
    for ev, dec, circ, bits, cts in jobs:
        out, _ = srv.wait(srv.submit(ev, circ, cts))
        alone.append(out)
    say("alone done")
    ids = [srv.submit(ev, circ, cts) for ev, dec, circ, bits, cts in jobs]
    mixed = [srv.wait(r)[0] for r in ids]
    st = srv.stats()
    srv.close()
    say(f"stats {st}")
    for (ev, dec, circ, bits, cts), a, m in zip(jobs, alone, mixed):
        assert np.array_equal(a, m)
        expect = circ.eval_plain(bits)
        got = [dec.decrypt(hb.LWECiphertext(m[k])) for k in range(m.shape[0])]
        assert list(expect) == got
    assert st["completed"] == 12 and st["gates_done"] == 2 * sum(j[2].n_gates for j in jobs)


def test_admission_by_hbm_budget(hb, clients):
    """A budget that fits one circuit at a time: requests queue, then run in order, with the
    same results as with a large budget."""
    from hasteboots_amd.serving import Server, random_nand_circuit
    rng = stream(51, "budget")
    ev, enc, dec = clients[0]
    circs = [random_nand_circuit(rng, 3, 8) for _ in range(4)]
    bits = [rng.integers(0, 2, size=8) for _ in circs]
    cts = [enc.encrypt_many(b) for b in bits]
    n1 = ev.params.n + 1
    need = (8 + 24) * n1 * 8 + 256 + 3 * 24 * 4
    small = Server(max_gates_per_launch=256, hbm_budget=need + 512)
    big = Server(max_gates_per_launch=256, hbm_budget=64 << 20)
    ids = [small.submit(ev, c, x) for c, x in zip(circs, cts)]
    st_mid = small.stats()
    outs_small = [small.wait(r)[0] for r in ids]
    outs_big = [big.wait(big.submit(ev, c, x))[0] for c, x in zip(circs, cts)]
    st = small.stats()
    small.close()
    big.close()
    assert st_mid["queued"] + st_mid["active"] + st_mid["completed"] == 4
    assert st["completed"] == 4 and st["admitted"] == 4
    for a, b, c, bb in zip(outs_small, outs_big, circs, bits):
        assert np.array_equal(a, b)
        assert [dec.decrypt(hb.LWECiphertext(a[k])) for k in range(a.shape[0])] == list(c.eval_plain(bb))


def test_oversized_circuit_is_rejected(hb, clients):
    from hasteboots_amd._lib import HbError
    from hasteboots_amd.serving import Server, random_nand_circuit
    ev, enc, dec = clients[0]
    srv = Server(max_gates_per_launch=64, hbm_budget=1 << 20)
    circ = random_nand_circuit(stream(52, "big"), 40, 40)
    with pytest.raises(HbError) as e:
        srv.submit(ev, circ, enc.encrypt_many(np.zeros(40, np.int64)))
    assert e.value.code == -4
    srv.close()


@pytest.mark.parametrize("set_id", [0, 1], ids=["babybear", "goldilocks"])
def test_traced_request_traces_match_oracle(hb, orc, keyset, set_id):
    """A traced circuit served next to an untraced one: every gate's emitted core trace,
    expanded, equals the oracle's traced records for that gate (rounds 0, 363, 727 and the
    tail), and the outputs equal the oracle's traced evaluation."""
    from hasteboots_amd.serving import Server, random_nand_circuit
    say(f"traced serving, set {set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    enc = hb.Encryptor(sk)
    rng = stream(54 + set_id, "traced")
    circ = random_nand_circuit(rng, 2, 3)
    bits = rng.integers(0, 2, size=3)
    cts = enc.encrypt_many(bits)
    other = random_nand_circuit(rng, 4, 6)
    srv = Server(max_gates_per_launch=64)
    r_other = srv.submit(ev, other, enc.encrypt_many(rng.integers(0, 2, size=6)))
    r = srv.submit(ev, circ, cts, traced=True)
    out, _, core = srv.wait(r)
    srv.wait(r_other)
    srv.close()
    ok = orc.OracleKeys(set_id, bsk, ksk)
    lut = orc.build_lut(orc.params(set_id), 0)
    rounds = [0, 363, 727]
    wires = list(cts)
    for g in range(circ.n_gates):
        a, b = wires[circ.in0[g]], wires[circ.in1[g]]
        c = ((a.astype(object) + b.astype(object)) % P.q).astype(np.uint64)
        ref_out, ref_rec, ref_tail = ok.bootstrap_traced(c, lut, rounds)
        rec, tail = ev.expand_trace("nand", a, b, core[g], rounds)
        assert np.array_equal(rec, ref_rec), g
        assert np.array_equal(tail, ref_tail), g
        wires.append(ref_out)
    assert np.array_equal(out, np.stack([wires[o] for o in circ.outputs]))
    dec = hb.Decryptor(sk)
    assert [dec.decrypt(hb.LWECiphertext(out[k])) for k in range(3)] == list(circ.eval_plain(bits))


def test_server_matches_oracle_gate_by_gate(hb, orc, keyset):
    from hasteboots_amd.serving import Server, random_nand_circuit
    P, sk, bsk, ksk = keyset(hb.SET_BABYBEAR)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    rng = stream(53, "oracle")
    circ = random_nand_circuit(rng, 3, 4)
    bits = rng.integers(0, 2, size=4)
    cts = hb.Encryptor(sk).encrypt_many(bits)
    srv = Server(max_gates_per_launch=64)
    out, _ = srv.wait(srv.submit(ev, circ, cts))
    srv.close()
    ok = orc.OracleKeys(hb.SET_BABYBEAR, bsk, ksk)
    wires = list(cts)
    for g in range(circ.n_gates):
        r = ok.gate_batch(0, wires[circ.in0[g]][None, :], wires[circ.in1[g]][None, :])[0]
        wires.append(r)
    ref = np.stack([wires[o] for o in circ.outputs])
    assert np.array_equal(out, ref)


def test_config_size_circuits_alone_equal_mixed(hb, orc, clients):
    """Circuits drawn as the serving workload draws them (depth uniform in 1-64, width in
    2-64, NAND only, Goldilocks keys of three clients): each circuit's outputs are the same
    words alone and submitted together, decrypt to the plaintext evaluation, and one
    moderate circuit matches the oracle gate by gate."""
    from hasteboots_amd.serving import Server, random_nand_circuit
    rng = stream(55, "config-size")
    jobs = []
    for i in range(3):
        ev, enc, dec = clients[i]
        circ = random_nand_circuit(rng, int(rng.integers(1, 65)), int(rng.integers(2, 65)))
        bits = rng.integers(0, 2, size=circ.n_inputs)
        jobs.append((ev, dec, circ, bits, enc.encrypt_many(bits)))
    say("config-size circuits: " + ", ".join(f"{j[2].depth()}x{j[2].n_inputs}" for j in jobs))
    srv = Server(max_gates_per_launch=256)
    alone = [srv.wait(srv.submit(ev, circ, cts))[0] for ev, dec, circ, bits, cts in jobs]
    ids = [srv.submit(ev, circ, cts) for ev, dec, circ, bits, cts in jobs]
    mixed = [srv.wait(r)[0] for r in ids]
    srv.close()
    for (ev, dec, circ, bits, cts), a, m in zip(jobs, alone, mixed):
        assert np.array_equal(a, m)
        got = [dec.decrypt(hb.LWECiphertext(m[k])) for k in range(m.shape[0])]
        assert got == list(circ.eval_plain(bits))
    # one 6-layer, 16-wide circuit gate by gate against the oracle (a layer per oracle call)
    ev, enc, dec = clients[1]
    circ = random_nand_circuit(rng, 6, 16)
    bits = rng.integers(0, 2, size=16)
    cts = enc.encrypt_many(bits)
    srv = Server(max_gates_per_launch=256)
    out, _ = srv.wait(srv.submit(ev, circ, cts))
    srv.close()
    bsk, ksk = ev.export_keys()
    ok = orc.OracleKeys(hb.SET_GOLDILOCKS, bsk, ksk)
    wires = np.concatenate([cts, np.zeros((circ.n_gates, cts.shape[1]), np.uint64)])
    for layer in range(6):
        g = np.arange(16 * layer, 16 * (layer + 1))
        wires[16 + g] = ok.gate_batch(0, wires[circ.in0[g]], wires[circ.in1[g]], threads=16)
    assert np.array_equal(out, wires[circ.outputs])


def test_throughput_mode_equals_latency_mode(hb, keyset, clients):
    """With a backlog of ready gates above two batched workgroups per CU, the server runs
    launches in the batched (4 gates per workgroup, 2 waves per SIMD) shape over mixed
    parameter sets and clients; every circuit's outputs, and a traced circuit's core
    traces, are the same words as when the circuit runs alone in latency mode."""
    from hasteboots_amd.serving import Server, random_nand_circuit
    P0, sk0, bsk0, ksk0 = keyset(hb.SET_BABYBEAR)
    bb = (hb.Evaluator(keys=(bsk0, ksk0), params=P0), hb.Encryptor(sk0), hb.Decryptor(sk0))
    pool = list(clients) + [bb]
    rng = stream(57, "throughput")
    jobs = []
    for i in range(36):
        ev, enc, dec = pool[i % len(pool)]
        circ = random_nand_circuit(rng, 2, 64)
        bits = rng.integers(0, 2, size=64)
        jobs.append((ev, dec, circ, bits, enc.encrypt_many(bits)))
    ev_t, enc_t, dec_t = clients[0]
    circ_t = random_nand_circuit(rng, 2, 8)
    bits_t = rng.integers(0, 2, size=8)
    cts_t = enc_t.encrypt_many(bits_t)
    srv = Server(max_gates_per_launch=4096)
    ids = [srv.submit(ev, circ, cts) for ev, dec, circ, bits, cts in jobs]
    rid_t = srv.submit(ev_t, circ_t, cts_t, traced=True)
    mixed = [srv.wait(r)[0] for r in ids]
    out_t, _, core_t = srv.wait(rid_t)
    st = srv.stats()
    srv.close()
    say(f"throughput stats {st}")
    assert st["batched_launches"] >= 1 and st["max_launch_gates"] >= 2048
    lat = Server(max_gates_per_launch=256)
    alone = [lat.wait(lat.submit(ev, circ, cts))[0] for ev, dec, circ, bits, cts in jobs]
    a_out, _, a_core = lat.wait(lat.submit(ev_t, circ_t, cts_t, traced=True))
    st_lat = lat.stats()
    lat.close()
    assert st_lat["batched_launches"] == 0 and st_lat["max_launch_gates"] <= 256
    for k, ((ev, dec, circ, bits, cts), a, m) in enumerate(zip(jobs, alone, mixed)):
        assert np.array_equal(a, m), k
        got = [dec.decrypt(hb.LWECiphertext(m[j])) for j in range(m.shape[0])]
        assert got == list(circ.eval_plain(bits)), k
    assert np.array_equal(out_t, a_out)
    assert np.array_equal(core_t, a_core)


def test_small_launches_pair_mode_equals_latency_mode(hb, keyset, clients):
    """Small launches (at most one gate per two CUs, one slot busy at a time) run in pair
    mode over BabyBear and Goldilocks circuits together; every circuit's outputs are the same
    words as with pair mode off (latency mode) and decrypt to the plaintext evaluation."""
    from hasteboots_amd.serving import Server, random_nand_circuit
    P0, sk0, bsk0, ksk0 = keyset(hb.SET_BABYBEAR)
    bb = (hb.Evaluator(keys=(bsk0, ksk0), params=P0), hb.Encryptor(sk0), hb.Decryptor(sk0))
    pool = [bb, clients[0], bb, clients[1]]
    rng = stream(58, "pair-serve")
    jobs = []
    for ev, enc, dec in pool:
        circ = random_nand_circuit(rng, 4, 12)
        bits = rng.integers(0, 2, size=12)
        jobs.append((ev, dec, circ, bits, enc.encrypt_many(bits)))

    def run():
        srv = Server(max_gates_per_launch=64)
        ids = [srv.submit(ev, circ, cts) for ev, dec, circ, bits, cts in jobs]
        outs = [srv.wait(r)[0] for r in ids]
        st = srv.stats()
        srv.close()
        return outs, st

    assert hb.pair_mode()
    paired, st = run()
    say(f"pair stats {st}")
    assert st["pair_launches"] >= 1
    hb.set_pair_mode(False)
    try:
        plain, st_off = run()
    finally:
        hb.set_pair_mode(True)
    assert st_off["pair_launches"] == 0
    for k, ((ev, dec, circ, bits, cts), a, m) in enumerate(zip(jobs, plain, paired)):
        assert np.array_equal(a, m), k
        got = [dec.decrypt(hb.LWECiphertext(m[j])) for j in range(m.shape[0])]
        assert got == list(circ.eval_plain(bits)), k


def test_pair_wait_timeout_stops_the_server_with_an_error(hb, clients):
    """A serving launch in pair mode whose waits run out (forced from the host with 0 polls
    per wait) makes the request's wait raise instead of returning the launch's wrong words;
    a new server with the default limit runs the same circuit to the plaintext result."""
    from hasteboots_amd._lib import HbError
    from hasteboots_amd.serving import Server, random_nand_circuit
    ev, enc, dec = clients[0]
    rng = stream(59, "pair-timeout")
    circ = random_nand_circuit(rng, 2, 6)
    bits = rng.integers(0, 2, size=6)
    cts = enc.encrypt_many(bits)
    hb.set_pair_wait_limit(0)
    try:
        srv = Server(max_gates_per_launch=64)
        rid = srv.submit(ev, circ, cts)
        with pytest.raises(HbError) as ei:
            srv.wait(rid)
        srv.close()
    finally:
        hb.set_pair_wait_limit()
    assert "timed out" in str(ei.value)
    srv = Server(max_gates_per_launch=64)
    out, _ = srv.wait(srv.submit(ev, circ, cts))
    st = srv.stats()
    srv.close()
    assert st["pair_launches"] >= 1
    got = [dec.decrypt(hb.LWECiphertext(out[k])) for k in range(out.shape[0])]
    assert got == list(circ.eval_plain(bits))
