"""GPU parity: the HIP kernels against the CPU oracle, bit for bit, through the C ABI.

Arithmetic is exact, so every comparison is equality of canonical words (tolerance 0).
"""
import sys

import numpy as np
import pytest

from conftest import stream, uniform_field

pytestmark = pytest.mark.gpu

GP = 2**64 - 2**32 + 1
BP = 2**31 - 2**27 + 1
NAND, AND, OR, NOR, XOR, XNOR, MAJ = range(7)
TRUTH = {NAND: lambda a, b, c: 1 - (a & b), AND: lambda a, b, c: a & b,
         OR: lambda a, b, c: a | b, NOR: lambda a, b, c: 1 - (a | b),
         XOR: lambda a, b, c: a ^ b, XNOR: lambda a, b, c: 1 - (a ^ b),
         MAJ: lambda a, b, c: int(a + b + c >= 2)}
NAMES = {NAND: "nand", AND: "and", OR: "or", NOR: "nor", XOR: "xor", XNOR: "xnor", MAJ: "majority"}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def dev_buffer(hb_lib, arr):
    import ctypes as C
    from hasteboots_amd._lib import check
    p = C.c_void_p()
    check(hb_lib.hb_dev_alloc(arr.nbytes, C.byref(p)))
    check(hb_lib.hb_memcpy_h2d(p, arr.ctypes.data, arr.nbytes))
    return p


def dev_read(hb_lib, p, like):
    from hasteboots_amd._lib import check
    out = np.zeros_like(like)
    check(hb_lib.hb_memcpy_d2h(out.ctypes.data, p, out.nbytes))
    return out


@pytest.fixture(scope="module")
def hlib(hb):
    from hasteboots_amd._lib import lib
    assert hb.device_count() > 0, "no GPU visible"
    return lib()


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)], ids=["goldilocks", "babybear"])
@pytest.mark.parametrize("log_n", [10, 11])
def test_ntt_matches_oracle(hlib, orc, gold, q, log_n):
    from hasteboots_amd._lib import check
    say(f"ntt gold={gold} log_n={log_n}")
    n = 1 << log_n
    x = uniform_field(stream(21, f"ntt{log_n}"), q, (64, n))
    d = dev_buffer(hlib, x)
    check(hlib.hb_ntt_dev(int(gold), 0, log_n, d, 64, None))
    check(hlib.hb_sync())
    y = dev_read(hlib, d, x)
    assert np.array_equal(y, orc.ntt(gold, x))
    check(hlib.hb_ntt_dev(int(gold), 1, log_n, d, 64, None))
    check(hlib.hb_sync())
    z = dev_read(hlib, d, x)
    assert np.array_equal(z, x)
    # edge words: all zeros, all q-1, a single spike
    e = np.zeros((4, n), np.uint64)
    e[1, :] = q - 1
    e[2, 0] = 1
    e[3, n - 1] = q - 1
    check(hlib.hb_memcpy_h2d(d, e.ctypes.data, e.nbytes))
    check(hlib.hb_ntt_dev(int(gold), 0, log_n, d, 4, None))
    check(hlib.hb_sync())
    assert np.array_equal(dev_read(hlib, d, e), orc.ntt(gold, e))
    check(hlib.hb_dev_free(d))


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)], ids=["goldilocks", "babybear"])
@pytest.mark.parametrize("log_n", [10, 11])
@pytest.mark.parametrize("count", [2049, 4099])
def test_ntt_batches_beyond_resident_waves(hlib, orc, gold, q, log_n, count):
    """Batches larger than the resident waves (two 4-wave workgroups per CU): waves take
    a second and third polynomial, which N = 1024 stages through LDS while transforming
    the one before; every polynomial equals the oracle's transform and round-trips."""
    from hasteboots_amd._lib import check
    say(f"ntt batch gold={gold} log_n={log_n} count={count}")
    n = 1 << log_n
    x = uniform_field(stream(23, f"ntt-batch{log_n}-{count}"), q, (count, n))
    d = dev_buffer(hlib, x)
    check(hlib.hb_ntt_dev(int(gold), 0, log_n, d, count, None))
    check(hlib.hb_sync())
    y = dev_read(hlib, d, x)
    assert np.array_equal(y, orc.ntt(gold, x))
    check(hlib.hb_ntt_dev(int(gold), 1, log_n, d, count, None))
    check(hlib.hb_sync())
    assert np.array_equal(dev_read(hlib, d, x), x)
    check(hlib.hb_dev_free(d))


@pytest.mark.parametrize("gold,q,bits,length", [(True, GP, 5, 13), (True, GP, 8, 8),
                                                (False, BP, 8, 4), (False, BP, 5, 7)])
def test_external_product_matches_oracle(hlib, orc, gold, q, bits, length):
    from hasteboots_amd._lib import check
    say(f"extprod gold={gold} bits={bits} len={length}")
    n = 1024
    rng = stream(22, f"ep{bits}")
    rgsw = uniform_field(rng, q, 2 * length * 2 * n)
    rl = uniform_field(rng, q, (32, 2 * n))
    d_k = dev_buffer(hlib, rgsw)
    d_r = dev_buffer(hlib, rl)
    check(hlib.hb_external_product_dev(int(gold), 10, bits, length, d_k, d_r, 32, None))
    check(hlib.hb_sync())
    got = dev_read(hlib, d_r, rl)
    for i in range(0, 32, 7):
        ea, eb = orc.external_product(gold, bits, length, rgsw, rl[i, :n], rl[i, n:])
        assert np.array_equal(got[i, :n], ea), i
        assert np.array_equal(got[i, n:], eb), i
    check(hlib.hb_dev_free(d_k))
    check(hlib.hb_dev_free(d_r))


def _encrypt_bits(hb, sk, bits):
    enc = hb.Encryptor(sk)
    return enc.encrypt_many(bits)


def test_babybear_single_nand_seed0(hb, orc, keyset):
    """One NAND with BABYBEAR_BINARY_128_BITS_PARAMETERS, keys and bits from seed 0."""
    say("babybear single nand")
    P, sk, bsk, ksk = keyset(hb.SET_BABYBEAR)
    bits = stream(0, "bits").integers(0, 2, size=2)
    c = _encrypt_bits(hb, sk, bits)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    out = ev.nand(hb.LWECiphertext(c[0]), hb.LWECiphertext(c[1]))
    ok = orc.OracleKeys(hb.SET_BABYBEAR, bsk, ksk)
    ref = ok.gate_batch(NAND, c[:1], c[1:2])[0]
    assert np.array_equal(out.words, ref)
    m, noise = hb.Decryptor(sk).decrypt_with_noise(out)
    assert m == 1 - (int(bits[0]) & int(bits[1]))
    assert noise < P.q / 16


@pytest.mark.parametrize("set_id,count", [(0, 96), (1, 48), (2, 16)],
                         ids=["babybear", "goldilocks", "zama"])
def test_nand_batch_matches_oracle(hb, orc, keyset, set_id, count):
    say(f"nand batch set={set_id} count={count}")
    P, sk, bsk, ksk = keyset(set_id)
    rng = stream(set_id, "batchbits")
    b0 = rng.integers(0, 2, size=count)
    b1 = rng.integers(0, 2, size=count)
    c0 = _encrypt_bits(hb, sk, b0)
    c1 = _encrypt_bits(hb, sk, b1)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    got = ev.nand_batch(c0, c1)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    ref = ok.gate_batch(NAND, c0, c1, threads=16)
    assert np.array_equal(got, ref)
    dec = hb.Decryptor(sk)
    for i in range(count):
        m, noise = dec.decrypt_with_noise(got[i])
        assert m == 1 - (int(b0[i]) & int(b1[i]))
        assert noise < P.q / 16
    # a gate's result is the same alone as inside the batch
    alone = ev.nand(hb.LWECiphertext(c0[3]), hb.LWECiphertext(c1[3]))
    assert np.array_equal(alone.words, got[3])


@pytest.mark.parametrize("gate", [AND, OR, NOR, XOR, XNOR, MAJ], ids=lambda g: NAMES[g])
def test_all_gates_babybear(hb, orc, keyset, gate):
    say(f"gate {NAMES[gate]}")
    P, sk, bsk, ksk = keyset(hb.SET_BABYBEAR)
    combos = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    a = np.array([x[0] for x in combos]); b = np.array([x[1] for x in combos])
    c = np.array([x[2] for x in combos])
    ca, cb, cc = (_encrypt_bits(hb, sk, v) for v in (a, b, c))
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    got = ev.gate_batch(NAMES[gate], ca, cb, cc if gate == MAJ else None)
    ok = orc.OracleKeys(hb.SET_BABYBEAR, bsk, ksk)
    ref = ok.gate_batch(gate, ca, cb, cc if gate == MAJ else None, threads=8)
    assert np.array_equal(got, ref)
    dec = hb.Decryptor(sk)
    for i, (x, y, z) in enumerate(combos):
        assert dec.decrypt(hb.LWECiphertext(got[i])) == TRUTH[gate](x, y, z)


@pytest.mark.parametrize("set_id", [0, 1])
def test_bootstrap_custom_lut(hb, orc, keyset, set_id):
    say(f"custom lut set={set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    rng = stream(5, "lut")
    lut = uniform_field(rng, P.q, P.N)
    c = _encrypt_bits(hb, sk, [1])[0]
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    got = ev.bootstrap(hb.LWECiphertext(c), lut)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    assert np.array_equal(got.words, ok.bootstrap(c, lut))


@pytest.mark.parametrize("latency", [True, False], ids=["latency-mode", "wave-per-gate"])
def test_goldilocks_raw_word_accumulator(hb, orc, latency):
    """Adversarial keys that drive the accumulator into raw words in [P, 2^64), so the
    reference's wrapping negation (ops.rs:69-71) of such words is exercised every round.

    Every RGSW is zero except c_m[0].b = Q-1 in all slots (the constant polynomial -1), so
    each round adds -digit_0((X^a - 1) ACC.b) to ACC.b; with a small-valued LUT the sums
    land in [P, 2^64). GPU and oracle must agree word for word, both through the
    latency-mode kernel and through the one-wavefront-per-gate kernel with byte digits."""
    say(f"raw-word accumulator, latency mode {latency}")
    old = hb.latency_modes()
    hb.restore_latency_modes(old if latency else (0, 0, 0))
    try:
        _raw_word_accumulator(hb, orc)
    finally:
        hb.restore_latency_modes(old)


def _raw_word_accumulator(hb, orc):
    set_id = hb.SET_GOLDILOCKS
    P = hb.Parameters.of(set_id)
    n, N, L = P.n, P.N, P.br_len
    bsk = np.zeros((n, 2, L, 2, N), np.uint64)
    bsk[:, 1, 0, 1, :] = GP - 1
    ksk = uniform_field(stream(6, "ksk"), GP, P.ks_len * 2 * N)
    bsk = bsk.reshape(-1)
    rng = stream(6, "adv")
    lut = rng.integers(1, 1000, size=N, dtype=np.uint64)
    c = uniform_field(rng, GP, (8, n + 1))
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    for i in range(8):
        orc.raw_neg_events(reset=True)
        ref = ok.bootstrap(c[i], lut)
        # the oracle negated raw words above P while computing this output
        assert orc.raw_neg_events() > 0, i
        got = ev.bootstrap(hb.LWECiphertext(c[i]), lut)
        assert np.array_equal(got.words, ref), i


@pytest.mark.parametrize("set_id", [0, 1])
def test_modulus_switch_rounding_boundaries(hb, orc, keyset, set_id):
    """Inputs whose switched value sits next to a .5 boundary of round(v * 2N / Q)."""
    say(f"mod switch boundaries set={set_id}")
    P, sk, bsk, ksk = keyset(set_id)
    q, n, N = P.q, P.n, P.N
    rng = stream(7, "ms")
    base = uniform_field(rng, q, (4, n + 1))
    for row in range(4):
        for k in range(0, n + 1, 3):
            j = int(rng.integers(0, 2 * N))
            center = ((2 * j + 1) * q) // (4 * N)
            base[row, k] = (center + int(rng.integers(-2, 3))) % q
    base[:, -1] = [0, q - 1, q // 2, 1]
    zero = np.zeros_like(base)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    got = ev.gate_batch("nand", base, zero)
    ref = ok.gate_batch(NAND, base, zero, threads=4)
    assert np.array_equal(got, ref)


def test_zama_batch_through_tuned_variant(hb, orc, keyset):
    """Zama-Goldilocks (13 five-bit digits, stored as byte digits) through the (4, 2)
    batched variant, latency mode off: more gates than SIMDs, the last workgroup partly
    empty."""
    P, sk, bsk, ksk = keyset(2)
    count = 1024 + 9
    say(f"zama batch {count} through (4, 2)")
    rng = stream(2, "zama-variant")
    b0, b1 = rng.integers(0, 2, size=(2, count))
    enc = hb.Encryptor(sk)
    c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    old, modes = hb.tuned_config(hb.SET_ZAMA), hb.latency_modes()
    hb.set_tuned_config(hb.SET_ZAMA, 4, 2)
    hb.set_latency_mode(0)
    try:
        out = ev.gate_batch("nand", c0, c1)
    finally:
        hb.set_tuned_config(hb.SET_ZAMA, old[0], old[1])
        hb.restore_latency_modes(modes)
    sample = np.r_[0:2, count - 2:count]
    ref = orc.OracleKeys(2, bsk, ksk).gate_batch(0, c0[sample], c1[sample], threads=4)
    assert np.array_equal(out[sample], ref)
    dec = hb.Decryptor(sk)
    assert all(dec.decrypt(hb.LWECiphertext(out[i])) == 1 - (int(b0[i]) & int(b1[i]))
               for i in range(count))
