"""Evaluation-key generation and the serialized formats, on the CPU.

Key generation (BinaryBlindRotationKey::generate, fhe_core/src/blind_rotation/rlwe/binary.rs:
207-238; KeySwitchingRLWEKey::generate, fhe_core/src/key_switch/rlwe.rs:37-112) is pinned by
what the keys must satisfy: decrypting a BSK row under the ring key gives its Gaussian error
(plus B^j at X^0 for the c_m rows of key bit 1), with the parameter set's sigma, the 6-sigma
cut and the normal law's mass within one sigma; the a-coefficients are uniform; the KSK rows
decrypt under NTT(s_lwe ‖ 0) to their error plus B_ks^i s_ring. The formats round-trip and
refuse corrupted, truncated, foreign or non-canonical data.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import stream


def centered(v, q):
    v = np.asarray(v, dtype=object)
    return np.where(v > q // 2, v - q, v).astype(np.float64)


def mulq(a, b, q):
    return (a.astype(object) * b.astype(object)) % q


@pytest.fixture(scope="module", params=[1, 0], ids=["goldilocks", "babybear"])
def keys(request, hb, keyset):
    P, sk, bsk, ksk = keyset(request.param)
    return P, sk, bsk.reshape(P.n, 2, P.br_len, 2, P.N), ksk.reshape(P.ks_len, 2, P.N)


def test_bsk_rows_decrypt_to_gaussian_errors(hb, orc, keys):
    P, sk, bsk, ksk = keys
    q, N, gl = P.q, P.N, P.is_goldilocks
    s_lwe = sk.lwe_secret_key()
    s_ntt = orc.ntt(gl, sk.ring_secret_key())
    errs, checked_value = [], 0
    for i in range(0, P.n, 23):
        for c in range(2):
            for j in range(0, P.br_len, 3):
                a, b = bsk[i, c, j, 0].copy(), bsk[i, c, j, 1]
                Bj = 1 << (P.br_bits * j)
                if c == 0 and s_lwe[i]:  # c_neg_s_m: a carries + B^j after b was formed
                    a = ((a.astype(object) - Bj) % q).astype(np.uint64)
                e_ntt = ((b.astype(object) - mulq(a, s_ntt, q)) % q).astype(np.uint64)
                e = centered(orc.ntt(gl, e_ntt, inverse=True), q)
                if c == 1 and s_lwe[i]:  # c_m of key bit 1: + B^j at X^0
                    e[0] -= Bj
                    checked_value += 1
                errs.append(e)
    e = np.concatenate(errs)
    sigma = {0: 41.9, 1: 2.3 * 2**37, 2: 2.3 * 2**37}[P.set_id]
    assert checked_value > 0
    assert np.all(np.abs(e) < 6 * sigma)
    assert np.allclose(e, np.round(e))
    n = e.size
    assert abs(e.mean()) < 5 * sigma / math.sqrt(n)
    assert abs(e.std() / sigma - 1) < 0.02, e.std() / sigma
    within = np.mean(np.abs(e) <= sigma)
    assert abs(within - 0.6827) < 0.01, within
    # a-coefficients uniform on [0, q): 16 equal bins
    a = np.concatenate([bsk[i, 1, 0, 0] for i in range(0, P.n, 7)]).astype(np.float64)
    counts = np.bincount(np.minimum((a / q * 16).astype(int), 15), minlength=16)
    expect = a.size / 16
    chi2 = float(np.sum((counts - expect) ** 2 / expect))
    assert chi2 < 50, chi2  # 15 degrees of freedom


def test_ksk_rows_decrypt_to_gaussian_errors(hb, orc, keys):
    P, sk, bsk, ksk = keys
    q, N, gl = P.q, P.N, P.is_goldilocks
    s_lwe = np.zeros(N, np.uint64)
    s_lwe[:P.n] = sk.lwe_secret_key()
    lwe_ntt = orc.ntt(gl, s_lwe)
    ring_ntt = orc.ntt(gl, sk.ring_secret_key())
    errs = []
    for i in range(P.ks_len):
        a, b = ksk[i, 0], ksk[i, 1]
        Bi = pow(1 << P.ks_bits, i, q)
        rhs = (mulq(a, lwe_ntt, q) + (ring_ntt.astype(object) * Bi)) % q
        e_ntt = ((b.astype(object) - rhs) % q).astype(np.uint64)
        errs.append(centered(orc.ntt(gl, e_ntt, inverse=True), q))
    e = np.concatenate(errs)
    sigma = {0: 11000.0, 1: 2.9 * 2**45, 2: 2.9 * 2**45}[P.set_id]
    assert np.all(np.abs(e) < 6 * sigma)
    assert abs(e.std() / sigma - 1) < 0.05, e.std() / sigma


def test_keys_depend_on_the_seed_only(hb):
    P = hb.Parameters.of(0)
    a = hb.KeyGen.generate_secret_key(P, seed=5).generate_keys_host()
    b = hb.KeyGen.generate_secret_key(P, seed=5).generate_keys_host()
    c = hb.KeyGen.generate_secret_key(P, seed=6).generate_keys_host()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_key_format_round_trip_and_refusals(hb, keyset):
    from hasteboots_amd._lib import HbError
    P, sk, bsk, ksk = keyset(hb.SET_BABYBEAR)
    data = hb.serialize_keys(P.set_id, bsk, ksk)
    assert data[:4] == b"HBTF" and len(data) == 72 + 8 * (bsk.size + ksk.size)
    s, b2, k2 = hb.deserialize_keys(data)
    assert s == P.set_id and np.array_equal(b2, bsk) and np.array_equal(k2, ksk)
    raw = bytearray(data)
    bad = bytearray(raw)
    bad[72 + 8 * 1000] ^= 1  # one payload bit
    with pytest.raises(HbError, match="checksum"):
        hb.deserialize_keys(bytes(bad))
    bad = bytearray(raw)
    bad[4] = 2  # version
    with pytest.raises(HbError, match="version"):
        hb.deserialize_keys(bytes(bad))
    bad = bytearray(raw)
    bad[12] ^= 1  # n
    with pytest.raises(HbError, match="parameters"):
        hb.deserialize_keys(bytes(bad))
    with pytest.raises(HbError, match="length"):
        hb.deserialize_keys(bytes(raw[:-8]))
    with pytest.raises(HbError, match="kind"):
        hb.deserialize_ciphertexts(data)
    big = bsk.copy()
    big[3] = P.q
    with pytest.raises(HbError, match="canonical"):
        hb.serialize_keys(P.set_id, big, ksk)


def test_ciphertext_format_round_trip_and_refusals(hb, keyset):
    from hasteboots_amd._lib import HbError
    for set_id in (0, 1, 2):
        P, sk, bsk, ksk = keyset(set_id)
        cts = hb.Encryptor(sk).encrypt_many(stream(9, "fmt").integers(0, 2, size=5))
        data = hb.serialize_ciphertexts(set_id, cts)
        s, back = hb.deserialize_ciphertexts(data)
        assert s == set_id and np.array_equal(back, cts)
        dec = hb.Decryptor(sk)
        assert [dec.decrypt(hb.LWECiphertext(c)) for c in back] == \
            [dec.decrypt(hb.LWECiphertext(c)) for c in cts]
    s, empty = hb.deserialize_ciphertexts(hb.serialize_ciphertexts(1, np.zeros((0, 729), np.uint64)))
    assert empty.shape == (0, 729)
    bad = bytearray(data)
    bad[-1] ^= 0x80
    with pytest.raises(HbError):
        hb.deserialize_ciphertexts(bytes(bad))
    from hasteboots_amd._lib import lib
    assert lib().hb_lwe_deserialize(bytes(data), len(data), None, None,
                                    (C.c_uint64 * 8)(), 8) == -4  # too little room
