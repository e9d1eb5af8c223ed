"""The C-ABI library loads and exports every function include/hb.h declares; host-side
entry points (parameters, key generation, encryption) behave like the reference's
(no GPU compute here)."""
import ctypes as C

import numpy as np
import pytest


def test_library_exports_every_declared_symbol(hb):
    from hasteboots_amd._lib import header_symbols, lib, _PROTOS
    l = lib()
    names = header_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(l, name), f"{name} declared in hb.h but not exported"
        assert name in _PROTOS, f"{name} has no ctypes prototype"


@pytest.mark.parametrize("set_id,q,br_len,ks_len,br_bits", [
    (0, 2**31 - 2**27 + 1, 4, 7, 8),
    (1, 2**64 - 2**32 + 1, 8, 13, 8),
    (2, 2**64 - 2**32 + 1, 13, 13, 5)])
def test_parameter_sets(hb, set_id, q, br_len, ks_len, br_bits):
    # tfhe/src/bfhe/parameters.rs:45-99; decompose_len = ceil(bits(Q) / basis bits)
    P = hb.Parameters.of(set_id)
    assert (P.n, P.N, P.t, P.q) == (728, 1024, 4, q)
    assert (P.br_len, P.ks_len, P.br_bits, P.ks_bits) == (br_len, ks_len, br_bits, 5)
    # Parameters::new validation: N a power of two, 2N | Q-1
    assert P.N & (P.N - 1) == 0 and (P.q - 1) % (2 * P.N) == 0
    # kernel shape assumptions: a quarter of a gadget's digits fits 32 bits (8-wave latency
    # mode), and Goldilocks with 8 digits of 8 bits takes the byte-digit path
    assert -(-P.br_len // 4) * P.br_bits <= 32
    assert (P.q > 2**32 and P.br_bits == 8 and P.br_len == 8) == (set_id == 1)


def test_bad_arguments_return_errors(hb):
    from hasteboots_amd._lib import lib, HbParams
    l = lib()
    assert l.hb_params(7, C.byref(HbParams())) == -1
    assert b"unknown" in l.hb_last_error()
    h = C.c_void_p()
    assert l.hb_sk_create(9, 0, C.byref(h)) == -1


def test_tuning_and_launch_arguments_are_checked_before_the_device(hb):
    """Bad parameter sets, kernel variants and handles are refused with HB_ERR_ARG and a
    message, without touching a GPU."""
    from hasteboots_amd._lib import lib
    l = lib()
    assert l.hb_set_tuned_config(3, 4, 2) == -1
    assert b"parameter set" in l.hb_last_error()
    for gpw, wps in ((3, 1), (4, 4), (1, 2), (8, 2), (0, 1)):
        assert l.hb_set_tuned_config(1, gpw, wps) == -1, (gpw, wps)
        assert b"no kernel variant" in l.hb_last_error()
    g, w = C.c_int(), C.c_int()
    assert l.hb_tuned_config(5, C.byref(g), C.byref(w), None) == -1
    assert l.hb_latency_mode(None) == -1
    assert l.hb_latency_mode_for(1, None) == -1
    assert l.hb_set_latency_mode_for(3, 10) == -1
    assert b"parameter set" in l.hb_last_error()
    assert l.hb_gate_batch_dev(None, 8, 4, None, None, None, None, None) == -1
    assert l.hb_stream_sync(None) == -1
    h = C.c_void_p()
    assert l.hb_server_create(0, 0, 0, C.byref(h)) == -1
    assert b"max_gates" in l.hb_last_error()
    assert l.hb_server_create(0, (1 << 20) + 1, 0, C.byref(h)) == -1
    assert l.hb_batch_graph_create(None, 0, 16, None, None, None, None, C.byref(h)) == -1
    assert b"null" in l.hb_last_error()


@pytest.mark.parametrize("set_id", [0, 1, 2])
def test_encrypt_decrypt_roundtrip(hb, set_id):
    P = hb.Parameters.of(set_id)
    sk = hb.KeyGen.generate_secret_key(P, seed=3)
    enc, dec = hb.Encryptor(sk), hb.Decryptor(sk)
    for m in (0, 1, 0, 1, 1):
        c = enc.encrypt(m)
        got, noise = dec.decrypt_with_noise(c)
        assert got == m
        assert noise < 6 * (11000 if set_id == 0 else 2.9 * 2**45) + 1
        assert (c.words < np.uint64(P.q)).all()


def test_keys_are_deterministic_in_the_seed(hb):
    P = hb.Parameters.of(0)
    a = hb.KeyGen.generate_secret_key(P, seed=42)
    b = hb.KeyGen.generate_secret_key(P, seed=42)
    c = hb.KeyGen.generate_secret_key(P, seed=43)
    assert np.array_equal(a.lwe_secret_key(), b.lwe_secret_key())
    assert np.array_equal(a.ring_secret_key(), b.ring_secret_key())
    assert not np.array_equal(a.lwe_secret_key(), c.lwe_secret_key())
    s = a.lwe_secret_key()
    assert set(np.unique(s).tolist()) <= {0, 1}
    r = a.ring_secret_key()
    assert set(np.unique(r).tolist()) <= {0, 1, P.q - 1}
    # ternary {0, 0, 1, -1}: about half zeros
    assert 0.4 < (r == 0).mean() < 0.6
