"""Key generation on the GPU: hb_keygen_dev's BSK and KSK equal hb_keygen_host's word for
word for every parameter set; an evaluator created from a secret key (device keygen and the
device-side conversion to the kernels' word form) and one read back from the serialized
format bootstrap exactly as the oracle does with the host keys."""
import sys
import time

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("set_id", [0, 1, 2], ids=["babybear", "goldilocks", "zama"])
def test_device_keys_equal_host_keys(hb, keyset, set_id):
    P, sk, bsk, ksk = keyset(set_id)
    t = time.perf_counter()
    dbsk, dksk = sk.generate_keys_device()
    say(f"set {set_id}: device keygen + copy {1e3 * (time.perf_counter() - t):.1f} ms")
    assert dbsk.shape == bsk.shape and dksk.shape == ksk.shape
    bad = np.flatnonzero(dbsk != bsk)
    assert bad.size == 0, (bad.size, bad[:8])
    assert np.array_equal(dksk, ksk)


@pytest.mark.parametrize("set_id", [0, 1], ids=["babybear", "goldilocks"])
def test_evaluator_from_secret_key_and_from_file(hb, orc, keyset, set_id):
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(sk)
    eb, ek = ev.export_keys()
    assert np.array_equal(eb, bsk) and np.array_equal(ek, ksk)
    data = ev.serialize()
    ev2 = hb.Evaluator.deserialize(data)
    assert ev2.params.set_id == set_id
    rng = stream(70 + set_id, "keygen-gates")
    enc = hb.Encryptor(sk)
    b0, b1 = rng.integers(0, 2, size=(2, 24))
    c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
    got = ev.gate_batch("nand", c0, c1)
    got2 = ev2.gate_batch("nand", c0, c1)
    ref = orc.OracleKeys(set_id, bsk, ksk).gate_batch(0, c0, c1, threads=8)
    assert np.array_equal(got, ref)
    assert np.array_equal(got2, ref)
