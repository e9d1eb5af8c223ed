"""hb_gate_batch_dev_noalloc: the batch entry point whose scratch is the library's fixed
device array. Through the C ABI, two calls in a row (the second smaller, so it runs over
what the first left in the array) give the same words as hb_gate_batch_dev on the same
inputs; sizes beyond the array are refused before anything is launched."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import stream

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(L, check, a):
    p = C.c_void_p()
    check(L.hb_dev_alloc(a.nbytes, C.byref(p)))
    check(L.hb_memcpy_h2d(p, a.ctypes.data, a.nbytes))
    return p


@pytest.mark.parametrize("set_id,sizes", [(1, (900, 40)), (0, (1500, 300))],
                         ids=["goldilocks", "babybear"])
def test_noalloc_matches_batch_dev(hb, keyset, set_id, sizes):
    from hasteboots_amd._lib import check, lib
    L = lib()
    P, sk, bsk, ksk = keyset(set_id)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    enc = hb.Encryptor(sk)
    rng = stream(12 + set_id, "noalloc")
    for rep, count in enumerate(sizes):
        say(f"no-alloc set {set_id}, call {rep}: {count} gates")
        b0, b1 = rng.integers(0, 2, size=(2, count))
        c0, c1 = enc.encrypt_many(b0), enc.encrypt_many(b1)
        d0, d1 = _dev(L, check, c0), _dev(L, check, c1)
        d_ref, d_new = _dev(L, check, np.zeros_like(c0)), _dev(L, check, np.zeros_like(c0))
        try:
            check(L.hb_gate_batch_dev(ev._h, 0, count, d0, d1, None, d_ref, None))
            check(L.hb_gate_batch_dev_noalloc(ev._h, 0, count, d0, d1, None, d_new, None))
            check(L.hb_sync())
            ref = np.zeros_like(c0)
            got = np.zeros_like(c0)
            check(L.hb_memcpy_d2h(ref.ctypes.data, d_ref, ref.nbytes))
            check(L.hb_memcpy_d2h(got.ctypes.data, d_new, got.nbytes))
        finally:
            for p in (d0, d1, d_ref, d_new):
                L.hb_dev_free(p)
        assert np.array_equal(got, ref), (set_id, rep)
        dec = hb.Decryptor(sk)
        assert all(dec.decrypt(hb.LWECiphertext(got[i])) == 1 - (int(b0[i]) & int(b1[i]))
                   for i in range(count))


def test_noalloc_refuses_sizes_beyond_the_array(hb, keyset):
    from hasteboots_amd._lib import lib
    L = lib()
    P, sk, bsk, ksk = keyset(1)
    ev = hb.Evaluator(keys=(bsk, ksk), params=P)
    one = C.c_void_p(16)  # never dereferenced: the call is refused before any launch
    assert L.hb_gate_batch_dev_noalloc(ev._h, 0, 1025, one, one, None, one, None) == -4
    assert b"fixed scratch" in L.hb_last_error()
    assert L.hb_gate_batch_dev_noalloc(ev._h, 8, 4, one, one, one, one, None) == -1
