"""Oracle negacyclic NTT pinned by mathematics.

The reference NTT lives in the un-vendored concrete-ntt crate; its slot order is unpinned,
so the engine documents its own: forward slot i = X(psi^(2*bitrev(i)+1)), psi = g^((P-1)/2N).
Checks: the O(N^2) evaluation formula, psi's order, round trips (the reference's own
goldilocks_ntt.rs:172-183 and babybear_ntt.rs:168-214 tests), the convolution theorem
against schoolbook negacyclic products, and the monomial closed form
(algebra/src/transformation/ntt_table.rs:202-244).
"""
import numpy as np
import pytest

from conftest import stream, uniform_field

GP = 2**64 - 2**32 + 1
BP = 2**31 - 2**27 + 1


def bitrev(x, bits):
    return int(format(x, f"0{bits}b")[::-1], 2)


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
@pytest.mark.parametrize("log_n", [4, 6, 10, 11])
def test_psi_is_primitive_2n_root(orc, gold, q, log_n):
    psi = orc.ntt_psi(gold, log_n)
    n = 1 << log_n
    assert pow(psi, n, q) == q - 1
    assert pow(psi, 2 * n, q) == 1


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
@pytest.mark.parametrize("log_n", [3, 5, 7])
def test_ntt_matches_evaluation_formula(orc, gold, q, log_n):
    n = 1 << log_n
    rng = stream(11, f"ev{log_n}{int(gold)}")
    x = uniform_field(rng, q, n)
    y = orc.ntt(gold, x)
    psi = orc.ntt_psi(gold, log_n)
    xs = [int(v) for v in x]
    for i in range(n):
        w = pow(psi, 2 * bitrev(i, log_n) + 1, q)
        expect = sum(c * pow(w, k, q) for k, c in enumerate(xs)) % q
        assert int(y[i]) == expect


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
def test_ntt_n1024_sampled_slots_match_formula(orc, gold, q):
    log_n, n = 10, 1024
    x = uniform_field(stream(12, "ev1024"), q, n)
    y = orc.ntt(gold, x)
    psi = orc.ntt_psi(gold, log_n)
    xs = [int(v) for v in x]
    for i in [0, 1, 2, 511, 512, 1000, 1023]:
        w = pow(psi, 2 * bitrev(i, log_n) + 1, q)
        acc, wk = 0, 1
        for c in xs:
            acc = (acc + c * wk) % q
            wk = wk * w % q
        assert int(y[i]) == acc


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
@pytest.mark.parametrize("log_n", [10, 11])
def test_ntt_roundtrip_and_range(orc, gold, q, log_n):
    n = 1 << log_n
    x = uniform_field(stream(13, "rt"), q, (8, n))
    y = orc.ntt(gold, x)
    assert (y < np.uint64(q)).all()
    z = orc.ntt(gold, y, inverse=True)
    assert np.array_equal(x, z)


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
def test_convolution_theorem_vs_schoolbook(orc, gold, q):
    n = 1024
    rng = stream(14, "conv")
    a = uniform_field(rng, q, n)
    b = uniform_field(rng, q, n)
    prod = orc.ntt(gold, _pointwise(orc.ntt(gold, a), orc.ntt(gold, b), q), inverse=True)
    assert np.array_equal(prod, orc.schoolbook(gold, a, b))


def _pointwise(x, y, q):
    return ((x.astype(object) * y.astype(object)) % q).astype(np.uint64)


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
def test_schoolbook_small_case_by_hand(orc, gold, q):
    # (1 + X) * X^(N-1) = X^(N-1) + X^N = -1 + X^(N-1) mod X^N + 1
    n = 16
    a = np.zeros(n, np.uint64); a[0] = 1; a[1] = 1
    b = np.zeros(n, np.uint64); b[n - 1] = 1
    c = orc.schoolbook(gold, a, b)
    expect = np.zeros(n, np.uint64); expect[0] = q - 1; expect[n - 1] = 1
    assert np.array_equal(c, expect)


@pytest.mark.parametrize("gold,q", [(True, GP), (False, BP)])
def test_monomial_closed_form(orc, gold, q):
    log_n, n = 10, 1024
    psi = orc.ntt_psi(gold, log_n)
    for d, coeff in [(0, 1), (1, 1), (513, q - 1), (1023, 7)]:
        x = np.zeros(n, np.uint64)
        x[d] = coeff
        y = orc.ntt(gold, x)
        for i in [0, 5, 700, 1023]:
            assert int(y[i]) == coeff * pow(psi, (2 * bitrev(i, log_n) + 1) * d, q) % q
