"""Oracle field arithmetic against closed forms (Python big integers).

Mirrors the reference's own modulus tests (algebra/src/modulus/goldilocks/mod.rs:175-278,
algebra/src/modulus/baby_bear/mod.rs:141-258) and adds the raw-word semantics that only the
source text fixes (algebra/src/modulus/goldilocks/ops.rs:27-71): which word in [0, 2^64) an
operation returns, including the wrapping negation of a raw word above P.
"""
import numpy as np

GP = 2**64 - 2**32 + 1
EPS = 2**32 - 1
BP = 2**31 - 2**27 + 1
M64 = 2**64

EDGES = [0, 1, 2, EPS, EPS + 1, GP - 2, GP - 1, GP, GP + 1, GP + 7, 2**63, M64 - 2, M64 - 1]


def ref_add(x, y):  # ops.rs:27-34, step by step
    s = (x + y) % M64
    over = x + y >= M64
    s2 = s + (EPS if over else 0)
    over2 = s2 >= M64
    s2 %= M64
    if over2:
        s2 += EPS
    return s2


def ref_sub(x, y):  # ops.rs:48-55
    d = (x - y) % M64
    under = x < y
    adj = EPS if under else 0
    d2 = (d - adj) % M64
    if d < adj:
        d2 -= EPS
    return d2


def ref_neg(x):  # ops.rs:69-71, release-mode wrapping
    return 0 if x == 0 else (GP - x) % M64


def test_goldilocks_raw_ops_exhaustive_edges(orc):
    xs = np.array([a for a in EDGES for _ in EDGES], dtype=np.uint64)
    ys = np.array([b for _ in EDGES for b in EDGES], dtype=np.uint64)
    add = orc.gl_op(0, xs, ys)
    sub = orc.gl_op(1, xs, ys)
    mul = orc.gl_op(3, xs, ys)
    for i, (x, y) in enumerate(zip(xs.tolist(), ys.tolist())):
        assert int(add[i]) == ref_add(x, y)
        assert int(sub[i]) == ref_sub(x, y)
        # raw results are congruent to the true values
        assert int(add[i]) % GP == (x + y) % GP
        assert int(sub[i]) % GP == (x - y) % GP
        assert int(mul[i]) % GP == (x * y) % GP
        assert 0 <= int(mul[i]) < M64


def test_goldilocks_neg_wraps_for_raw_words(orc):
    xs = np.array(EDGES, dtype=np.uint64)
    neg = orc.gl_op(2, xs)
    for x, r in zip(EDGES, neg.tolist()):
        assert r == ref_neg(x)
        if x <= GP:
            assert r % GP == (-x) % GP
        else:
            # the reference's raw negation is off by 2^32 - 1 for x > P
            assert r % GP == (-x + EPS) % GP


def test_goldilocks_random_mul_and_value(orc):
    rng = np.random.default_rng(1)
    x = rng.integers(0, 2**64, size=20000, dtype=np.uint64)
    y = rng.integers(0, 2**64, size=20000, dtype=np.uint64)
    mul = orc.gl_op(3, x, y)
    val = orc.gl_op(4, x)
    for a, b, m, v in zip(x.tolist(), y.tolist(), mul.tolist(), val.tolist()):
        assert m % GP == (a * b) % GP
        assert v == a % GP


def test_goldilocks_reduce128_closed_form(orc):
    rng = np.random.default_rng(2)
    lo = rng.integers(0, 2**64, size=5000, dtype=np.uint64)
    hi = rng.integers(0, 2**64, size=5000, dtype=np.uint64)
    r = orc.gl_op(5, lo, hi)
    for a, b, c in zip(lo.tolist(), hi.tolist(), r.tolist()):
        assert c % GP == (a + (b << 64)) % GP


def test_babybear_ops_closed_form(orc):
    rng = np.random.default_rng(3)
    x = rng.integers(0, BP, size=20000, dtype=np.uint64)
    y = rng.integers(0, BP, size=20000, dtype=np.uint64)
    edge = np.array([0, 1, BP - 1, BP - 2, 2**30, 2**27], dtype=np.uint64)
    x = np.concatenate([x, np.repeat(edge, len(edge))])
    y = np.concatenate([y, np.tile(edge, len(edge))])
    add, sub, neg, mul = (orc.bb_op(k, x, y) for k in (0, 1, 2, 3))
    xi, yi = x.astype(object), y.astype(object)
    assert (add.astype(object) == (xi + yi) % BP).all()
    assert (sub.astype(object) == (xi - yi) % BP).all()
    assert (neg.astype(object) == (-xi) % BP).all()
    assert (mul.astype(object) == (xi * yi) % BP).all()
