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


def test_device_add_and_rotation_shortcuts(orc):
    """The kernels' rewrites of the reference's raw-word steps give the same words.

    add_cy (hb_field.hpp): for a canonical addend the carry's EPS cannot carry again, so one
    correction gives add's word (ops.rs:27-34). rotate_sub (Goldilocks): neg_raw(v) is
    replaced by P - v mod 2^64 (they differ only at v = 0, where P = 0 mod P) and the
    difference is one raw sub (ops.rs:48-55) then a canonicalisation; the reference
    (rlwe.rs:445-481) negates the raw word, so for v > P the result is -v + EPS, not -v."""
    rng = np.random.default_rng(5)
    rand = [int(v) for v in rng.integers(0, 2**63, 200, dtype=np.uint64) * 2 + 1]
    words = EDGES + rand
    canon = [w for w in words if w < GP]
    xs = np.array([x for x in words for _ in canon], dtype=np.uint64)
    ys = np.array([y for _ in words for y in canon], dtype=np.uint64)
    add = orc.gl_op(0, xs, ys)
    for x, y, r in zip(xs.tolist(), ys.tolist(), add.tolist()):
        s = (x + y) % M64
        assert r == (s + (EPS if x + y >= M64 else 0)) % M64
    # rotation: dst = canon(sub(P - v, a)) against (neg_raw(v) - a) mod P and
    # canon(sub(v, a)) against (v - a) mod P, for raw v and a
    vs = np.array([v for v in words for _ in words], dtype=np.uint64)
    As = np.array([a for _ in words for a in words], dtype=np.uint64)
    negv = np.array([(GP - v) % M64 for v in vs.tolist()], dtype=np.uint64)
    d_neg = orc.gl_op(1, negv, As)
    d_pos = orc.gl_op(1, vs, As)
    for v, a, dn, dp in zip(vs.tolist(), As.tolist(), d_neg.tolist(), d_pos.tolist()):
        assert dn % GP == (ref_neg(v) - a) % GP
        assert dp % GP == (v - a) % GP


def test_rotation_index_of_canon_and_negated_halves():
    """rotate_sub_cn reads word (k - r) mod 2N of [C | M] (C = canon(acc), M =
    canon(neg_raw(acc))): the same source word and sign as the reference's negacyclic
    rotation (rlwe.rs:445-481: for r <= N the words k < r wrap with a sign flip, for r > N
    the rotation by r - N flips the words k >= r - N)."""
    for n in (8, 64, 1024):
        for r in range(1, 2 * n):
            for k in range(n):
                if r <= n:
                    neg, src = k < r, (k - r + n if k < r else k - r)
                else:
                    rr = r - n
                    neg, src = k >= rr, (k - rr if k >= rr else k - rr + n)
                t = (k - r) & (2 * n - 1)
                assert (t >= n) == neg and t % n == src, (n, r, k)
            if n == 1024 and r > 64:
                break
