"""The oracle's traced bootstrap (vfhe path) pinned by the relations the reference's trace
crate states (trace/src/acc_trace.rs:26-47) and by closed forms, on BabyBear keys, Goldilocks
keys, and Goldilocks keys that drive the accumulator into raw words above P (where the
traced path's exact monomial products and the untraced path's raw negations differ):
  input_acc[r+1] = output_acc[r];  output_acc = input_acc + sum_prod.poly;
  monomial_times_acc = monomial * input_acc (schoolbook);  ext_input = product - input_acc;
  every NTT half = NTT(its polynomial);  Σ_j bit_j B^j = ext_input;
  sum_prod.ntt = Σ_j bit_ntt_j ⊙ key_j;  INTT(key) = recorded key polynomial;
  modulus-switch output = round(v 2N / Q);  sample extraction = [A0, -A_{N-1}, ..., -A_1].
"""
import math

import numpy as np
import pytest

from conftest import stream, uniform_field

BP = 2**31 - 2**27 + 1
GP = 2**64 - 2**32 + 1


@pytest.fixture(scope="module", params=["babybear", "goldilocks", "goldilocks-raw-words"])
def traced(request, hb, orc, keyset):
    kind = request.param
    set_id = hb.SET_BABYBEAR if kind == "babybear" else hb.SET_GOLDILOCKS
    rounds = [0, 1, 2, 3, 500, 726, 727]
    if kind != "goldilocks-raw-words":
        P, sk, bsk, ksk = keyset(set_id)
        enc = hb.Encryptor(sk)
        c0, c1 = enc.encrypt(1).words, enc.encrypt(0).words
        c = ((c0.astype(object) + c1.astype(object)) % P.q).astype(np.uint64)
        lut = orc.build_lut(orc.params(set_id), 0)
    else:
        # every RGSW zero except c_m[0].b = Q-1 in all slots: each round adds
        # -digit_0((X^a - 1) ACC.b) to ACC.b, and with small LUT words the sums land in
        # [P, 2^64) in the untraced path
        P = hb.Parameters.of(set_id)
        n, N, L = P.n, P.N, P.br_len
        bsk = np.zeros((n, 2, L, 2, N), np.uint64)
        bsk[:, 1, 0, 1, :] = GP - 1
        bsk = bsk.reshape(-1)
        ksk = uniform_field(stream(6, "ksk"), GP, P.ks_len * 2 * N)
        rng = stream(6, "adv")
        lut = rng.integers(1, 1000, size=N, dtype=np.uint64)
        c = uniform_field(rng, GP, n + 1)
        c0 = c1 = None
    ok = orc.OracleKeys(set_id, bsk, ksk)
    out, rec, tail = ok.bootstrap_traced(c, lut, rounds)
    return dict(kind=kind, gl=set_id != hb.SET_BABYBEAR, P=P, bsk=bsk, ksk=ksk, ok=ok, c=c,
                c0=c0, c1=c1, lut=lut, rounds=rounds, out=out, rec=rec, tail=tail)


def fields(rec, N, L):
    f = {}
    f["in_a"], f["in_b"], f["in_na"], f["in_nb"] = (rec[k * N:(k + 1) * N] for k in range(4))
    f["out_a"], f["out_b"], f["out_na"], f["out_nb"] = (rec[(4 + k) * N:(5 + k) * N] for k in range(4))
    f["mono"], f["mono_ntt"] = rec[8 * N:9 * N], rec[9 * N:10 * N]
    f["deg"], f["coef"] = int(rec[10 * N]), int(rec[10 * N + 1])
    o = 10 * N + 2
    f["prod_a"], f["prod_b"], f["prod_na"], f["prod_nb"] = (rec[o + k * N:o + (k + 1) * N] for k in range(4))
    o = 14 * N + 2
    f["ext_a"], f["ext_b"], f["ext_na"], f["ext_nb"] = (rec[o + k * N:o + (k + 1) * N] for k in range(4))
    o = 18 * N + 2
    f["had"] = [rec[o + 6 * j * N:o + 6 * (j + 1) * N].reshape(6, N) for j in range(2 * L)]
    o = (18 + 12 * L) * N + 2
    f["sp_na"], f["sp_nb"], f["sp_a"], f["sp_b"] = (rec[o + k * N:o + (k + 1) * N] for k in range(4))
    return f


def addm(x, y, q):
    return ((x.astype(object) + y.astype(object)) % q).astype(np.uint64)


def subm(x, y, q):
    return ((x.astype(object) - y.astype(object)) % q).astype(np.uint64)


def test_traced_output_against_untraced(traced, orc):
    """Real keys: both paths give the same words. Raw-word keys: they differ (the untraced
    path negates raw words above P, which the traced path never holds)."""
    if traced["kind"] == "goldilocks-raw-words":
        orc.raw_neg_events(reset=True)
        ref = traced["ok"].bootstrap(traced["c"], traced["lut"])
        assert orc.raw_neg_events() > 0
        assert not np.array_equal(traced["out"], ref)
    else:
        ref = traced["ok"].gate_batch(0, traced["c0"][None], traced["c1"][None])[0]
        assert np.array_equal(traced["out"], ref)


def test_round_relations(traced, orc):
    P = traced["P"]
    N, L, q = P.N, P.br_len, P.q
    gl = traced["gl"]
    recs = [fields(r, N, L) for r in traced["rec"]]
    for idx, (r, f) in enumerate(zip(traced["rounds"], recs)):
        for a, na in [("in_a", "in_na"), ("in_b", "in_nb"), ("out_a", "out_na"), ("out_b", "out_nb"),
                      ("mono", "mono_ntt"), ("prod_a", "prod_na"), ("prod_b", "prod_nb"),
                      ("ext_a", "ext_na"), ("ext_b", "ext_nb")]:
            assert np.array_equal(orc.ntt(gl, f[a]), f[na]), (r, a)
        assert np.array_equal(addm(f["in_a"], f["sp_a"], q), f["out_a"])
        assert np.array_equal(addm(f["in_b"], f["sp_b"], q), f["out_b"])
        mono = np.zeros(N, np.uint64)
        mono[f["deg"]] = f["coef"]
        assert np.array_equal(mono, f["mono"])
        assert f["coef"] in (1, q - 1)
        assert np.array_equal(orc.schoolbook(gl, f["mono"], f["in_a"]), f["prod_a"])
        assert np.array_equal(orc.schoolbook(gl, f["mono"], f["in_b"]), f["prod_b"])
        assert np.array_equal(subm(f["prod_a"], f["in_a"], q), f["ext_a"])
        assert np.array_equal(subm(f["prod_b"], f["in_b"], q), f["ext_b"])
        # decomposition and the hadamard sum
        B = 1 << P.br_bits
        for half, src in ((0, "ext_a"), (1, "ext_b")):
            rebuilt = sum(f["had"][half * L + j][0].astype(object) * B ** j for j in range(L))
            assert np.array_equal(np.array(rebuilt, dtype=object) % q, f[src].astype(object))
        acc_a = np.zeros(N, dtype=object)
        acc_b = np.zeros(N, dtype=object)
        for h in f["had"]:
            assert np.array_equal(orc.ntt(gl, h[0]), h[1])
            assert np.array_equal(orc.ntt(gl, h[2], inverse=True), h[4])
            assert np.array_equal(orc.ntt(gl, h[3], inverse=True), h[5])
            assert (h[0] < B).all()
            acc_a = (acc_a + h[1].astype(object) * h[2].astype(object)) % q
            acc_b = (acc_b + h[1].astype(object) * h[3].astype(object)) % q
        assert np.array_equal(acc_a.astype(np.uint64), f["sp_na"])
        assert np.array_equal(acc_b.astype(np.uint64), f["sp_nb"])
        assert np.array_equal(orc.ntt(gl, f["sp_na"], inverse=True), f["sp_a"])
        # consecutive rounds chain
        if idx + 1 < len(recs) and traced["rounds"][idx + 1] == r + 1:
            g = recs[idx + 1]
            assert np.array_equal(f["out_a"], g["in_a"]) and np.array_equal(f["out_b"], g["in_b"])


def round_half_away(x: float) -> int:
    """f64 rounding half away from zero (x >= 0), as Rust's f64::round."""
    f = math.floor(x)
    return f + (1 if x - f >= 0.5 else 0)


def test_tail_relations(traced, orc):
    P = traced["P"]
    N, n, q = P.N, P.n, P.q
    gl = traced["gl"]
    t = traced["tail"]
    ms_in, ms_out = t[:n + 1], t[n + 1:2 * (n + 1)]
    assert np.array_equal(ms_in, traced["c"])
    expect = [round_half_away(float(v) * (2 * N) / float(q)) % (2 * N) for v in ms_in]
    assert list(ms_out) == expect
    o = 2 * (n + 1)
    dec = t[o:o + N]
    o += N
    had = [t[o + 6 * j * N:o + 6 * (j + 1) * N].reshape(6, N) for j in range(P.ks_len)]
    o += 6 * P.ks_len * N
    sp_na, sp_nb, sp_a, sp_b, se_in, se_out = (t[o + k * N:o + (k + 1) * N] for k in range(6))
    B = 1 << P.ks_bits
    rebuilt = sum(h[0].astype(object) * B ** j for j, h in enumerate(had)) % q
    assert np.array_equal(rebuilt.astype(np.uint64), dec)
    assert np.array_equal(orc.ntt(gl, sp_na, inverse=True), sp_a)
    assert np.array_equal(se_in, sp_a)
    assert se_out[0] == se_in[0]
    assert np.array_equal(se_out[1:], ((-se_in[1:][::-1].astype(object)) % q).astype(np.uint64))
    assert np.array_equal(traced["out"][:n], se_out[:n])
