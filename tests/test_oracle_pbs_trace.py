"""The oracle's finalized PBSTrace containers (what the prover takes after
BlindRotationTrace::finalize(n) and ModulusSwitchingTrace::finalize(n + 1),
vfhe/examples/zk_binary_pbs/zk_nand_goldilocks.rs:97-105) pinned by the relations the trace
crate states, on Goldilocks and BabyBear keys:
  initial_acc = input_acc[0], final_acc = output_acc[n-1], input_acc[r+1] = output_acc[r]
    (trace/src/acc_trace.rs:26-47);
  input_acc_permuted = permutation_matrix * input_acc on its poly planes, its last row
    = output_acc last + initial - final, its ntt planes output_acc's (acc_trace.rs:157-190);
  permutation_info = rotation left by one of the first n rows (row_perm_trace.rs:41-52);
  output_acc = input_acc + sum_prod.poly; monomial_times_acc = monomial * input_acc;
    external_product_input = monomial_times_acc - input_acc; sum_prod.ntt = Σ_j bit.ntt ⊙
    rlwe.ntt; Σ_j bit_j B^j = external_product_input; rlwe.ntt = the round's key rows;
  every ntt plane = NTT(its poly plane); rows n..R-1 zero (rlwe_trace.rs:103-192);
  modulus switching: b' = b or q when b = 0, e = a - (2b'-1)k - 1 lies in [0, 2k] (the
    rounding bound the trace crate states), 2qk = Q-1, padding b' = q and e = k
    (modulus_switching_trace.rs:25-49, 112-133).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

SAMPLE_ROWS = (0, 1, 363, 726, 727)


def part_id(orc, name):
    return orc.TRACE_PARTS.index(name)


@pytest.fixture(scope="module", params=["goldilocks", "babybear"])
def fin(request, hb, orc, keyset):
    set_id = hb.SET_GOLDILOCKS if request.param == "goldilocks" else hb.SET_BABYBEAR
    P, sk, bsk, ksk = keyset(set_id)
    enc = hb.Encryptor(sk)
    c0, c1 = enc.encrypt(0).words, enc.encrypt(1).words
    c = ((c0.astype(object) + c1.astype(object)) % P.q).astype(np.uint64)
    lut = orc.build_lut(orc.params(set_id), 0)
    ok = orc.OracleKeys(set_id, bsk, ksk)
    N, n, L = P.N, P.n, P.br_len
    R = 1 << (n - 1).bit_length()
    jobs = [(name, 0) for name in orc.TRACE_PARTS if name != "hadamard"]
    jobs += [("hadamard", j) for j in range(2 * L)]

    def run(job):
        name, j = job
        x = ok.pbs_trace_part(c, lut, part_id(orc, name), j)
        if name in ("permutation_info", "modulus_switching"):
            return job, x
        if name == "monomial_representation":
            return job, x.reshape(2, R)
        planes = {"monomial": 2, "hadamard": 6}.get(name, 4)
        x = x.reshape(planes, -1, N)
        if name == "hadamard":  # keep sampled rows and the padding check only
            assert not x[:, n:].any()
            bits_ok = bool((x[0, :n] < (1 << P.br_bits)).all())
            return job, (x[:, list(SAMPLE_ROWS)].copy(), x[0, :n].copy(), bits_ok)
        return job, x

    with ThreadPoolExecutor(8) as pool:
        got = dict(pool.map(run, jobs))
    out, rec, tail = ok.bootstrap_traced(c, lut, list(SAMPLE_ROWS))
    return dict(P=P, bsk=bsk.reshape(n, 2, L, 2, N), c=c, got=got, gl=set_id != hb.SET_BABYBEAR,
                rec=rec, tail=tail, R=R)


def acc(f, name):
    return f["got"][(name, 0)]


def addq(x, y, q):
    return ((x.astype(object) + y.astype(object)) % q).astype(np.uint64)


def subq(x, y, q):
    return ((x.astype(object) - y.astype(object)) % q).astype(np.uint64)


def test_shapes_and_padding(fin):
    P, R = fin["P"], fin["R"]
    n, N = P.n, P.N
    assert R == 1024
    for name in ("input_acc", "output_acc", "input_acc_permuted", "monomial_times_acc",
                 "external_product_input", "sum_prod", "monomial"):
        x = acc(fin, name)
        assert x.shape[1:] == (R, N), name
        assert not x[:, n:].any(), name
    assert acc(fin, "initial_acc").shape == (4, 1, N)
    assert acc(fin, "final_acc").shape == (4, 1, N)
    assert not acc(fin, "monomial_representation")[:, n:].any()


def test_acc_chain_and_endpoints(fin):
    n = fin["P"].n
    ia, oa = acc(fin, "input_acc"), acc(fin, "output_acc")
    assert np.array_equal(acc(fin, "initial_acc")[:, 0], ia[:, 0])
    assert np.array_equal(acc(fin, "final_acc")[:, 0], oa[:, n - 1])
    assert np.array_equal(ia[:, 1:n], oa[:, :n - 1])
    # rows agree with the per-round records (their own relations are pinned elsewhere)
    N = fin["P"].N
    for k, r in enumerate(SAMPLE_ROWS):
        rec = fin["rec"][k]
        assert np.array_equal(ia[:, r].reshape(-1), rec[:4 * N])
        assert np.array_equal(oa[:, r].reshape(-1), rec[4 * N:8 * N])


def test_input_acc_permuted(fin, orc):
    P, R = fin["P"], fin["R"]
    n, q = P.n, P.q
    perm = acc(fin, "permutation_info").astype(np.int64)
    assert sorted(perm.tolist()) == list(range(R))
    assert perm[n - 1] == 0 and (perm[:n - 1] == np.arange(1, n)).all()
    assert (perm[n:] == np.arange(n, R)).all()
    ia, oa, iap = acc(fin, "input_acc"), acc(fin, "output_acc"), acc(fin, "input_acc_permuted")
    init, final = acc(fin, "initial_acc")[:, 0], acc(fin, "final_acc")[:, 0]
    # input_acc_permuted = permutation_matrix * input_acc (poly planes)
    assert np.array_equal(iap[:2], ia[:2, perm])
    # = output_acc + Z, Z zero but for its last row initial - final
    for h in range(2):
        assert np.array_equal(iap[h, n - 1], addq(oa[h, n - 1], subq(init[h], final[h], q), q))
    assert np.array_equal(iap[:2, :n - 1], oa[:2, :n - 1])
    # the ntt planes are output_acc's, untouched by the last-row update
    assert np.array_equal(iap[2:], oa[2:])
    assert not np.array_equal(iap[2, n - 1], orc.ntt(fin["gl"], iap[0, n - 1]))


def test_ntt_planes_are_transforms(fin, orc):
    gl = fin["gl"]
    for name in ("input_acc", "output_acc", "monomial_times_acc", "external_product_input",
                 "sum_prod"):
        x = acc(fin, name)
        for r in SAMPLE_ROWS:
            for h in range(2):
                assert np.array_equal(orc.ntt(gl, x[h, r]), x[2 + h, r]), (name, r, h)
    m = acc(fin, "monomial")
    for r in SAMPLE_ROWS:
        assert np.array_equal(orc.ntt(gl, m[0, r]), m[1, r])
    for j in range(2 * fin["P"].br_len):
        h = fin["got"][("hadamard", j)][0]
        for k in range(len(SAMPLE_ROWS)):
            assert np.array_equal(orc.ntt(gl, h[0, k]), h[1, k])
            assert np.array_equal(orc.ntt(gl, h[2, k]), h[4, k])
            assert np.array_equal(orc.ntt(gl, h[3, k]), h[5, k])


def test_round_products(fin, orc):
    P = fin["P"]
    n, N, q, L, gl = P.n, P.N, P.q, P.br_len, fin["gl"]
    ia, oa, sp = acc(fin, "input_acc"), acc(fin, "output_acc"), acc(fin, "sum_prod")
    mono, rep = acc(fin, "monomial"), acc(fin, "monomial_representation")
    mta, ext = acc(fin, "monomial_times_acc"), acc(fin, "external_product_input")
    for h in range(2):
        assert np.array_equal(oa[h, :n], addq(ia[h, :n], sp[h, :n], q))
        assert np.array_equal(ext[h, :n], subq(mta[h, :n], ia[h, :n], q))
    deg, coef = rep[0, :n].astype(np.int64), rep[1, :n]
    assert (deg < N).all() and np.isin(coef, [1, q - 1]).all()
    expect = np.zeros((n, N), np.uint64)
    expect[np.arange(n), deg] = coef
    assert np.array_equal(mono[0, :n], expect)
    for r in SAMPLE_ROWS:
        for h in range(2):
            assert np.array_equal(orc.schoolbook(gl, mono[0, r], ia[h, r]), mta[h, r])
    # gadget decomposition of every round: Σ_j bit_j B^j = external_product_input
    B = 1 << P.br_bits
    for half in range(2):
        rebuilt = np.zeros((n, N), dtype=object)
        for d in range(L):
            _, bits, ok = fin["got"][("hadamard", half * L + d)]
            assert ok
            rebuilt = rebuilt + bits.astype(object) * B ** d
        assert np.array_equal(rebuilt % q, ext[half, :n].astype(object))
    # key rows and the NTT-domain sum, sampled rows
    for k, r in enumerate(SAMPLE_ROWS):
        sa = np.zeros(N, dtype=object)
        sb = np.zeros(N, dtype=object)
        for j in range(2 * L):
            h = fin["got"][("hadamard", j)][0][:, k]
            half, d = divmod(j, L)
            assert np.array_equal(h[4], fin["bsk"][r, half, d, 0])
            assert np.array_equal(h[5], fin["bsk"][r, half, d, 1])
            sa = (sa + h[1].astype(object) * h[4].astype(object)) % q
            sb = (sb + h[1].astype(object) * h[5].astype(object)) % q
        assert np.array_equal(sa.astype(np.uint64), sp[2, r])
        assert np.array_equal(sb.astype(np.uint64), sp[3, r])


def test_modulus_switching_finalize(fin):
    P = fin["P"]
    n, N, Q = P.n, P.N, P.q
    M = 1 << n.bit_length()
    ms = acc(fin, "modulus_switching")
    assert ms.size == 4 * M + 3
    a, b, wit, e = (ms[k * M:(k + 1) * M].astype(object) for k in range(4))
    q2, k, two_k = (int(x) for x in ms[4 * M:])
    assert q2 == 2 * N and 2 * q2 * k == Q - 1 and two_k == 2 * k
    assert np.array_equal(a[:n + 1], fin["c"].astype(object))
    assert np.array_equal(b[:n + 1], fin["tail"][n + 1:2 * (n + 1)].astype(object))
    for i in range(n + 1):
        assert wit[i] == (q2 if b[i] == 0 else b[i])
        assert 1 <= wit[i] <= q2
        assert (a[i] - (2 * wit[i] - 1) * k - 1 - e[i]) % Q == 0
        assert 0 <= e[i] <= 2 * k  # the rounding bound: b = round(a q / Q)
    assert not a[n + 1:].any() and not b[n + 1:].any()
    assert (wit[n + 1:] == q2).all() and (e[n + 1:] == k).all()


def test_planted_mistakes_fail_the_pins(fin, orc):
    """The checks above catch a shifted permutation and an ntt plane rebuilt from the
    modified poly row (mistakes planted in copies of the oracle's output)."""
    n = fin["P"].n
    ia, iap, oa = acc(fin, "input_acc"), acc(fin, "input_acc_permuted"), acc(fin, "output_acc")
    perm = acc(fin, "permutation_info").astype(np.int64)
    wrong = np.roll(perm, 1)
    assert not np.array_equal(iap[:2], ia[:2, wrong])
    bad = iap.copy()
    bad[2, n - 1] = orc.ntt(fin["gl"], iap[0, n - 1])
    assert not np.array_equal(bad[2:], oa[2:])
