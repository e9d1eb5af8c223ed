"""The oracle's gate bootstrap pinned semantically on the CPU: for every input pair, the
NAND output decrypts to 1 - (a & b) with noise below Q/16, the reference's own example
check (tfhe/examples/nand_goldilocks.rs:16-24, 51-53), on BabyBear and Goldilocks keys.
A planted mistake anywhere on the path (rotation sign, digit order, key row, LUT offset,
key switch) breaks the decryption or the noise bound."""
import numpy as np
import pytest


@pytest.mark.parametrize("set_id", [0, 1])
def test_oracle_nand_truth_table_and_noise(hb, orc, keyset, set_id):
    P, sk, bsk, ksk = keyset(set_id)
    enc, dec = hb.Encryptor(sk), hb.Decryptor(sk)
    pairs = [(0, 0), (0, 1), (1, 0), (1, 1)]
    c0 = np.stack([enc.encrypt(a).words for a, _ in pairs])
    c1 = np.stack([enc.encrypt(b).words for _, b in pairs])
    out = orc.OracleKeys(set_id, bsk, ksk).gate_batch(0, c0, c1, threads=4)
    for (a, b), words in zip(pairs, out):
        m, noise = dec.decrypt_with_noise(hb.LWECiphertext(words))
        assert m == 1 - (a & b), (set_id, a, b)
        assert noise < P.q / 16, (set_id, a, b, noise)
