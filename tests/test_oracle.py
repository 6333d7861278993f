"""Pins the CPU oracle (oracle/zoda_oracle.c) before anything is checked against it.

Each check is independent of the oracle's own code path: brute-force O(n^2)
DFT goldens (tests/golden/make_golden.py), the reference's own test vectors
(src/ntt/cpu.rs:112-121,158-187, src/field/babybear.rs:84-105), round trips,
and ZODA invariants that the mathematics fixes (systematic even rows, codeword
degree, erasure reconstruction, RLC commuting with the extension).  A few
planted mistakes show that a plausible slip in the oracle would be caught.
"""
import numpy as np
import pytest

import oracle as O

P = O.P


def test_field_and_generator(gold):
    assert gold["generator"] == 31
    assert O.find_generator(P) == 31
    assert O.prime_factors(P - 1) == [2, 3, 5]
    assert O.find_generator(257) == gold["generator_257"] == 3
    assert O.pow_mod(2, P - 1) == 1
    assert O.pow_mod(P - 1, 2) == 1
    kat = gold["field_kat"]
    # mul and add of the field through the oracle's RLC (src/zoda/rlc.rs:32-40)
    assert O.rlc_row([100], np.array([[200]], np.uint32))[0] == kat["mul_100_200"]
    assert O.rlc_row([1, 1], np.array([[100], [200]], np.uint32))[0] == kat["add_100_200"]
    assert O.rlc_row([P - 1], np.array([[P - 1]], np.uint32))[0] == kat["neg1_squared"] == 1
    assert O.rlc_row([1, 1], np.array([[P - 1], [1]], np.uint32))[0] == kat["pm1_plus_1"] == 0


def test_roots(gold):
    assert list(O.roots_of_unity(8)) == gold["roots_8"]
    for n in (1, 2, 64, 1 << 12):
        r = O.roots_of_unity(n).astype(np.uint64)
        w = int(r[1]) if n > 1 else 1
        assert O.pow_mod(w, n) == 1
        if n > 1:
            assert O.pow_mod(w, n // 2) == P - 1  # primitive
    with pytest.raises(ValueError):
        O.roots_of_unity(7)


def test_reverse_is_bit_reversal():
    for log_n in range(0, 10):
        n = 1 << log_n
        v = O.reverse(np.arange(n))
        expect = [int(format(i, f"0{log_n}b")[::-1], 2) if log_n else 0 for i in range(n)]
        assert list(v) == expect


def test_ntt_goldens(gold):
    for case in gold["ntt"]:
        x = case["input"]
        assert list(O.ntt_babybear(x)) == case["ntt"], case["n"]
        assert list(O.intt_babybear(x)) == case["intt"], case["n"]


def test_naive_evaluation(gold):
    # src/ntt/cpu.rs:170-187
    c = gold["naive_eval_8"]
    assert list(O.ntt_babybear(c["input"])) == c["ntt"]


def test_generic_modulus_257(gold):
    # src/ntt/cpu.rs:112-121
    c = gold["p257"]
    roots = O.roots_of_unity(8, 257)
    assert list(roots) == c["roots"]
    out = O.ntt(c["input"], roots, 257)
    assert list(out) == c["ntt"]
    assert list(O.intt(out, roots, 257)) == c["input"]


@pytest.mark.parametrize("log_n", [0, 1, 4, 8, 12])
def test_babybear_roundtrip(log_n):
    # src/ntt/cpu.rs:158-168
    n = 1 << log_n
    x = (np.arange(n, dtype=np.uint64) * 7 + 3) % P
    assert np.array_equal(O.intt_babybear(O.ntt_babybear(x)), x.astype(np.uint32))


def test_encode_goldens(gold):
    for case in gold["encode"]:
        data = np.array(case["data"], dtype=np.uint32)
        out = O.encode(data, n=case["n"])
        assert out.tolist() == case["out"], (case["k"], case["n"])
        col = O.encode(data, n=case["n"], layout="col")
        assert np.array_equal(col.T, out)


def test_ramp_golden(gold):
    c = gold["ramp_4"]
    data = O.ramp_input(c["k"], c["cols"])
    out = O.encode(data)
    for col in c["sample_cols"]:
        assert out[:, col].tolist() == c["out_cols"][str(col)]


def _uniform(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, P, size=shape, dtype=np.uint32)


def test_encode_invariants():
    k, cols = 64, 96
    data = _uniform((k, cols), 1)
    out = O.encode(data, threads=4)
    # systematic: even rows reproduce the data (src/zoda/babybear.rs:201-205)
    assert np.array_equal(out[0::2], data)
    # codeword: coefficients above n_k vanish (src/zoda/babybear.rs:206-211)
    for c in (0, 17, 95):
        coeffs = O.intt_babybear(out[:, c])
        assert not coeffs[k:].any()
    # threads do not change the answer
    assert np.array_equal(O.encode(data, threads=1), out)


def _interp_eval(xs, ys, x):
    """Lagrange interpolation through (xs, ys), evaluated at x (Python ints)."""
    total = 0
    for i, (xi, yi) in enumerate(zip(xs, ys)):
        num, den = 1, 1
        for j, xj in enumerate(xs):
            if j != i:
                num = num * (x - xj) % P
                den = den * (xi - xj) % P
        total = (total + yi * num * pow(den, P - 2, P)) % P
    return total


def test_erasure_reconstruction():
    # any k of the 2k extended rows determine the rest
    k = 8
    data = _uniform((k, 3), 2)
    out = O.encode(data)
    n_kn = out.shape[0]
    w = int(O.roots_of_unity(n_kn)[1])
    pts = [pow(w, r, P) for r in range(n_kn)]
    rng = np.random.default_rng(3)
    keep = sorted(rng.choice(n_kn, size=k, replace=False).tolist())
    for c in range(3):
        ys = [int(out[r, c]) for r in keep]
        for r in range(n_kn):
            assert _interp_eval([pts[i] for i in keep], ys, pts[r]) == int(out[r, c])


def test_rlc_commutes_with_extension():
    # src/benchmarks/utils.rs:221-272
    k, cols = 16, 40
    data = _uniform((k, cols), 4)
    out = O.encode(data)
    root = O.compute_data_root(out)
    coeffs = O.generate_deterministic_coefficients(root, cols)
    limb_rows = np.array([O.rlc_row(data[r], coeffs) for r in range(k)], dtype=np.uint32)
    ext = O.encode(limb_rows)               # extend each limb column
    for r in range(out.shape[0]):
        assert np.array_equal(O.rlc_row(out[r], coeffs), ext[r])


def test_rlc_coefficients_known_answer():
    # hashlib pins SHA-256; this pins the framing of src/zoda/rlc.rs:48-66
    import hashlib
    import struct
    root = "ab" * 32
    c = O.generate_deterministic_coefficients(root, 2)
    dg = hashlib.sha256(root.encode() + struct.pack("<Q", 1)).digest()
    assert int(c[1, 3]) == int.from_bytes(dg[24:32], "big") % P


def test_planted_mistakes_are_caught(gold):
    """Plausible slips in an oracle fail the goldens: a missing n^-1 scale, the
    forward root used for the inverse, and no bit reversal."""
    case = next(c for c in gold["ntt"] if c["n"] == 16)
    x = np.array(case["input"], dtype=np.uint64)
    good = O.intt_babybear(x).astype(np.uint64)
    unscaled = (good * 16) % P
    assert unscaled.tolist() != case["intt"]
    wrong_root = O.ntt_babybear(x)
    assert wrong_root.tolist() != case["intt"]
    roots = O.roots_of_unity(16)
    no_reverse = O.reverse(O.ntt(O.reverse(x), roots, P))  # DIT on natural-order input
    assert no_reverse.tolist() != case["ntt"]
