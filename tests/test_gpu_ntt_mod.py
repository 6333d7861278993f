"""The reference's generic transforms over other moduli (src/ntt/cpu.rs:1-60)
on the GPU: equal to the oracle's restatement for the reference's own p = 257
test (cpu.rs:112-121), another NTT prime and BabyBear, forward and inverse,
and over strided batches."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zh():
    import zoda_hip
    assert zoda_hip.gpu_available(), "no HIP device"
    return zoda_hip


def test_reference_roundtrip_p257(zh):
    poly = np.array([1, 3, 4, 5, 6, 7, 9, 10], dtype=np.uint32)
    roots = O.roots_of_unity(8, 257)
    fwd = zh.ntt_mod(poly.copy(), roots, 257)
    assert np.array_equal(fwd, O.ntt(poly, roots, 257))
    back = zh.ntt_mod(fwd.copy(), roots, 257, inverse=True)
    assert np.array_equal(back, poly)


@pytest.mark.parametrize("p", [257, 7681, O.P])
@pytest.mark.parametrize("n", [1, 2, 16, 256])
def test_matches_oracle(zh, p, n):
    if (p - 1) % n:
        pytest.skip("no root of unity of that order")
    print(f"[case] ntt_mod p={p} n={n}", flush=True)
    rng = np.random.default_rng(p + n)
    polys, stride = 5, n + 3
    vals = rng.integers(0, p, size=polys * stride, dtype=np.uint32)
    roots = O.roots_of_unity(n, p)
    for inverse in (False, True):
        got = zh.ntt_mod(vals.copy(), roots, p, n=n, stride=stride, inverse=inverse)
        for i in range(polys):
            seg = vals[i * stride:i * stride + n]
            ref = O.intt(seg, roots, p) if inverse else O.ntt(seg, roots, p)
            assert np.array_equal(got[i * stride:i * stride + n], ref), (inverse, i)
            assert np.array_equal(got[i * stride + n:(i + 1) * stride], vals[i * stride + n:(i + 1) * stride])


def test_largest_size_babybear(zh):
    n = 8192
    rng = np.random.default_rng(1)
    vals = rng.integers(0, O.P, size=2 * n, dtype=np.uint32)
    roots = O.roots_of_unity(n, O.P)
    got = zh.ntt_mod(vals.copy(), roots, O.P)
    assert np.array_equal(got[:n], O.ntt(vals[:n], roots, O.P))
    assert np.array_equal(got[n:], O.ntt(vals[n:], roots, O.P))


def test_argument_errors(zh):
    with pytest.raises(zh.ZodaError):
        zh.ntt_mod(np.zeros(12, dtype=np.uint32), np.zeros(12, dtype=np.uint32), 257)   # n not a power of two
    with pytest.raises(zh.ZodaError):
        zh.ntt_mod(np.zeros(8, dtype=np.uint32), np.ones(8, dtype=np.uint32), 1 << 31)  # modulus too large
