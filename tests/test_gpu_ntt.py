"""GPU parity of the drop-in batched NTT/INTT (zoda_hip_ntt / zoda_hip_intt,
replacing cuda_ntt / cuda_intt) against the oracle.  The first five tests are
the reference's own GPU tests (src/ntt/cuda.rs:150-237) with their inputs."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module")
def zh():
    import zoda_hip
    assert zoda_hip.gpu_available(), "no HIP device"
    return zoda_hip


def sample(n):
    return (np.arange(n, dtype=np.uint64) * 7 + 3).astype(np.uint32) % P


def one_poly(zh, values):
    sq = zh.PinnedSquare(len(values), 1)
    sq.array[:] = values
    return sq


@pytest.mark.parametrize("log_n", [1, 4, 8, 12])
def test_ntt_vs_cpu(zh, log_n):
    # src/ntt/cuda.rs:150-164
    n = 1 << log_n
    sq = one_poly(zh, sample(n))
    zh.ntt_gpu(sq.array, n, n)
    assert np.array_equal(sq.array, O.ntt_babybear(sample(n)))


@pytest.mark.parametrize("log_n", [1, 4, 8, 12])
def test_intt_vs_cpu(zh, log_n):
    # src/ntt/cuda.rs:166-180
    n = 1 << log_n
    sq = one_poly(zh, sample(n))
    zh.intt_gpu(sq.array, n, n)
    assert np.array_equal(sq.array, O.intt_babybear(sample(n)))


def test_roundtrip(zh):
    # src/ntt/cuda.rs:182-192
    n = 1 << 10
    sq = one_poly(zh, sample(n))
    zh.ntt_gpu(sq.array, n, n)
    zh.intt_gpu(sq.array, n, n)
    assert np.array_equal(sq.array, sample(n))


def test_many_columns_vs_cpu(zh):
    # src/ntt/cuda.rs:195-214: 37 polynomials of n = 256
    n, cols = 1 << 8, 37
    sq = zh.PinnedSquare(n, cols)
    for col in range(cols):
        sq.array[col * n:(col + 1) * n] = (np.arange(n) * 31 + col * 7 + 1) % P
    before = sq.array.copy()
    zh.ntt_gpu(sq.array, n, n)
    for col in range(cols):
        assert np.array_equal(sq.column(col), O.ntt_babybear(before[col * n:(col + 1) * n])), col


def test_strided_vs_cpu(zh):
    # src/ntt/cuda.rs:216-237: stride 256 > n = 64, rows n..stride untouched
    n, stride, cols = 1 << 6, 1 << 8, 5
    sq = zh.PinnedSquare(stride, cols)
    for col in range(cols):
        sq.array[col * stride:(col + 1) * stride] = (np.arange(stride) * 13 + col * 3 + 2) % P
    before = sq.array.copy()
    zh.ntt_gpu(sq.array, n, stride)
    assert np.array_equal(sq.array, O.ntt_polys(before, n, stride))


@pytest.mark.parametrize("log_n", [0, 2, 13, 15])
def test_sizes_outside_reference_tests(zh, log_n):
    # n = 1 (the reference kernel shifts by 32 there) and sizes past the LDS kernel
    n = 1 << log_n
    polys = 3
    rng = np.random.default_rng(log_n)
    x = rng.integers(0, P, size=polys * n, dtype=np.uint32)
    got = x.copy()
    zh.ntt_gpu(got, n, n)
    assert np.array_equal(got, O.ntt_polys(x, n, n))
    zh.intt_gpu(got, n, n)
    assert np.array_equal(got, x)


def test_caller_roots_are_used(zh):
    # the reference kernel transforms with whatever roots it is given
    n = 16
    w = O.roots_of_unity(n)
    other = np.array([O.pow_mod(int(w[1]), 3 * i) for i in range(n)], dtype=np.uint32)  # w^3 is primitive
    x = sample(n)
    got = x.copy()
    zh.ntt_gpu(got, n, n, roots=other)
    assert np.array_equal(got, O.ntt(x, other, P))
    got2 = x.copy()
    zh.ntt_gpu(got2, n, n)
    assert np.array_equal(got2, O.ntt_babybear(x))


def test_many_polys_chunked(zh):
    # more than one 128 MiB chunk, pageable memory, stride > n
    n, stride = 1 << 11, (1 << 11) + 5
    polys = (160 << 20) // (n * 4)
    rng = np.random.default_rng(5)
    x = rng.integers(0, P, size=polys * stride, dtype=np.uint32)
    got = x.copy()
    zh.intt_gpu(got, n, stride)
    sel = [0, 1, polys // 2, polys - 1]
    for p in sel:
        assert np.array_equal(got[p * stride:p * stride + n], O.intt_babybear(x[p * stride:p * stride + n]))
        assert np.array_equal(got[p * stride + n:(p + 1) * stride], x[p * stride + n:(p + 1) * stride])


def test_device_entry_point(zh):
    import torch
    n, polys, stride = 1 << 9, 100, 600
    rng = np.random.default_rng(9)
    x = rng.integers(0, P, size=polys * stride, dtype=np.uint32)
    t = torch.from_numpy(x.view(np.int32)).cuda()
    zh.ntt_device(t.data_ptr(), n, polys, stride, inverse=False,
                  stream=torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = t.cpu().numpy().view(np.uint32)
    assert np.array_equal(got, O.ntt_polys(x, n, stride))
