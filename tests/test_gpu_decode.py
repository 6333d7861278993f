"""GPU erasure decoder: from any k of the 2k rows of an extended square the
decoder returns the data square bit for bit (the inverse of the codeword map
the reference checks at src/zoda/babybear.rs:195-211).  Rows not listed are
filled with garbage first, so a decoder that read them would fail."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def zh():
    import zoda_hip
    assert zoda_hip.gpu_available(), "no HIP device"
    return zoda_hip


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


def patterns(k, rng):
    n = 2 * k
    yield "odd", np.arange(1, n, 2)
    yield "first", np.arange(k)
    yield "last", np.arange(k, n)
    yield "random", rng.choice(n, size=k, replace=False)


@pytest.mark.parametrize("k,cols", [(1, 5), (2, 33), (64, 1000), (128, 77), (256, 4097), (512, 130), (1024, 300), (4096, 64), (8192, 5), (16384, 3)])
def test_decode_from_any_k_rows(zh, torch, k, cols):
    rng = np.random.default_rng(k * 7 + cols)
    t_in = torch.empty((k, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=k + cols)
    enc = zh.encode_tensor(t_in)
    for name, known in patterns(k, rng):
        print(f"[case] decode k={k} cols={cols} rows={name}", flush=True)
        damaged = enc.clone()
        erased = np.setdiff1d(np.arange(2 * k), known)
        if erased.size:
            damaged[torch.from_numpy(erased.astype(np.int64)).cuda()] = -1
        got = zh.decode_tensor(damaged, rng.permutation(known))
        torch.cuda.synchronize()
        assert torch.equal(got, t_in), (k, cols, name)


@pytest.mark.parametrize("k,cols", [(128, 320), (256, 192), (512, 64), (1024, 192), (2048, 128), (4096, 4480),
                                    (8192, 2240)])
def test_decode_chunked_path(zh, torch, k, cols):
    # whole 64-column blocks run the chunked FP64 path (gather, column kernel
    # and scatter of successive chunks in one launch); 4096 x 4480 and
    # 8192 x 2240 span two chunks, the second partial
    rng = np.random.default_rng(k * 11 + cols)
    t_in = torch.empty((k, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=k * 3 + cols)
    enc = zh.encode_tensor(t_in)
    for name, known in patterns(k, rng):
        print(f"[case] chunked decode k={k} cols={cols} rows={name}", flush=True)
        damaged = enc.clone()
        erased = np.setdiff1d(np.arange(2 * k), known)
        damaged[torch.from_numpy(erased.astype(np.int64)).cuda()] = -1
        got = zh.decode_tensor(damaged, rng.permutation(known))
        torch.cuda.synchronize()
        assert torch.equal(got, t_in), (k, cols, name)


def test_decode_2048_square(zh, torch):
    side, cols = 2048, 128 * 2048
    print(f"[case] decode {side}x{cols}", flush=True)
    t_in = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=99)
    enc = zh.encode_tensor(t_in)
    known = np.random.default_rng(5).choice(2 * side, size=side, replace=False)
    erased = np.setdiff1d(np.arange(2 * side), known)
    enc[torch.from_numpy(erased.astype(np.int64)).cuda()] = -1
    got = zh.decode_tensor(enc, known)
    torch.cuda.synchronize()
    assert torch.equal(got, t_in)


def test_decode_argument_errors(zh, torch):
    enc = torch.zeros((16, 8), dtype=torch.int32, device="cuda")
    with pytest.raises(zh.ZodaError) as e:
        zh.decode_tensor(enc, np.arange(7))            # fewer than k distinct rows
    assert e.value.code == -1
    with pytest.raises(zh.ZodaError):
        zh.decode_tensor(enc, np.array([0, 1, 2, 3, 4, 5, 6, 6]))
    with pytest.raises(zh.ZodaError):
        zh.decode_tensor(enc, np.array([0, 1, 2, 3, 4, 5, 6, 16]))   # outside the 16 rows
    with pytest.raises(zh.ZodaError):
        zh.decode_tensor(torch.zeros((12, 8), dtype=torch.int32, device="cuda"), np.arange(6))  # k = 6
    with pytest.raises(zh.ZodaError):
        zh.decode_tensor(torch.zeros((65536, 1), dtype=torch.int32, device="cuda"), np.arange(32768))  # k > 16384
