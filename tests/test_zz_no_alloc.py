"""The RLC entry point that allocates nothing (zoda_hip_rlc_rows_noalloc)
against zoda_hip_rlc_rows, through the C ABI: the same results on two calls
in a row, the second smaller, so that the shared static scratch array holds
the first call's partial sums when the second runs."""
import ctypes

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


def rows_both(zh, torch, square, coeffs, index):
    rows, cols = square.shape
    count = rows if index is None else index.numel()
    stream = torch.cuda.current_stream().cuda_stream
    idx = None if index is None else index.data_ptr()
    ref = torch.full((count, 4), -1, dtype=torch.int32, device="cuda")
    got = torch.full((count, 4), -2, dtype=torch.int32, device="cuda")
    rc = zh.lib.zoda_hip_rlc_rows(square.data_ptr(), rows, cols, idx, count, coeffs.data_ptr(), ref.data_ptr(),
                                  stream)
    assert rc == 0, zh.last_error()
    rc = zh.lib.zoda_hip_rlc_rows_noalloc(square.data_ptr(), rows, cols, idx, count, coeffs.data_ptr(),
                                          got.data_ptr(), stream)
    assert rc == 0, zh.last_error()
    torch.cuda.synchronize()
    return ref, got


def test_noalloc_matches_allocating_entry_point(zh, torch):
    cases = [(1024, 20000, 700, 3), (64, 3000, None, 4)]   # the second call is smaller
    for rows, cols, nidx, seed in cases:
        print(f"[case] rlc noalloc rows={rows} cols={cols} index={nidx}", flush=True)
        square = torch.empty((rows, cols), dtype=torch.int32, device="cuda")
        zh.fill_uniform(square, seed=seed)
        coeffs = zh.rlc_coefficients("ab" * 32, cols)
        index = None
        if nidx is not None:
            pick = np.random.default_rng(seed).choice(rows + 8, size=nidx, replace=False).astype(np.int32)
            index = torch.from_numpy(pick).cuda()   # includes rows >= rows: zeros
        ref, got = rows_both(zh, torch, square, coeffs, index)
        assert torch.equal(ref, got)
        assert (ref >= 0).all()


def test_noalloc_size_limit(zh, torch):
    square = torch.zeros((16, 2048), dtype=torch.int32, device="cuda")
    coeffs = zh.rlc_coefficients("00" * 32, 2048)
    out = torch.zeros((4,), dtype=torch.int32, device="cuda")
    count = (1 << 20) + 1                      # one row-block over the scratch limit (one column block)
    index = torch.zeros((count,), dtype=torch.int32, device="cuda")
    rc = zh.lib.zoda_hip_rlc_rows_noalloc(square.data_ptr(), 16, 2048, index.data_ptr(), count,
                                          coeffs.data_ptr(), out.data_ptr(), None)
    assert rc == -1 and "scratch" in zh.last_error()
