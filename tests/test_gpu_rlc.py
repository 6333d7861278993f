"""GPU parity of the random linear combinations (src/zoda/rlc.rs:32-66) and
the commitment check they serve (src/benchmarks/utils.rs:221-272): the
device-derived coefficients equal the oracle's SHA-256 derivation, the device
row RLC equals the oracle's rlc_row, and on the 2048^2 and 4096^2 squares the
RLC of every extended row equals the extension of the original rows' RLC (the
two are F_p-linear and commute)."""
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


@pytest.fixture(scope="module")
def torch():
    import torch
    return torch


def u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("root", ["", "c0ffee", "a" * 64, "3f" * 40, "e" * 128])
def test_coefficients_match_oracle(zh, torch, root):
    print(f"[case] rlc coefficients root_len={len(root)}", flush=True)
    cols = 3000
    got = u32(zh.rlc_coefficients(root, cols))
    assert np.array_equal(got, O.generate_deterministic_coefficients(root, cols))


@pytest.mark.parametrize("rows,cols", [(1, 1), (37, 5000), (16, 4096), (17, 4097), (5, 70000)])
def test_rlc_rows_match_oracle(zh, torch, rows, cols):
    print(f"[case] rlc rows {rows}x{cols}", flush=True)
    rng = np.random.default_rng(rows * 100003 + cols)
    sq = rng.integers(0, P, size=(rows, cols), dtype=np.uint32)
    # extreme values in the first row: every product near (p-1)^2
    sq[0, :] = P - 1
    coeffs = zh.rlc_coefficients("f" * 64, cols)
    coeffs[: min(cols, 3)] = P - 1
    t_sq = torch.from_numpy(sq.view(np.int32)).cuda()
    got = u32(zh.rlc_rows(t_sq, coeffs))
    c = u32(coeffs)
    for r in range(rows):
        assert np.array_equal(got[r], O.rlc_row(sq[r], c)), r
    # a row subset, repeated rows and an index past the last row (zeros)
    idx = np.array([rows - 1, 0, 0, rows + 5], dtype=np.int32)
    got = u32(zh.rlc_rows(t_sq, coeffs, torch.from_numpy(idx).cuda()))
    assert np.array_equal(got[0], O.rlc_row(sq[rows - 1], c))
    assert np.array_equal(got[1], O.rlc_row(sq[0], c)) and np.array_equal(got[2], got[1])
    assert not got[3].any()


@pytest.mark.parametrize("rows,cols", [(70, 4196), (3, 2048)])
def test_rlc_rows_any_u32(zh, torch, rows, cols):
    # rlc_row is defined for any u32 values and coefficients (each reduced
    # mod p in the product), not only canonical ones: full-range words,
    # including 2^32 - 1 and values just above p and 2p.
    print(f"[case] rlc rows any-u32 {rows}x{cols}", flush=True)
    rng = np.random.default_rng(rows * 7 + cols)
    sq = rng.integers(0, 2**32, size=(rows, cols), dtype=np.uint64).astype(np.uint32)
    sq[0, :4] = [2**32 - 1, P, 2 * P, 2 * P - 1]
    c = rng.integers(0, 2**32, size=(cols, 4), dtype=np.uint64).astype(np.uint32)
    c[0] = [2**32 - 1, P, 2 * P, 3 * P - 1 - 2**31]
    t_sq = torch.from_numpy(sq.view(np.int32)).cuda()
    coeffs = torch.from_numpy(c.view(np.int32)).cuda()
    got = u32(zh.rlc_rows(t_sq, coeffs))
    for r in range(rows):
        assert np.array_equal(got[r], O.rlc_row(sq[r], c)), r


def test_rlc_rows_unaligned_square(zh, torch):
    # a square starting 4 bytes past a 16-byte boundary, with whole 2048-column
    # blocks: the kernel must not take its 16-byte load path there
    rows, cols = 20, 4096
    print(f"[case] rlc rows unaligned {rows}x{cols}", flush=True)
    rng = np.random.default_rng(4097)
    sq = rng.integers(0, P, size=(rows, cols), dtype=np.uint32)
    flat = torch.empty(rows * cols + 1, dtype=torch.int32, device="cuda")
    t_sq = flat[1:].view(rows, cols)
    assert t_sq.data_ptr() % 16 == 4
    t_sq.copy_(torch.from_numpy(sq.view(np.int32)))
    coeffs = zh.rlc_coefficients("9" * 64, cols)
    got = u32(zh.rlc_rows(t_sq, coeffs))
    c = u32(coeffs)
    for r in range(rows):
        assert np.array_equal(got[r], O.rlc_row(sq[r], c)), r


@pytest.mark.parametrize("side", [2048, 4096])
def test_commutes_with_the_extension(zh, torch, side):
    # RLC(extended row i) = sum_j G[i][j] RLC(original row j) for every row of
    # the 2048^2 (4096 rows) and 4096^2 (8192 rows) squares: each limb of the
    # original rows' RLC, extended by the oracle (INTT-s, pad, NTT-2s), must
    # equal the RLC of the device-encoded rows.  A wrong value anywhere in the
    # output changes its row's RLC, so this checks every value the encode
    # writes.  Three encodes into a zeroed output: a fault that depends on
    # timing (such as registers of a store overwritten before the store has
    # read them) shows as a difference in any one of them.
    cols = 128 * side
    print(f"[case] rlc commute {side}x{cols}", flush=True)
    t_in = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=side)
    coeffs = zh.rlc_coefficients("ab" * 32, cols)
    orig = u32(zh.rlc_rows(t_in, coeffs))          # (side, 4)
    expect = np.stack([O.encode(np.ascontiguousarray(orig[:, limb]).reshape(side, 1), threads=4).ravel()
                       for limb in range(4)], axis=1)
    t_out = torch.empty((2 * side, cols), dtype=torch.int32, device="cuda")
    for rep in range(3):
        t_out.zero_()
        zh.encode_tensor(t_in, out=t_out)
        ext = u32(zh.rlc_rows(t_out, coeffs))      # (2 side, 4)
        bad = np.nonzero((ext != expect).any(axis=1))[0]
        assert bad.size == 0, f"encode {rep}: {bad.size} rows differ, e.g. {bad[:8].tolist()}"
        # and the systematic rows: the even extended rows are the original rows
        assert np.array_equal(ext[0::2], orig)
    del t_in, t_out
    torch.cuda.empty_cache()
