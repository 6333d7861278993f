"""GPU parity of the fused encode (zoda_hip_encode_device / _host, replacing
encode_gpu of src/benchmarks/utils.rs:100-146) against the oracle.

Whole squares are compared bit for bit up to 1024^2.  At 2048^2 and 4096^2
(the sizes the oracle cannot finish in seconds) the checks are
size-independent: seeded sampled columns against the oracle, the systematic
even rows, a random linear combination of all columns commuting with the
extension (src/benchmarks/utils.rs:221-272), and reconstruction of erased rows
from a random s of the 2s rows."""
import numpy as np
import pytest

import oracle as O
from helpers import lagrange_eval, rlc_rows

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


def to_dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32)).cuda()


def to_host(t):
    return t.cpu().numpy().view(np.uint32)


def gpu_encode(zh, torch, data, n=None, layout=0):
    out = zh.encode_tensor(to_dev(torch, data), n=n, layout=layout)
    torch.cuda.synchronize()
    return to_host(out)


def print_case(name, **kw):
    print(f"[case] {name} {kw}", flush=True)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 8, 16, 17, 32, 64, 100, 128, 256, 512, 1000, 1024, 2048, 4096])
def test_small_widths_all_sides(zh, torch, k):
    rng = np.random.default_rng(k)
    cols = 37 if k <= 1024 else 13
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    print_case("small", k=k, cols=cols)
    expect = O.encode(data, threads=8)
    assert np.array_equal(gpu_encode(zh, torch, data), expect)
    assert np.array_equal(gpu_encode(zh, torch, data, layout=1).T, expect)


@pytest.mark.parametrize("side", [64, 128, 256, 512, 1024])
def test_whole_square_celestia(zh, torch, side):
    # side x side square of 512-byte shares: side rows x 128*side columns
    cols = 128 * side
    data = np.empty((side, cols), dtype=np.uint32)
    t = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t, seed=side)
    torch.cuda.synchronize()
    data[:] = to_host(t)
    print_case("square", side=side)
    expect = O.encode(data, threads=16)
    got = zh.encode_tensor(t)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(got), expect)


def test_reference_ramp_input(zh, torch):
    # src/benchmarks/utils.rs:81-84 input on a 256^2 square, validated the
    # way validate_zoda_encoding does (64 evenly spaced columns, all rows)
    side = 256
    cols = 128 * side
    data = O.ramp_input(side, cols)
    got = gpu_encode(zh, torch, data)
    idx = [(i * cols) // 64 for i in range(64)]
    assert np.array_equal(got[:, idx], O.encode_columns(data, idx))


def test_uniform_generator_matches_host(zh, torch):
    t = torch.empty(1 << 20, dtype=torch.int32, device="cuda")
    zh.fill_uniform(t, seed=1234)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(t), zh.uniform_numpy(1 << 20, 1234))


@pytest.mark.parametrize("k,n", [(5, 4), (6, 10), (100, 29), (1000, 1000), (33, 32), (8, 1), (8, 0),
                                 (3, 20), (700, 2000), (5000, 5000)])
def test_uneven_and_general_shapes(zh, torch, k, n):
    # n_kn = 2 n_k runs the fused kernel; other shapes (n_kn = n_k, n_kn >= 4 n_k,
    # n_k > 4096) run the staged transpose + batched-transform path
    rng = np.random.default_rng(k + 7 * n)
    cols = 21
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    print_case("shape", k=k, n=n)
    expect = O.encode(data, n=n, threads=8)
    assert np.array_equal(gpu_encode(zh, torch, data, n=n), expect)
    assert np.array_equal(gpu_encode(zh, torch, data, n=n, layout=1).T, expect)


def test_edge_values(zh, torch):
    for fill in (0, 1, P - 1):
        data = np.full((2048, 64), fill, dtype=np.uint32)
        assert np.array_equal(gpu_encode(zh, torch, data), O.encode(data, threads=8))


def test_host_end_to_end(zh, torch):
    side = 256
    cols = 128 * side
    rng = np.random.default_rng(3)
    data = rng.integers(0, P, size=(side, cols), dtype=np.uint32)
    expect = O.encode(data, threads=16)
    assert np.array_equal(zh.encode_host(data), expect)
    assert np.array_equal(zh.encode_host(data, layout=1).T, expect)
    # pinned input and output buffers
    pin_in = zh.PinnedSquare(side * cols, 1)
    pin_out = zh.PinnedSquare(2 * side * cols, 1)
    pin_in.array[:] = data.ravel()
    zh.encode_host(pin_in.array.reshape(side, cols), out=pin_out.array.reshape(2 * side, cols))
    assert np.array_equal(pin_out.array.reshape(2 * side, cols), expect)


@pytest.mark.parametrize("side,tiles", [(64, 5), (128, 7), (256, 9), (512, 37), (1024, 13),
                                        (2048, 29), (4096, 11)])
def test_whole_tiles_column_major(zh, torch, side, tiles):
    # cols a multiple of 64 (every plan's tile width divides it) and k = 2^L:
    # the whole-tile instantiations of the column-major layout, the form
    # zoda_hip_encode_host(..., ZODA_LAYOUT_COL) takes for its slabs
    cols = 64 * tiles
    rng = np.random.default_rng(side + tiles)
    data = rng.integers(0, P, size=(side, cols), dtype=np.uint32)
    print_case("whole-tiles-col", side=side, cols=cols)
    expect = O.encode(data, threads=16)
    got = gpu_encode(zh, torch, data, layout=1)
    assert got.shape == (cols, 2 * side)
    assert np.array_equal(got.T, expect)
    # and the row layout of the same whole tiles
    assert np.array_equal(gpu_encode(zh, torch, data), expect)


@pytest.mark.parametrize("tiles", [1, 7, 255, 256, 257, 600])
def test_slice_kernel_tile_counts(zh, torch, tiles):
    # side 2048 in the row layout with whole 32-column tiles runs the
    # slice-pipelined kernel: one persistent work-group per CU, so fewer tiles
    # than CUs, exactly one per CU, one more, and several per CU; every value
    # of the encoded square against the oracle
    side, cols = 2048, 32 * tiles
    rng = np.random.default_rng(tiles)
    data = rng.integers(0, P, size=(side, cols), dtype=np.uint32)
    print_case("slice-tiles", side=side, cols=cols)
    assert np.array_equal(gpu_encode(zh, torch, data), O.encode(data, threads=16))


@pytest.mark.parametrize("tiles", [1, 7, 255, 256, 257, 600])
def test_slice_kernel_tile_counts_4096(zh, torch, tiles):
    # side 4096 in the row layout with whole 16-column tiles runs the
    # slice-pipelined kernel's 4096 plan (radix-32 passes split between two
    # threads, a buffer base per step); every value against the oracle
    side, cols = 4096, 16 * tiles
    rng = np.random.default_rng(tiles + 4096)
    data = rng.integers(0, P, size=(side, cols), dtype=np.uint32)
    print_case("slice-tiles", side=side, cols=cols)
    assert np.array_equal(gpu_encode(zh, torch, data), O.encode(data, threads=16))


@pytest.mark.parametrize("side,cols", [(256, 64 * 1100), (512, 64 * 37), (1024, 64 * 300), (2048, 64),
                                       (2048, 64 * 1000), (2048, 128 * 2048), (4096, 64), (4096, 64 * 300)])
def test_column_kernel_matches_row_layout(zh, torch, side, cols):
    # whole squares with cols a multiple of 64 go into the column-major layout
    # by transpose + column kernel: fewer tiles than work-groups, several per
    # work-group, and the whole 2048^2 square of 512-byte shares; every value
    # against the row-layout encode of the same data, sampled columns against
    # the oracle
    t = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t, seed=side + cols)
    print_case("cols-kernel", side=side, cols=cols)
    col = zh.encode_tensor(t, layout=1)
    row = zh.encode_tensor(t, layout=0)
    torch.cuda.synchronize()
    assert tuple(col.shape) == (cols, 2 * side)
    assert torch.equal(col.t().contiguous(), row)
    idx = sorted({0, 1, cols // 3, cols - 1})
    data = to_host(t[:, idx].contiguous())
    assert np.array_equal(to_host(col[idx, :].t().contiguous()), O.encode(data, threads=4))


def test_column_layout_on_two_streams(zh, torch):
    # two column-major encodes in flight at once on two streams (default
    # priority), each with its own scratch from the library's stream-ordered
    # pool, twice in a row (the second pair reuses the pool's blocks); every
    # value against the row-layout encode of the same data
    shapes = [(1024, 64 * 300), (2048, 64 * 200)]
    ins = []
    for i, (side, cols) in enumerate(shapes):
        t = torch.empty((side, cols), dtype=torch.int32, device="cuda")
        zh.fill_uniform(t, seed=7 + i)
        ins.append(t)
    refs = [zh.encode_tensor(t, layout=0) for t in ins]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rnd in range(2):
        print_case("col-two-streams", round=rnd)
        outs = [torch.empty((cols, 2 * side), dtype=torch.int32, device="cuda") for side, cols in shapes]
        torch.cuda.synchronize()
        for (side, cols), t, o, s in zip(shapes, ins, outs, streams):
            zh.encode_device(t.data_ptr(), side, cols, side, o.data_ptr(), 1, s.cuda_stream)
        torch.cuda.synchronize()
        for o, ref in zip(outs, refs):
            assert torch.equal(o.t().contiguous(), ref)


def test_column_layout_under_graph_capture(zh, torch):
    # captured into a graph, the column-major encode of whole 64-column blocks
    # leaves the chunked path (its scratch comes from the stream-ordered
    # allocator) for the one-kernel form; three replays, new inputs and eager
    # work in between, each against the eager (chunked) encode
    side, cols = 2048, 4096
    s = torch.cuda.current_stream().cuda_stream
    t_in = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    t_out = torch.empty((cols, 2 * side), dtype=torch.int32, device="cuda")
    ref = torch.empty_like(t_out)
    zh.fill_uniform(t_in, seed=1)
    zh.encode_device(t_in.data_ptr(), side, cols, side, ref.data_ptr(), 1, s)
    part = torch.empty((cols - 16, 2 * side), dtype=torch.int32, device="cuda")
    narrow = t_in[:, :cols - 16].contiguous()  # the one-kernel form's set-up, eagerly
    zh.encode_device(narrow.data_ptr(), side, cols - 16, side, part.data_ptr(), 1, s)
    torch.cuda.synchronize()
    print_case("col-graph", side=side, cols=cols)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        zh.encode_device(t_in.data_ptr(), side, cols, side, t_out.data_ptr(), 1,
                         torch.cuda.current_stream().cuda_stream)
    for rep in range(3):
        zh.fill_uniform(t_in, seed=10 + rep)
        t_out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        zh.encode_device(t_in.data_ptr(), side, cols, side, ref.data_ptr(), 1, s)
        torch.cuda.synchronize()
        assert torch.equal(t_out, ref), rep
    data = to_host(t_in[:, [0, cols - 1]].contiguous())
    assert np.array_equal(to_host(t_out[[0, cols - 1], :].t().contiguous()), O.encode(data, threads=2))
    del g


def test_column_major_slab_2048(zh, torch):
    # the exact slab zoda_hip_encode_host encodes at 2048^2 (4096 columns of
    # 2048 rows, whole tiles, column-major output), every value checked
    side, cols = 2048, 4096
    t = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t, seed=77)
    print_case("slab-col", side=side, cols=cols)
    out = zh.encode_tensor(t, layout=1)
    torch.cuda.synchronize()
    data = to_host(t)
    assert np.array_equal(to_host(out).T, O.encode(data, threads=16))


@pytest.mark.parametrize("layout", [0, 1])
def test_host_end_to_end_2048(zh, torch, layout):
    # zoda_hip_encode_host on the whole 2048^2 square from pinned host memory
    # (the drop-in for encode_gpu, src/benchmarks/utils.rs:100-146): seeded
    # sampled columns against the oracle and the systematic rows everywhere
    side = 2048
    cols = 128 * side
    rows_out = 2 * side
    t = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t, seed=4242 + layout)
    pin_in = zh.PinnedSquare(side * cols, 1)
    pin_out = zh.PinnedSquare(rows_out * cols, 1)
    try:
        a_in = pin_in.array.reshape(side, cols)
        a_in[:] = to_host(t)
        del t
        torch.cuda.empty_cache()
        shape = (rows_out, cols) if layout == 0 else (cols, rows_out)
        a_out = pin_out.array.reshape(shape)
        print_case("host-2048", layout=layout)
        zh.encode_host(a_in, layout=layout, out=a_out)
        rng = np.random.default_rng(side + layout)
        idx = sorted(set(rng.choice(cols, size=32, replace=False).tolist()) | {0, 4095, 4096, cols - 1})
        got_s = a_out[:, idx] if layout == 0 else a_out[idx, :].T
        assert np.array_equal(got_s, O.encode(np.ascontiguousarray(a_in[:, idx]), threads=16))
        for c0 in range(0, cols, 8192):
            if layout == 0:
                assert np.array_equal(a_out[0::2, c0:c0 + 8192], a_in[:, c0:c0 + 8192])
            else:
                assert np.array_equal(a_out[c0:c0 + 8192, 0::2], a_in[:, c0:c0 + 8192].T)
    finally:
        pin_in.free()
        pin_out.free()


def big_square_checks(zh, torch, side, seed):
    cols = 128 * side
    t_in = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=seed)
    print_case("big-encode", side=side)
    t_out = zh.encode_tensor(t_in)
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    # 1. sampled columns, bit-exact against the oracle
    cols_idx = sorted(set(rng.choice(cols, size=24, replace=False).tolist()) | {0, cols - 1})
    idx = torch.tensor(cols_idx, device="cuda")
    data_s = to_host(t_in[:, idx].contiguous())
    got_s = to_host(t_out[:, idx].contiguous())
    print_case("big-sampled", side=side, ncols=len(cols_idx))
    assert np.array_equal(got_s, O.encode(data_s, threads=16))
    # 2. systematic even rows equal the data everywhere
    assert torch.equal(t_out[0::2], t_in)
    # 3. RLC: sum_c coeff_c * out[:, c] is the extension of sum_c coeff_c * in[:, c]
    coeffs = rng.integers(0, P, size=(cols, 4), dtype=np.uint32)
    c_dev = torch.from_numpy(coeffs.astype(np.int64)).cuda()

    def rlc_dev(rows_t):
        acc = torch.zeros((rows_t.shape[0], 4), dtype=torch.int64, device="cuda")
        for c0 in range(0, cols, 1 << 16):
            blk = rows_t[:, c0:c0 + (1 << 16)].to(torch.int64) & 0xFFFFFFFF
            for limb in range(4):
                prod = (blk * c_dev[c0:c0 + (1 << 16), limb]) % P
                acc[:, limb] = (acc[:, limb] + prod.sum(dim=1)) % P
        return acc.cpu().numpy().astype(np.uint32)

    print_case("big-rlc", side=side)
    rlc_in = rlc_dev(t_in)                       # side x 4
    ext = O.encode(rlc_in)                        # 2 side x 4: the extended RLC vector
    rows_idx = sorted(rng.choice(2 * side, size=64, replace=False).tolist())
    rlc_out = rlc_dev(t_out[torch.tensor(rows_idx, device="cuda")])
    assert np.array_equal(rlc_out, ext[rows_idx])
    # cross-check the device RLC on the CPU for a few rows
    host_rows = to_host(t_out[torch.tensor(rows_idx[:2], device="cuda")])
    assert np.array_equal(rlc_rows(host_rows, coeffs), rlc_out[:2])
    # 4. erasure: any s of the 2s rows give back the rest (two sampled columns)
    print_case("big-erasure", side=side)
    w = int(O.roots_of_unity(2 * side)[1])
    pts = np.array([O.pow_mod(w, r) for r in range(2 * side)], dtype=np.uint64)
    keep = np.sort(rng.choice(2 * side, size=side, replace=False))
    lost = np.setdiff1d(np.arange(2 * side), keep)[:32]
    for j in (0, len(cols_idx) // 2):
        col = got_s[:, j]
        rec = lagrange_eval(pts[keep], col[keep], pts[lost])
        assert np.array_equal(rec, col[lost])
    del t_in, t_out
    torch.cuda.empty_cache()


def test_big_square_2048(zh, torch):
    big_square_checks(zh, torch, 2048, seed=2048)


def test_big_square_4096(zh, torch):
    big_square_checks(zh, torch, 4096, seed=4096)



@pytest.mark.parametrize("side,cols", [(2048, 262208), (2048, 262224), (1024, 524512)])
def test_raw_buffer_offset_limit(zh, torch, side, cols):
    # Whole tiles of the persistent kernel around the raw-buffer form's limit
    # (((2 side - 1) cols + W) * 4 < 2^32): 2048 x 262208 and 1024 x 524512
    # are inside it (the last output word's offset from its tile's base is
    # 196 bytes and 257 KiB below 2^32); 2048 x 262224 is past it and takes
    # 64-bit pointers.
    t_in = torch.empty((side, cols), dtype=torch.int32, device="cuda")
    zh.fill_uniform(t_in, seed=cols)
    print_case("wide-encode", side=side, cols=cols)
    t_out = zh.encode_tensor(t_in)
    torch.cuda.synchronize()
    rng = np.random.default_rng(cols)
    cols_idx = sorted(set(rng.choice(cols, size=16, replace=False).tolist()) | {0, cols - 1})
    idx = torch.tensor(cols_idx, device="cuda")
    data_s = to_host(t_in[:, idx].contiguous())
    got_s = to_host(t_out[:, idx].contiguous())
    assert np.array_equal(got_s, O.encode(data_s, threads=16))
    assert torch.equal(t_out[0::2], t_in)
    del t_in, t_out
    torch.cuda.empty_cache()
