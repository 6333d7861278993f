"""CPU-only checks of the product's host side: the C ABI library loads and
exports every entry point include/zoda_hip.h declares, host-only entry points
and argument validation behave, and the encode kernel's per-thread code (run
on the CPU through tests/sim) reproduces the oracle bit for bit."""
import os
import re

import numpy as np
import pytest

import oracle as O
from helpers import ROOT, sim_encode, sim_lib

P = O.P


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "zoda_hip.h")).read()
    return sorted(set(re.findall(r"\b(zoda_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import zoda_hip
    names = _declared_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(zoda_hip.lib, name), name


def test_device_count_without_gpu():
    import ctypes
    import zoda_hip
    c = ctypes.c_int(-1)
    assert zoda_hip.lib.zoda_hip_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0
    assert zoda_hip.lib.zoda_hip_device_count(None) == -1


def test_roots_of_unity_matches_oracle(gold):
    import zoda_hip
    assert list(zoda_hip.roots_of_unity(8)) == gold["roots_8"]
    for n in (1, 2, 1024, 1 << 16):
        assert np.array_equal(zoda_hip.roots_of_unity(n), O.roots_of_unity(n))
    with pytest.raises(zoda_hip.ZodaError) as e:
        zoda_hip.roots_of_unity(12)
    assert e.value.code == -1


def test_argument_errors_do_not_touch_the_device():
    import zoda_hip
    lib = zoda_hip.lib
    buf = np.zeros(64, dtype=np.uint32)
    roots = np.zeros(64, dtype=np.uint32)
    # n not a power of two, stride < n, NULL pointers -> ZODA_EINVAL, no abort
    assert lib.zoda_hip_ntt(buf.ctypes.data, roots.ctypes.data, 3, 1, 3) == -1
    assert "power of two" in zoda_hip.last_error()
    assert lib.zoda_hip_intt(buf.ctypes.data, roots.ctypes.data, 64, 1, 32) == -1
    assert "stride" in zoda_hip.last_error()
    assert lib.zoda_hip_ntt(None, roots.ctypes.data, 8, 1, 8) == -1
    assert lib.zoda_hip_encode_device(None, 4, 4, 4, None, 0, None) == -1
    assert lib.zoda_hip_encode_device(buf.ctypes.data, 0, 4, 4, buf.ctypes.data, 0, None) == -1
    assert lib.zoda_hip_encode_device(buf.ctypes.data, 4, 4, 4, buf.ctypes.data, 7, None) == -1
    assert lib.zoda_hip_encode_host(buf.ctypes.data, 4, 0, 4, buf.ctypes.data, 0) == -1
    # the Python mirror keeps the reference's asserts (src/ntt/cuda.rs:106-108)
    with pytest.raises(AssertionError):
        zoda_hip.ntt_gpu(buf, 6, 6)
    with pytest.raises(AssertionError):
        zoda_hip.ntt_gpu(buf, 64, 32)
    with pytest.raises(AssertionError):
        zoda_hip.ntt_gpu(buf[:60], 8, 16)


def test_runtime_failures_return_the_hip_status():
    """Without a visible device every entry point that reaches the HIP runtime
    returns the runtime's positive hipError_t (hipErrorNoDevice) and a message,
    never a ZODA_E* code and never an abort."""
    import ctypes
    import zoda_hip
    if zoda_hip.gpu_available():
        pytest.skip("a device is visible")
    lib = zoda_hip.lib
    buf = np.zeros(1 << 12, dtype=np.uint32)
    roots = np.ones(64, dtype=np.uint32)
    p, r = buf.ctypes.data, roots.ctypes.data
    rows = (ctypes.c_uint32 * 64)(*range(64))
    calls = {
        "ntt": lambda: lib.zoda_hip_ntt(p, r, 64, 1, 64),
        "intt": lambda: lib.zoda_hip_intt(p, r, 64, 1, 64),
        "ntt_mod": lambda: lib.zoda_hip_ntt_mod(p, r, 64, 1, 64, 257, 0),
        "encode_host": lambda: lib.zoda_hip_encode_host(p, 16, 16, 32, p, 0),
        "encode_device": lambda: lib.zoda_hip_encode_device(p, 16, 16, 32, p, 0, None),
        "ntt_device": lambda: lib.zoda_hip_ntt_device(p, 64, 1, 64, 0, None),
        "fill_uniform": lambda: lib.zoda_hip_fill_uniform(p, 64, 1, None),
        "rlc_coefficients": lambda: lib.zoda_hip_rlc_coefficients(b"ab" * 32, 64, 8, p, None),
        "rlc_rows": lambda: lib.zoda_hip_rlc_rows(p, 8, 8, None, 8, p, p, None),
        "decode_device": lambda: lib.zoda_hip_decode_device(p, 64, 8, rows, 64, p, None),
        "autotune": lambda: lib.zoda_hip_autotune((ctypes.c_uint32 * 1)(64), 1, 1, None),
        "server_create": lambda: lib.zoda_server_create(1 << 30, 2, ctypes.byref(ctypes.c_void_p())),
    }
    for name, call in calls.items():
        rc = call()
        assert rc > 0, (name, rc)
        assert zoda_hip.last_error(), name


def test_kernel_emulations_under_sanitizers():
    """Every kernel emulation (fused plans in both layouts, batched plans,
    persistent and slice kernels, the decoder) rebuilt with AddressSanitizer
    and UndefinedBehaviorSanitizer and run against the oracle: an LDS or global
    index outside its buffer, or undefined integer arithmetic, fails the run."""
    import subprocess
    sim = os.path.join(os.path.dirname(__file__), "sim")
    subprocess.run(["make", "-s", "-C", sim, "sanitize"], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    res = subprocess.run([os.path.join(sim, "sim_check_san")], capture_output=True, text=True, env=env,
                         timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    assert "0 mismatches" in res.stdout


def test_noalloc_rlc_rejects_sizes_beyond_its_scratch():
    """The allocation-free RLC entry point refuses, before touching a device,
    a call whose partial sums exceed its static scratch array."""
    import zoda_hip
    buf = np.zeros(16, dtype=np.uint32)
    p = buf.ctypes.data
    count = (1 << 20) + 1  # rows x column blocks just over the limit (one block)
    rc = zoda_hip.lib.zoda_hip_rlc_rows_noalloc(p, count, 2048, None, count, p, p, None)
    assert rc == -1 and "scratch" in zoda_hip.last_error()


def test_uniform_generator_host_restatement():
    import zoda_hip
    a = zoda_hip.uniform_numpy(100000, 7)
    assert a.max() < P
    assert np.array_equal(a, zoda_hip.uniform_numpy(100000, 7))
    assert not np.array_equal(a, zoda_hip.uniform_numpy(100000, 8))
    # roughly uniform over [0, p)
    hist = np.histogram(a, bins=16, range=(0, P))[0]
    assert hist.min() > 100000 / 16 * 0.9


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 7, 8, 16, 31, 32, 33, 64, 100, 128, 256, 300, 512,
                               1024, 2048, 4096])
def test_kernel_emulation_matches_oracle(k):
    rng = np.random.default_rng(k)
    cols = 37 if k <= 1024 else 9   # ragged against every tile width
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    expect = O.encode(data, threads=4)
    assert np.array_equal(sim_encode(data), expect)
    assert np.array_equal(sim_encode(data, layout=1).T, expect)


@pytest.mark.parametrize("k,n", [(5, 4), (6, 10), (33, 40), (100, 29)])
def test_kernel_emulation_uneven_extension(k, n):
    rng = np.random.default_rng(k * 100 + n)
    data = rng.integers(0, P, size=(k, 11), dtype=np.uint32)
    assert np.array_equal(sim_encode(data, n=n), O.encode(data, n=n))


def test_kernel_emulation_extreme_values():
    k, cols = 512, 40
    for fill in (0, 1, P - 1):
        data = np.full((k, cols), fill, dtype=np.uint32)
        assert np.array_equal(sim_encode(data), O.encode(data, threads=4))


@pytest.mark.parametrize("k,cols", [(1, 256), (16, 256), (32, 512), (64, 128), (256, 64), (512, 64),
                                    (1024, 64), (2048, 32), (4096, 16)])
def test_kernel_emulation_full_tiles(k, cols):
    # whole tiles and k = 2^L select the variant compiled without bounds checks
    rng = np.random.default_rng(k + cols)
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    expect = O.encode(data, threads=4)
    assert np.array_equal(sim_encode(data), expect)
    # the column-major layout of the same whole tiles
    assert np.array_equal(sim_encode(data, layout=1).T, expect)


@pytest.mark.parametrize("k,cols", [(2048, 32), (2048, 64), (2048, 96), (4096, 16), (4096, 32), (4096, 80)])
def test_slice_kernel_emulation(k, cols):
    # the slice-pipelined kernel's per-thread code and order of work (front of
    # tile i beside the back of tile i - 1, held registers handed over slot by
    # slot, input and odd rows staged in LDS; at side 4096 the radix-32 passes
    # split between two threads and per-step buffer bases), run on the host
    from helpers import sim_encode_slice
    rng = np.random.default_rng(cols + k)
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    assert np.array_equal(sim_encode_slice(data), O.encode(data, threads=4))


@pytest.mark.parametrize("k,cols", [(2048, 32), (4096, 16)])
def test_slice_kernel_emulation_extreme_values(k, cols):
    from helpers import sim_encode_slice
    for fill in (0, 1, P - 1):
        data = np.full((k, cols), fill, dtype=np.uint32)
        assert np.array_equal(sim_encode_slice(data), O.encode(data, threads=4))
    data = np.zeros((k, cols), dtype=np.uint32)
    data[::3] = P - 1
    data[1::3] = P - 2
    assert np.array_equal(sim_encode_slice(data), O.encode(data, threads=4))


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("k", [1, 3, 8, 16, 17, 32, 64, 100, 128, 256, 512, 1024, 2048, 4096])
def test_batched_plans_emulation(k, variant):
    # the plans of the persistent batched kernel (every autotuner variant)
    rng = np.random.default_rng(10 * k + variant)
    cols = 70 if k <= 1024 else 12
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    assert np.array_equal(sim_encode(data, variant=variant), O.encode(data, threads=4))


@pytest.mark.parametrize("k,cols", [(256, 320), (512, 160), (1024, 96), (2048, 48), (4096, 24)])
def test_persistent_kernel_emulation(k, cols):
    # first pass from registers loaded while the previous tile was in flight,
    # in parts over the middle passes; even-row stores during the first pass
    from helpers import sim_encode_persistent
    rng = np.random.default_rng(k)
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    assert np.array_equal(sim_encode_persistent(data), O.encode(data, threads=4))


@pytest.mark.parametrize("log_n", list(range(2, 13)))
@pytest.mark.parametrize("inverse", [0, 1])
def test_fp64_poly_transform_emulation(log_n, inverse):
    # The drop-in transform's FP64 digit-pass kernel (ntt_fp64_core.hpp), run
    # on the host phase by phase, against the oracle's ntt / intt
    # (src/ntt/cpu.rs): a ragged last tile, a stride past n, one poly.
    n = 1 << log_n
    rng = np.random.default_rng(log_n * 2 + inverse)
    lib = sim_lib()
    for polys, extra in ((3, 0), (4096 // n + 1, 4), (1, 12)):
        stride = n + extra
        v = rng.integers(0, O.P, size=polys * stride, dtype=np.uint32)
        got = v.copy()
        assert lib.zoda_sim_ntt_fp64(got.ctypes.data, log_n, polys, stride, inverse) == 0
        assert np.array_equal(got, O.ntt_polys(v, n, stride, inverse=bool(inverse))), (polys, stride)
    # extreme inputs: all p - 1 and alternating 0 / p - 1
    for pattern in (np.full(n, O.P - 1, dtype=np.uint32), np.where(np.arange(n) % 2 == 1, O.P - 1, 0).astype(np.uint32)):
        got = pattern.copy()
        assert lib.zoda_sim_ntt_fp64(got.ctypes.data, log_n, 1, n, inverse) == 0
        assert np.array_equal(got, O.ntt_polys(pattern, n, n, inverse=bool(inverse)))


@pytest.mark.parametrize("log_k", list(range(6, 13)))
def test_column_plan_emulation(log_k):
    # The column-major encode (encode_cols_core.hpp) run on the host as the
    # device schedules it: the row-major input transposed block by block into
    # double-buffered chunks, then the column kernel (INTT passes, the twist
    # folded into the last of them, the middle pass, NTT passes landing in
    # natural order) over each chunk's tiles; against the oracle, for one
    # chunk, chunks with a partial last one, and columns of p - 1, zeros and
    # alternating 0 / p - 1.
    from helpers import sim_encode_cols
    k = 1 << log_k
    rng = np.random.default_rng(log_k)
    for cols, chunk in ((64, None), (192, 128), (320, 64)):
        data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
        assert np.array_equal(sim_encode_cols(data, chunk).T, O.encode(data, threads=4)), (cols, chunk)
    ext = np.zeros((k, 64), dtype=np.uint32)
    ext[:, 0::3] = P - 1
    ext[:, 2::3] = np.where(np.arange(k) % 2 == 1, P - 1, 0)[:, None]
    assert np.array_equal(sim_encode_cols(ext).T, O.encode(ext, threads=4))


@pytest.mark.parametrize("tables", [0, 1])
def test_constant_multiplication_edges(tables):
    # the kernels' four-operation multiplication by a constant: exact, congruent
    # to x * w (mod p) and within p/2 + p/256, for |x| < 2^45 at the ends of the
    # range, around multiples of p, and at random, against every kind of w
    import ctypes
    from helpers import sim_lib
    lib = sim_lib()
    rng = np.random.default_rng(7 + tables)
    xs = [0, 1, -1, P - 1, P, P + 1, -P, 2 ** 45 - 1, -(2 ** 45 - 1), 2 ** 44, -(2 ** 44) + 3,
          (2 ** 45 // P) * P, -(2 ** 45 // P) * P + 1]
    xs += rng.integers(-(2 ** 45) + 1, 2 ** 45, size=4000).tolist()
    ws = [0, 1, 2, P - 1, P - 2, (P - 1) // 2, (P + 1) // 2] + rng.integers(0, P, size=200).tolist()
    xa = np.array([x for x in xs for _ in ws], dtype=np.float64)
    wa = np.array([w for _ in xs for w in ws], dtype=np.float64)
    out = np.zeros_like(xa)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.zoda_sim_mulmod(xa.ctypes.data_as(dp), wa.ctypes.data_as(dp), ctypes.c_size_t(xa.size), tables,
                        out.ctypes.data_as(dp))
    assert np.all(out == np.round(out))
    assert np.all(np.abs(out) <= P // 2 + P // 256 + 1)
    xi = xa.astype(np.int64)
    wi = wa.astype(np.int64)
    got = out.astype(np.int64) % P
    expect = ((xi % P) * wi) % P
    assert np.array_equal(got, expect)


def test_canonical_reduction_edges():
    # the kernels' floor-division reduction: every multiple of p in range and
    # its neighbours (j = 0, 1, p - 1 in x = n p + j), and random values over
    # the whole input range |x| < 2^36, against Python's exact modulo
    import ctypes
    from helpers import sim_lib
    lib = sim_lib()
    lib.zoda_sim_canonical.restype = None
    lib.zoda_sim_canonical.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    lim = 1 << 36
    ns = np.arange(-(lim // P) - 1, lim // P + 2, dtype=np.int64)
    xs = np.concatenate([ns * P, ns * P + 1, ns * P - 1, ns * P + (P - 1),
                         np.random.default_rng(3).integers(-lim + 1, lim, size=200000)])
    xs = xs[np.abs(xs) < lim]
    d = xs.astype(np.float64)
    assert np.array_equal(d.astype(np.int64), xs)
    out = np.zeros(xs.size, dtype=np.uint32)
    lib.zoda_sim_canonical(d.ctypes.data, xs.size, out.ctypes.data)
    assert np.array_equal(out.astype(np.int64), np.mod(xs, P))


@pytest.mark.parametrize("k,cols", [(128, 9), (256, 5), (512, 6), (1024, 3), (2048, 2), (4096, 1), (8192, 1)])
def test_decoder_kernel_emulation(k, cols):
    # the FP64 column kernel of the erasure decoder, step by step on the host,
    # from the odd rows, a random half, and the last k rows; ragged last tiles
    from helpers import sim_decode
    rng = np.random.default_rng(k + cols)
    data = rng.integers(0, P, size=(k, cols), dtype=np.uint32)
    data[0, 0] = P - 1
    enc = O.encode(data, threads=4)
    for known in (np.arange(1, 2 * k, 2), rng.choice(2 * k, size=k, replace=False), np.arange(k, 2 * k)):
        assert np.array_equal(sim_decode(enc, rng.permutation(known)), data)


# ---- SHA-256, RLC coefficients and the data root (src/zoda/rlc.rs, utils.rs:51-59) ----

def _sim_sha256(data: bytes) -> bytes:
    import ctypes
    from helpers import sim_lib
    lib = sim_lib()
    out = (ctypes.c_uint8 * 32)()
    buf = ctypes.create_string_buffer(data, len(data) + 1)
    lib.zoda_sim_sha256(buf, ctypes.c_size_t(len(data)), out)
    return bytes(out)


def test_device_sha256_matches_hashlib():
    # the device code's SHA-256, run on the host, against hashlib on every
    # length around the block and padding boundaries
    import hashlib
    rng = np.random.default_rng(3)
    for n in list(range(0, 140)) + [1000, 4097]:
        data = rng.integers(0, 256, size=n, dtype=np.uint8).tobytes()
        assert _sim_sha256(data) == hashlib.sha256(data).digest(), n


@pytest.mark.parametrize("root", ["", "ab", "0" * 63, "f" * 64, "9d" * 50, "e" * 128])
def test_rlc_coefficient_derivation_matches_oracle(root):
    # roots of whole 64-byte blocks ("", 64 and 128 characters) take the
    # shared-state path, the others the whole hash
    import ctypes
    from helpers import sim_lib
    lib = sim_lib()
    expect = O.generate_deterministic_coefficients(root, 300)
    buf = ctypes.create_string_buffer(root.encode(), len(root) + 1)
    out = (ctypes.c_uint32 * 4)()
    for c in [0, 1, 2, 255, 256, 299]:
        lib.zoda_sim_rlc_coefficient(buf, len(root), ctypes.c_uint64(c), out)
        assert list(out) == expect[c].tolist(), c


def test_data_root_matches_oracle():
    import zoda_hip
    rng = np.random.default_rng(4)
    for n in [0, 1, 15, 16, 17, 1000, 70000]:
        rows = rng.integers(0, P, size=n, dtype=np.uint32)
        assert zoda_hip.data_root(rows) == O.compute_data_root(rows), n
