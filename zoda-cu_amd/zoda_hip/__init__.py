"""Python host side of the MI355X ZODA encoder (ctypes over libzoda_hip.so).

Mirrors the reference crate's GPU API (src/ntt/cuda.rs, src/ntt/mod.rs) under
the same argument meaning and error behaviour:

    gpu_available()                 <- cuda_available()        src/ntt/cuda.rs:25-29
    ntt_gpu(values, n, stride)      <- ntt_cuda                src/ntt/cuda.rs:124-126
    intt_gpu(values, n, stride)     <- intt_cuda               src/ntt/cuda.rs:129-131
    PinnedSquare(rows, cols)        <- PinnedSquare            src/ntt/cuda.rs:34-97
    roots_of_unity(n)               <- roots_of_unity          src/ntt/cpu.rs:97-110

plus the device-resident entry points (encode_device, ntt_device,
fill_uniform) that take raw device pointers, and torch conveniences that pass
a tensor's pointer and the current torch stream.  The library must be built
(`python zoda-cu_amd/build_lib.py`); importing fails loudly when it is not.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

try:
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Loading it first makes libzoda_hip bind to that copy, so torch and this
    # library share one runtime (two copies in one process cannot both open
    # the device).
    import torch  # noqa: F401
except ImportError:  # the C ABI works without torch
    pass

P = 2013265921
LAYOUT_ROW = 0
LAYOUT_COL = 1

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libzoda_hip.so")

_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class ZodaError(RuntimeError):
    """A non-zero status from libzoda_hip (negative: argument or resource
    error; positive: a hipError_t)."""

    def __init__(self, code: int, where: str, message: str):
        super().__init__(f"{where} failed with status {code}: {message}")
        self.code = code


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libzoda_hip.so is not built ({LIB_PATH}); run zoda-cu_amd/build_lib.py")
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "zoda_hip_last_error": (ctypes.c_char_p, []),
        "zoda_hip_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
        "zoda_hip_host_alloc": (_vp, [ctypes.c_size_t]),
        "zoda_hip_host_free": (None, [_vp]),
        "zoda_hip_ntt": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
        "zoda_hip_intt": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
        "zoda_hip_roots_of_unity": (ctypes.c_int, [ctypes.c_uint32, _vp]),
        "zoda_hip_ntt_device": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                               ctypes.c_int, _vp]),
        "zoda_hip_encode_device": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  _vp, ctypes.c_int, _vp]),
        "zoda_hip_encode_host": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                _vp, ctypes.c_int]),
        "zoda_hip_fill_uniform": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_uint64, _vp]),
        "zoda_hip_encode_batch": (ctypes.c_int, [_vp, ctypes.c_uint32, _vp]),
        "zoda_hip_rlc_coefficients": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32, _vp, _vp]),
        "zoda_hip_rlc_rows": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp, _vp,
                                             _vp]),
        "zoda_hip_rlc_rows_noalloc": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32,
                                                     _vp, _vp, _vp]),
        "zoda_data_root": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_char_p]),
        "zoda_hip_ntt_mod": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_int]),
        "zoda_hip_decode_device": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp, ctypes.c_uint32, _vp,
                                                  _vp]),
        "zoda_hip_batch_plan_create": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp)]),
        "zoda_hip_batch_plan_launch": (ctypes.c_int, [_vp, _vp]),
        "zoda_hip_batch_plan_destroy": (ctypes.c_int, [_vp]),
        "zoda_hip_autotune": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_int, _vp]),
        "zoda_hip_get_variant": (ctypes.c_int, [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
        "zoda_hip_set_variant": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int]),
        "zoda_hip_batch_variant_info": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_int, _vp]),
        "zoda_hip_tune_save": (ctypes.c_int, [ctypes.c_char_p]),
        "zoda_hip_autotune_batch": (ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_int, _vp]),
        "zoda_hip_tune_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)]),
        "zoda_server_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(_vp)]),
        "zoda_server_submit": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_uint64)]),
        "zoda_server_poll": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_double)]),
        "zoda_server_wait_all": (ctypes.c_int, [_vp]),
        "zoda_server_output": (ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
        "zoda_server_release": (ctypes.c_int, [_vp, ctypes.c_uint64]),
        "zoda_server_stats": (ctypes.c_int, [_vp, _vp]),
        "zoda_server_destroy": (ctypes.c_int, [_vp]),
        "zoda_hip_serve_trace": (ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                                ctypes.c_int, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    return lib.zoda_hip_last_error().decode()


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise ZodaError(rc, where, last_error())


def _np_ptr(a: np.ndarray) -> int:
    if a.dtype != np.uint32 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("expected a C-contiguous uint32 array")
    return a.ctypes.data


# ---- reference API mirror (src/ntt/cuda.rs) -------------------------------------

def gpu_available() -> bool:
    """cuda_available(): a HIP device is visible."""
    c = ctypes.c_int(0)
    return lib.zoda_hip_device_count(ctypes.byref(c)) == 0 and c.value > 0


def roots_of_unity(n: int) -> np.ndarray:
    """roots[i] = w^i, w = 31^((p-1)/n) (src/ntt/cpu.rs:97-110)."""
    out = np.zeros(max(n, 1), dtype=np.uint32)
    _check(lib.zoda_hip_roots_of_unity(n, _np_ptr(out)), "zoda_hip_roots_of_unity")
    return out[:n]


def _run(values: np.ndarray, n: int, stride: int, inverse: bool, roots=None) -> None:
    # src/ntt/cuda.rs:99-119: the same asserts, then the FFI call
    if n <= 0 or n & (n - 1):
        raise AssertionError("NTT size must be power of 2")
    if stride < n:
        raise AssertionError("stride must be at least n")
    if stride <= 0 or values.size % stride != 0:
        raise AssertionError("values must be whole polynomials of stride elements")
    polys = values.size // stride
    if roots is None:
        roots = roots_of_unity(n)
    roots = np.ascontiguousarray(roots, dtype=np.uint32)
    fn = lib.zoda_hip_intt if inverse else lib.zoda_hip_ntt
    rc = fn(_np_ptr(values), _np_ptr(roots), n, polys, stride)
    name = "zoda_hip_intt" if inverse else "zoda_hip_ntt"
    if rc != 0:
        raise ZodaError(rc, name, last_error())


def ntt_gpu(values: np.ndarray, n: int, stride: int, roots=None) -> None:
    """Forward NTT in place (ntt_cuda, src/ntt/cuda.rs:124-126)."""
    _run(values, n, stride, False, roots)


def intt_gpu(values: np.ndarray, n: int, stride: int, roots=None) -> None:
    """Inverse NTT in place (intt_cuda, src/ntt/cuda.rs:129-131)."""
    _run(values, n, stride, True, roots)


class PinnedSquare:
    """`cols` polynomials of `rows` values in pinned host memory; element
    (row, col) at col*rows + row (src/ntt/cuda.rs:31-97).  `.array` is a
    numpy view of the buffer (zero-filled)."""

    def __init__(self, rows: int, cols: int):
        count = max(rows * cols, 1)
        ptr = lib.zoda_hip_host_alloc(count)
        if not ptr:
            raise MemoryError(f"pinned allocation of {count} u32 failed: {last_error()}")
        self._ptr = ptr
        self.rows, self.cols = rows, cols
        buf = (ctypes.c_uint32 * count).from_address(ptr)
        self._flat = np.frombuffer(buf, dtype=np.uint32, count=count)
        self._flat[:] = 0
        self.array = self._flat[: rows * cols]

    def set(self, row: int, col: int, value: int) -> None:
        self.array[col * self.rows + row] = value

    def get(self, row: int, col: int) -> int:
        return int(self.array[col * self.rows + row]) % P

    def column(self, col: int) -> np.ndarray:
        return self.array[col * self.rows:(col + 1) * self.rows]

    def row(self, row: int) -> np.ndarray:
        return self.array[row::self.rows].copy()

    def free(self) -> None:
        if self._ptr:
            self.array = None
            self._flat = None
            lib.zoda_hip_host_free(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---- device-resident entry points (raw device pointers) -------------------------

def n_kn_for(k: int, n: int) -> int:
    r = 1
    while r < k + n:
        r <<= 1
    return r


def encode_device(d_in: int, k: int, cols: int, n: int, d_out: int, layout: int = LAYOUT_ROW,
                  stream: int = 0) -> None:
    _check(lib.zoda_hip_encode_device(d_in, k, cols, n, d_out, layout, stream or None),
           "zoda_hip_encode_device")


def ntt_device(d_values: int, n: int, polys: int, stride: int, inverse: bool = False,
               stream: int = 0) -> None:
    _check(lib.zoda_hip_ntt_device(d_values, n, polys, stride, 1 if inverse else 0, stream or None),
           "zoda_hip_ntt_device")


def fill_uniform_device(d_out: int, count: int, seed: int, stream: int = 0) -> None:
    _check(lib.zoda_hip_fill_uniform(d_out, count, seed, stream or None), "zoda_hip_fill_uniform")


def encode_host(data: np.ndarray, n: int | None = None, layout: int = LAYOUT_ROW,
                out: np.ndarray | None = None) -> np.ndarray:
    """End-to-end encode of a row-major k x cols host array (pinned or not)."""
    if data.dtype != np.uint32 or data.ndim != 2 or not data.flags["C_CONTIGUOUS"]:
        raise TypeError("data must be a C-contiguous 2-D uint32 array")
    k, cols = data.shape
    if n is None:
        n = k
    rows = n_kn_for(k, n)
    shape = (rows, cols) if layout == LAYOUT_ROW else (cols, rows)
    if out is None:
        out = np.empty(shape, dtype=np.uint32)
    elif out.shape != shape or out.dtype != np.uint32 or not out.flags["C_CONTIGUOUS"]:
        raise TypeError(f"out must be a C-contiguous uint32 array of shape {shape}")
    _check(lib.zoda_hip_encode_host(data.ctypes.data, k, cols, n, out.ctypes.data, layout),
           "zoda_hip_encode_host")
    return out


# ---- torch conveniences -----------------------------------------------------------

def _stream_of(t) -> int:
    import torch
    return torch.cuda.current_stream(t.device).cuda_stream


def encode_tensor(t_in, n: int | None = None, layout: int = LAYOUT_ROW, out=None):
    """Encode a k x cols int32/uint32 CUDA tensor (canonical values); returns
    the n_kn x cols (or cols x n_kn) tensor, on the current torch stream."""
    import torch
    if t_in.dim() != 2 or not t_in.is_cuda or not t_in.is_contiguous() or t_in.element_size() != 4:
        raise TypeError("t_in must be a contiguous 2-D 4-byte CUDA tensor")
    k, cols = t_in.shape
    if n is None:
        n = k
    rows = n_kn_for(k, n)
    shape = (rows, cols) if layout == LAYOUT_ROW else (cols, rows)
    if out is None:
        out = torch.empty(shape, dtype=t_in.dtype, device=t_in.device)
    encode_device(t_in.data_ptr(), k, cols, n, out.data_ptr(), layout, _stream_of(t_in))
    return out


def fill_uniform(t, seed: int) -> None:
    """Fill a contiguous 4-byte CUDA tensor with the seeded uniform generator."""
    if not t.is_cuda or not t.is_contiguous() or t.element_size() != 4:
        raise TypeError("t must be a contiguous 4-byte CUDA tensor")
    fill_uniform_device(t.data_ptr(), t.numel(), seed, _stream_of(t))


def uniform_numpy(count: int, seed: int) -> np.ndarray:
    """Host restatement of the device generator (zoda_hip_fill_uniform)."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    idx = np.arange(count, dtype=np.uint64)
    out = np.zeros(count, dtype=np.uint32)
    todo = np.arange(count)
    with np.errstate(over="ignore"):
        for a in range(64):
            z = np.uint64(seed) + (np.uint64(64) * idx[todo] + np.uint64(a + 1)) * gamma
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            v = (z >> np.uint64(33)).astype(np.uint32)
            ok = v < P
            out[todo[ok]] = v[ok]
            todo = todo[~ok]
            if todo.size == 0:
                break
    return out


# ---- batched encode ------------------------------------------------------------------

class Square(ctypes.Structure):
    """zoda_square: device buffers of one square of a batch."""
    _fields_ = [("inp", ctypes.c_void_p), ("out", ctypes.c_void_p), ("k", ctypes.c_uint32),
                ("cols", ctypes.c_uint32), ("n", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


def _squares(specs):
    arr = (Square * len(specs))()
    for i, (d_in, d_out, k, cols, n) in enumerate(specs):
        arr[i] = Square(d_in, d_out, k, cols, n, 0)
    return arr


def encode_batch(specs, stream: int = 0) -> None:
    """specs: list of (d_in, d_out, k, cols, n) device pointers and shapes."""
    arr = _squares(specs)
    _check(lib.zoda_hip_encode_batch(ctypes.addressof(arr), len(specs), stream or None),
           "zoda_hip_encode_batch")


class BatchPlan:
    """A batch captured into a hipGraph once and replayed with launch()."""

    def __init__(self, specs):
        self._arr = _squares(specs)
        self._plan = ctypes.c_void_p()
        _check(lib.zoda_hip_batch_plan_create(ctypes.addressof(self._arr), len(specs),
                                              ctypes.byref(self._plan)), "zoda_hip_batch_plan_create")

    def launch(self, stream: int = 0) -> None:
        _check(lib.zoda_hip_batch_plan_launch(self._plan, stream or None), "zoda_hip_batch_plan_launch")

    def destroy(self) -> None:
        if self._plan:
            _check(lib.zoda_hip_batch_plan_destroy(self._plan), "zoda_hip_batch_plan_destroy")
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


BATCH_VARIANTS = 4
BATCH_DIRECT = 4          # tuning choice: the side's own fused launch
TUNE_CANDIDATES = BATCH_VARIANTS + 1


def candidate_info(side: int, variant: int) -> dict:
    """Tile width, threads, digit split and twiddle placement of a tuning
    candidate (variant 0..3: a work-queue plan; BATCH_DIRECT: the side's own
    fused kernel)."""
    info = (ctypes.c_int * 6)()
    _check(lib.zoda_hip_batch_variant_info(side.bit_length() - 1, variant, info), "zoda_hip_batch_variant_info")
    return {"kind": "direct" if variant == BATCH_DIRECT else "queue", "W": info[0], "threads": info[1],
            "digits": [info[2], info[3], info[4]], "twiddles": "lds" if info[5] else "global"}


def autotune(sides, iters: int = 3) -> dict:
    """Times every candidate per side (the four work-queue plans and the
    side's own fused launch) and records the fastest in the device's table;
    returns {side: {"choice": c, "times_us": [5 floats], "candidates": [...]}}
    (choice 0..3: queue plan, BATCH_DIRECT: own launch)."""
    sides = list(sides)
    arr = (ctypes.c_uint32 * len(sides))(*sides)
    times = (ctypes.c_float * (TUNE_CANDIDATES * len(sides)))()
    _check(lib.zoda_hip_autotune(ctypes.addressof(arr), len(sides), iters, ctypes.addressof(times)),
           "zoda_hip_autotune")
    out = {}
    for i, side in enumerate(sides):
        out[side] = {"choice": get_variant(side),
                     "times_us": [round(times[TUNE_CANDIDATES * i + v], 2) for v in range(TUNE_CANDIDATES)],
                     "candidates": [candidate_info(side, v) for v in range(TUNE_CANDIDATES)]}
    return out


def autotune_batch(specs, iters: int = 3) -> dict:
    """Tunes the device's table for this batch as it runs (coordinate descent
    over its sides, largest first; see zoda_hip_autotune_batch); returns
    {side: {"choice": c, "times_us": [5 floats], "candidates": [...]}} for the
    sides the batch holds."""
    arr = _squares(specs)
    times = (ctypes.c_float * (13 * TUNE_CANDIDATES))()
    _check(lib.zoda_hip_autotune_batch(ctypes.addressof(arr), len(specs), iters, ctypes.addressof(times)),
           "zoda_hip_autotune_batch")
    out = {}
    for side in sorted({int(sp[2]) for sp in specs}):
        side = 1 << (side - 1).bit_length()
        L = side.bit_length() - 1
        out[side] = {"choice": get_variant(side),
                     "times_us": [round(times[TUNE_CANDIDATES * L + v], 2) for v in range(TUNE_CANDIDATES)],
                     "candidates": [candidate_info(side, v) for v in range(TUNE_CANDIDATES)]}
    return out


def tune_save(path: str) -> None:
    _check(lib.zoda_hip_tune_save(path.encode()), "zoda_hip_tune_save")


def tune_load(path: str) -> int:
    n = ctypes.c_uint32(0)
    _check(lib.zoda_hip_tune_load(path.encode(), ctypes.byref(n)), "zoda_hip_tune_load")
    return n.value


def get_variant(side: int) -> int:
    v = ctypes.c_int(0)
    _check(lib.zoda_hip_get_variant(side.bit_length() - 1, ctypes.byref(v)), "zoda_hip_get_variant")
    return v.value


def set_variant(side: int, variant: int) -> None:
    _check(lib.zoda_hip_set_variant(side.bit_length() - 1, variant), "zoda_hip_set_variant")


# ---- serving --------------------------------------------------------------------------

class ServeStats(ctypes.Structure):
    _fields_ = [("makespan_us", ctypes.c_double), ("batches", ctypes.c_uint32),
                ("max_batch", ctypes.c_uint32), ("peak_bytes", ctypes.c_uint64),
                ("rejected", ctypes.c_uint32), ("pool_peak_bytes", ctypes.c_uint64),
                ("pool_held_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def serve_trace(specs, arrival_us, hbm_budget: int, slots: int = 3, mode: int = 0):
    """Replays a trace (specs as for encode_batch, arrivals in microseconds);
    mode 0 = continuous batching, 1 = plain loop.  Returns (latencies_us, stats)."""
    arr = _squares(specs)
    arrivals = np.ascontiguousarray(arrival_us, dtype=np.float64)
    lat = np.zeros(len(specs), dtype=np.float64)
    st = ServeStats()
    _check(lib.zoda_hip_serve_trace(ctypes.addressof(arr), arrivals.ctypes.data, len(specs), hbm_budget,
                                    slots, mode, lat.ctypes.data, ctypes.addressof(st)),
           "zoda_hip_serve_trace")
    return lat, st.as_dict()


class Server:
    """Online continuous-batching server (background scheduler thread)."""

    def __init__(self, hbm_budget: int, slots: int = 3):
        self._srv = ctypes.c_void_p()
        _check(lib.zoda_server_create(hbm_budget, slots, ctypes.byref(self._srv)), "zoda_server_create")

    def submit(self, d_in: int, d_out: int, k: int, cols: int, n: int) -> int:
        sq = Square(d_in, d_out, k, cols, n, 0)
        t = ctypes.c_uint64(0)
        _check(lib.zoda_server_submit(self._srv, ctypes.addressof(sq), ctypes.byref(t)), "zoda_server_submit")
        return t.value

    def poll(self, ticket: int):
        state = ctypes.c_int(0)
        lat = ctypes.c_double(0)
        _check(lib.zoda_server_poll(self._srv, ticket, ctypes.byref(state), ctypes.byref(lat)),
               "zoda_server_poll")
        return state.value, lat.value

    def output(self, ticket: int) -> int:
        """Device pointer of a finished request's pool output (submitted with
        d_out = 0); valid until release(ticket).  Reads of it must be complete
        (stream synchronized) before the release: the block may go to the
        next request at once."""
        p = _vp()
        _check(lib.zoda_server_output(self._srv, ticket, ctypes.byref(p)), "zoda_server_output")
        return p.value

    def release(self, ticket: int) -> None:
        _check(lib.zoda_server_release(self._srv, ticket), "zoda_server_release")

    def wait_all(self) -> None:
        _check(lib.zoda_server_wait_all(self._srv), "zoda_server_wait_all")

    def stats(self) -> dict:
        st = ServeStats()
        _check(lib.zoda_server_stats(self._srv, ctypes.addressof(st)), "zoda_server_stats")
        return st.as_dict()

    def close(self) -> None:
        if self._srv:
            lib.zoda_server_destroy(self._srv)
            self._srv = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_view(d_ptr: int, rows: int, cols: int):
    """A torch int32 tensor over `rows x cols` u32 of device memory this
    library handed out (e.g. Server.output); it does not own the memory."""
    import torch

    class _Dev:
        __cuda_array_interface__ = {"shape": (rows, cols), "typestr": "<i4", "data": (d_ptr, False),
                                    "version": 3, "strides": None}
    return torch.as_tensor(_Dev(), device="cuda")


# ---- random linear combinations (src/zoda/rlc.rs) ----

def rlc_coefficients(data_root: str, cols: int, out=None):
    """Per-column RLC coefficients from the data root text (rlc.rs:48-66) as a
    (cols, 4) int32 torch tensor on the current device (u32 bit patterns)."""
    import torch
    if out is None:
        out = torch.empty((cols, 4), dtype=torch.int32, device="cuda")
    root = data_root.encode()
    _check(lib.zoda_hip_rlc_coefficients(root, len(root), cols, out.data_ptr(),
                                         torch.cuda.current_stream().cuda_stream),
           "zoda_hip_rlc_coefficients")
    return out


def rlc_rows(square, coeffs, rows=None):
    """rlc_row (rlc.rs:32-40) of rows of a row-major (R, C) int32 device tensor
    (all rows, or the row indices given as a device int32 tensor): an
    (n, 4) int32 tensor of u32 bit patterns."""
    import torch
    assert square.is_contiguous() and square.dim() == 2
    R, C = square.shape
    assert coeffs.shape == (C, 4) and coeffs.is_contiguous()
    count = R if rows is None else rows.numel()
    out = torch.empty((count, 4), dtype=torch.int32, device=square.device)
    _check(lib.zoda_hip_rlc_rows(square.data_ptr(), R, C, None if rows is None else rows.data_ptr(), count,
                                 coeffs.data_ptr(), out.data_ptr(), torch.cuda.current_stream().cuda_stream),
           "zoda_hip_rlc_rows")
    return out


def data_root(values) -> str:
    """compute_data_root (utils.rs:51-59) of host u32 values in row order."""
    a = np.ascontiguousarray(values, dtype=np.uint32)
    buf = ctypes.create_string_buffer(65)
    _check(lib.zoda_data_root(a.ctypes.data, a.size, buf), "zoda_data_root")
    return buf.value.decode()


def decode_tensor(encoded, known_rows, out=None):
    """The k x cols data square back from the 2k x cols extended square
    (int32 device tensor, row-major), using the first k distinct rows listed
    in known_rows; the other rows are never read."""
    import torch
    assert encoded.is_contiguous() and encoded.dim() == 2
    n, cols = encoded.shape
    k = n // 2
    if out is None:
        out = torch.empty((k, cols), dtype=torch.int32, device=encoded.device)
    rows = np.ascontiguousarray(known_rows, dtype=np.uint32)
    _check(lib.zoda_hip_decode_device(encoded.data_ptr(), k, cols, rows.ctypes.data, rows.size, out.data_ptr(),
                                      torch.cuda.current_stream().cuda_stream),
           "zoda_hip_decode_device")
    return out


def ntt_mod(values, roots, modulus: int, n: int | None = None, stride: int | None = None, inverse: bool = False):
    """ntt / intt of src/ntt/cpu.rs over any modulus < 2^31 with the caller's
    roots, on the GPU; transforms the first n of every `stride` values of a
    host uint32 array in place and returns it."""
    v = np.ascontiguousarray(values, dtype=np.uint32)
    r = np.ascontiguousarray(roots, dtype=np.uint32)
    n = len(r) if n is None else n
    stride = n if stride is None else stride
    assert len(r) == n and v.size % stride == 0
    _check(lib.zoda_hip_ntt_mod(v.ctypes.data, r.ctypes.data, n, v.size // stride, stride, modulus,
                                1 if inverse else 0), "zoda_hip_ntt_mod")
    return v
