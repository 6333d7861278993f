"""Shared test helpers: vectorised BabyBear arithmetic in numpy (used by the
size-independent checks: RLC, erasure reconstruction) and the loader of the
CPU emulation of the encode kernel."""
from __future__ import annotations

import ctypes
import os

import numpy as np

P = 2013265921
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def mulmod(a, b):
    return (np.asarray(a, dtype=np.uint64) * np.asarray(b, dtype=np.uint64)) % np.uint64(P)


def powmod(a, e: int):
    a = np.asarray(a, dtype=np.uint64) % np.uint64(P)
    r = np.ones_like(a)
    while e:
        if e & 1:
            r = mulmod(r, a)
        a = mulmod(a, a)
        e >>= 1
    return r


def inv(a):
    return powmod(a, P - 2)


def rlc_rows(rows: np.ndarray, coeffs: np.ndarray) -> np.ndarray:
    """rows: r x cols uint32; coeffs: cols x limbs -> r x limbs, exact mod p."""
    rows = np.asarray(rows, dtype=np.uint64)
    out = np.zeros((rows.shape[0], coeffs.shape[1]), dtype=np.uint64)
    for limb in range(coeffs.shape[1]):
        c = np.asarray(coeffs[:, limb], dtype=np.uint64)
        # chunk columns so partial sums of values < p stay below 2^64
        acc = np.zeros(rows.shape[0], dtype=np.uint64)
        for c0 in range(0, rows.shape[1], 1 << 20):
            prod = (rows[:, c0:c0 + (1 << 20)] * c[c0:c0 + (1 << 20)]) % np.uint64(P)
            acc = (acc + prod.sum(axis=1) % np.uint64(P)) % np.uint64(P)
        out[:, limb] = acc
    return out.astype(np.uint32)


def lagrange_eval(xs, ys, targets):
    """Evaluate the unique polynomial of degree < len(xs) through (xs, ys) at
    `targets`, all mod p (barycentric form, O(len(xs) * (len(xs)+len(targets))))."""
    xs = np.asarray(xs, dtype=np.uint64)
    ys = np.asarray(ys, dtype=np.uint64)
    m = len(xs)
    # w_i = 1 / prod_{j != i} (x_i - x_j)
    w = np.ones(m, dtype=np.uint64)
    for j in range(m):
        d = (xs + np.uint64(P) - xs[j]) % np.uint64(P)
        d[j] = 1
        w = mulmod(w, d)
    w = inv(w)
    out = []
    for t in np.asarray(targets, dtype=np.uint64):
        d = (np.uint64(t) + np.uint64(P) - xs) % np.uint64(P)
        hit = np.nonzero(d == 0)[0]
        if hit.size:
            out.append(int(ys[hit[0]]))
            continue
        ell = np.uint64(1)
        for v in d:
            ell = (ell * v) % np.uint64(P)
        s = mulmod(mulmod(ys, w), inv(d)).sum() % np.uint64(P)
        out.append(int((ell * s) % np.uint64(P)))
    return np.array(out, dtype=np.uint32)


_sim = None


def sim_lib():
    global _sim
    if _sim is None:
        path = os.path.join(ROOT, "tests", "sim", "libencode_sim.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-s", "-C", os.path.dirname(path)], check=True)
        lib = ctypes.CDLL(path)
        u32p = ctypes.c_void_p
        lib.zoda_sim_encode.restype = ctypes.c_int
        lib.zoda_sim_encode.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        u32p, ctypes.c_int]
        lib.zoda_sim_encode_variant.restype = ctypes.c_int
        lib.zoda_sim_encode_variant.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                u32p, ctypes.c_int, ctypes.c_int]
        lib.zoda_sim_sha256.restype = None
        lib.zoda_sim_sha256.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p]
        lib.zoda_sim_rlc_coefficient.restype = None
        lib.zoda_sim_rlc_coefficient.argtypes = [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p]
        lib.zoda_sim_decode.restype = ctypes.c_int
        lib.zoda_sim_decode.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32, u32p]
        lib.zoda_sim_encode_slice.restype = ctypes.c_int
        lib.zoda_sim_encode_slice.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
        lib.zoda_sim_encode_persistent.restype = ctypes.c_int
        lib.zoda_sim_encode_persistent.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
        lib.zoda_sim_ntt_fp64.restype = ctypes.c_int
        lib.zoda_sim_ntt_fp64.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        lib.zoda_sim_encode_cols.restype = ctypes.c_int
        lib.zoda_sim_encode_cols.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_uint32]
        _sim = lib
    return _sim


def sim_encode_persistent(data: np.ndarray) -> np.ndarray:
    k, cols = data.shape
    out = np.zeros((2 * k, cols), dtype=np.uint32)
    d = np.ascontiguousarray(data, dtype=np.uint32)
    rc = sim_lib().zoda_sim_encode_persistent(d.ctypes.data, k, cols, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"no persistent plan for k={k}, cols={cols} (status {rc})")
    return out


def sim_encode_slice(data: np.ndarray) -> np.ndarray:
    """The slice kernel (side 2048, 32-column tiles) run on the host."""
    k, cols = data.shape
    out = np.zeros((2 * k, cols), dtype=np.uint32)
    d = np.ascontiguousarray(data, dtype=np.uint32)
    rc = sim_lib().zoda_sim_encode_slice(d.ctypes.data, k, cols, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"no slice plan for k={k}, cols={cols} (status {rc})")
    return out


def sim_encode_cols(data: np.ndarray, chunk: int | None = None) -> np.ndarray:
    """The column-major encode (chunked transposes + column kernel) run on the
    host: the (cols, 2k) column-major extension of a k x cols square, cols a
    multiple of 64, chunks of `chunk` columns (default: one chunk)."""
    k, cols = data.shape
    out = np.zeros((cols, 2 * k), dtype=np.uint32)
    d = np.ascontiguousarray(data, dtype=np.uint32)
    rc = sim_lib().zoda_sim_encode_cols(d.ctypes.data, k, cols, out.ctypes.data, chunk or cols)
    if rc != 0:
        raise ValueError(f"no column plan for k={k} (status {rc})")
    return out


def sim_encode(data: np.ndarray, n: int | None = None, layout: int = 0, variant: int = -1) -> np.ndarray:
    k, cols = data.shape
    if n is None:
        n = k
    nkn = 1
    while nkn < k + n:
        nkn <<= 1
    out = np.zeros((nkn, cols) if layout == 0 else (cols, nkn), dtype=np.uint32)
    d = np.ascontiguousarray(data, dtype=np.uint32)
    rc = sim_lib().zoda_sim_encode_variant(d.ctypes.data, k, cols, n, out.ctypes.data, layout, variant)
    if rc != 0:
        raise ValueError(f"shape not covered by the fused kernel (k={k}, n={n})")
    return out


def sim_decode(enc: np.ndarray, known_rows) -> np.ndarray:
    """The decoder's FP64 column kernel run on the host (tests/sim): the k x
    cols data from the 2k x cols extended square and its known rows."""
    n, cols = enc.shape
    k = n // 2
    out = np.zeros((k, cols), dtype=np.uint32)
    e = np.ascontiguousarray(enc, dtype=np.uint32)
    rows = np.ascontiguousarray(known_rows, dtype=np.uint32)
    rc = sim_lib().zoda_sim_decode(e.ctypes.data, k, cols, rows.ctypes.data, rows.size, out.ctypes.data)
    if rc != 0:
        raise ValueError(f"no FP64 decoder plan for k={k} (status {rc})")
    return out
