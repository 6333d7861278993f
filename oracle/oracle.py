"""CPU oracle for the ZODA column encode -- TEST INFRASTRUCTURE ONLY.

Only the tests, the smoke check and the CPU-baseline leg of the benchmark may
import this module, and only as the checker.  The product library never routes
through it.

It wraps ``oracle/liboracle.so`` (built from ``oracle/zoda_oracle.c``, a step by
step restatement of the reference's ``src/ntt/cpu.rs``, ``src/zoda/rlc.rs`` and
the encode of ``src/benchmarks/utils.rs:100-146``) and adds the small pure
Python pieces of the reference's validation procedure that are not arithmetic
hot loops: the SHA-256 data root (``src/benchmarks/utils.rs:51-59``) and the
RLC coefficient derivation (``src/zoda/rlc.rs:48-66``).
"""
from __future__ import annotations

import ctypes
import hashlib
import os
import struct
import subprocess

import numpy as np

P = 2013265921
RLC_LIMBS = 4  # src/zoda/rlc.rs:18

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_u32p = ctypes.POINTER(ctypes.c_uint32)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "-B", "liboracle.so"], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.zo_pow_mod.restype = ctypes.c_uint32
    lib.zo_pow_mod.argtypes = [ctypes.c_uint32] * 3
    lib.zo_prime_factors.restype = ctypes.c_uint32
    lib.zo_prime_factors.argtypes = [ctypes.c_uint32, _u32p, ctypes.c_uint32]
    lib.zo_find_generator.restype = ctypes.c_uint32
    lib.zo_find_generator.argtypes = [ctypes.c_uint32]
    lib.zo_roots_of_unity.restype = ctypes.c_int
    lib.zo_roots_of_unity.argtypes = [ctypes.c_uint32, ctypes.c_uint32, _u32p]
    lib.zo_reverse.restype = None
    lib.zo_reverse.argtypes = [_u32p, ctypes.c_uint32]
    lib.zo_ntt.restype = None
    lib.zo_ntt.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32]
    lib.zo_intt.restype = None
    lib.zo_intt.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32]
    lib.zo_ntt_babybear.restype = ctypes.c_int
    lib.zo_ntt_babybear.argtypes = [_u32p, ctypes.c_uint32]
    lib.zo_intt_babybear.restype = ctypes.c_int
    lib.zo_intt_babybear.argtypes = [_u32p, ctypes.c_uint32]
    lib.zo_encode_column.restype = ctypes.c_int
    lib.zo_encode_column.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
    lib.zo_encode.restype = ctypes.c_int
    lib.zo_encode.argtypes = [_u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                              _u32p, ctypes.c_int, ctypes.c_int]
    lib.zo_rlc_row.restype = None
    lib.zo_rlc_row.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
    _lib = lib
    return lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def _u32(values) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(values, dtype=np.uint64).astype(np.uint32))


def next_pow2(x: int) -> int:
    r = 1
    while r < x:
        r <<= 1
    return r


# ---- src/ntt/cpu.rs ---------------------------------------------------------

def pow_mod(b: int, e: int, p: int = P) -> int:
    """src/ntt/cpu.rs:62-71"""
    return int(_load().zo_pow_mod(b, e, p))


def prime_factors(m: int) -> list[int]:
    """src/ntt/cpu.rs:74-86"""
    buf = np.zeros(32, dtype=np.uint32)
    cnt = _load().zo_prime_factors(m, _ptr(buf), 32)
    return [int(x) for x in buf[:cnt]]


def find_generator(p: int = P) -> int:
    """src/ntt/cpu.rs:89-94"""
    return int(_load().zo_find_generator(p))


def roots_of_unity(n: int, p: int = P) -> np.ndarray:
    """src/ntt/cpu.rs:97-110"""
    out = np.zeros(n, dtype=np.uint32)
    if _load().zo_roots_of_unity(n, p, _ptr(out)) != 0:
        raise ValueError("n must divide p - 1")
    return out


def reverse(values) -> np.ndarray:
    """src/ntt/cpu.rs:34-46"""
    v = _u32(values).copy()
    _load().zo_reverse(_ptr(v), len(v))
    return v


def ntt(values, roots, modulus: int) -> np.ndarray:
    """src/ntt/cpu.rs:1-32 (generic modulus)"""
    v = _u32(values).copy()
    r = _u32(roots)
    assert len(v) == len(r)
    _load().zo_ntt(_ptr(v), _ptr(r), len(v), modulus)
    return v


def intt(values, roots, modulus: int) -> np.ndarray:
    """src/ntt/cpu.rs:48-60 (generic modulus)"""
    v = _u32(values).copy()
    r = _u32(roots)
    assert len(v) == len(r)
    _load().zo_intt(_ptr(v), _ptr(r), len(v), modulus)
    return v


def ntt_babybear(values) -> np.ndarray:
    """src/ntt/cpu.rs:133-141"""
    v = _u32(values).copy()
    if _load().zo_ntt_babybear(_ptr(v), len(v)) != 0:
        raise ValueError("NTT size must be power of 2")
    return v


def intt_babybear(values) -> np.ndarray:
    """src/ntt/cpu.rs:144-152"""
    v = _u32(values).copy()
    if _load().zo_intt_babybear(_ptr(v), len(v)) != 0:
        raise ValueError("NTT size must be power of 2")
    return v


def ntt_polys(values: np.ndarray, n: int, stride: int, inverse: bool = False) -> np.ndarray:
    """The reference GPU contract (src/ntt/cuda.rs:99-131) computed with the CPU
    oracle: polynomials `stride` apart, first `n` of each transformed in place,
    rows n..stride untouched."""
    v = _u32(values).copy()
    polys = len(v) // stride
    for c in range(polys):
        seg = v[c * stride:c * stride + n]
        v[c * stride:c * stride + n] = intt_babybear(seg) if inverse else ntt_babybear(seg)
    return v


# ---- encode: src/benchmarks/utils.rs:100-146,206-210 ---------------------------

def encode_column(col, n: int | None = None) -> np.ndarray:
    """One column: resize to n_k, INTT, resize to n_kn, NTT.  Default n = k."""
    c = _u32(col)
    k = len(c)
    if n is None:
        n = k
    out = np.zeros(next_pow2(k + n), dtype=np.uint32)
    _load().zo_encode_column(_ptr(c), k, n, _ptr(out))
    return out


def encode(data: np.ndarray, n: int | None = None, layout: str = "row", threads: int = 1) -> np.ndarray:
    """Row-major k x cols input -> n_kn x cols output (layout "row") or the
    column-major PinnedSquare buffer [cols][n_kn] (layout "col")."""
    d = np.ascontiguousarray(data, dtype=np.uint32)
    k, cols = d.shape
    if n is None:
        n = k
    n_kn = next_pow2(k + n)
    if layout == "row":
        out = np.zeros((n_kn, cols), dtype=np.uint32)
        lay = 0
    else:
        out = np.zeros((cols, n_kn), dtype=np.uint32)
        lay = 1
    rc = _load().zo_encode(_ptr(d), k, cols, n, _ptr(out), lay, threads)
    if rc != 0:
        raise ValueError("bad encode arguments")
    return out


def encode_columns(data: np.ndarray, cols_idx, n: int | None = None) -> np.ndarray:
    """Encode only the listed columns: returns n_kn x len(cols_idx)."""
    d = np.asarray(data, dtype=np.uint32)
    sub = np.ascontiguousarray(d[:, list(cols_idx)])
    return encode(sub, n=n)


# ---- validation helpers: src/benchmarks/utils.rs:51-59, src/zoda/rlc.rs ----------

def compute_data_root(rows: np.ndarray) -> str:
    """src/benchmarks/utils.rs:51-59: SHA-256 over every value, little-endian u32,
    row after row; lowercase hex digest."""
    r = np.ascontiguousarray(rows, dtype="<u4")
    return hashlib.sha256(r.tobytes()).hexdigest()


def generate_deterministic_coefficients(data_root: str, num_columns: int) -> np.ndarray:
    """src/zoda/rlc.rs:48-66: per column, SHA-256(root_hex_bytes || i as u64 LE);
    the four 8-byte windows read big-endian and reduced mod p."""
    out = np.zeros((num_columns, RLC_LIMBS), dtype=np.uint32)
    root = data_root.encode()
    for i in range(num_columns):
        dg = hashlib.sha256(root + struct.pack("<Q", i)).digest()
        for limb in range(RLC_LIMBS):
            out[i, limb] = int.from_bytes(dg[limb * 8:(limb + 1) * 8], "big") % P
    return out


def rlc_row(row, coeffs: np.ndarray) -> np.ndarray:
    """src/zoda/rlc.rs:32-40"""
    r = _u32(row)
    c = np.ascontiguousarray(coeffs, dtype=np.uint32)
    acc = np.zeros(c.shape[1], dtype=np.uint32)
    _load().zo_rlc_row(_ptr(r), _ptr(c), len(r), c.shape[1], _ptr(acc))
    return acc


# ---- inputs: src/benchmarks/utils.rs:81-84 and the seeded uniform generator -------

def ramp_input(k: int, cols: int) -> np.ndarray:
    """src/benchmarks/utils.rs:81-84: element (row, col) = (row*cols+col) mod p."""
    return (np.arange(k * cols, dtype=np.uint64) % P).astype(np.uint32).reshape(k, cols)
