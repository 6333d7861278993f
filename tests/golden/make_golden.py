"""Writes tests/golden/golden.json: known-answer vectors computed by brute force.

Nothing here runs or imports the reference crate (it is Rust and cannot be
built in this image).  Every vector is computed from the mathematics the
reference's own tests assert:

* the field known answers of src/field/babybear.rs:84-105;
* the smallest generator of F_p^* (src/ntt/cpu.rs:89-94) by direct order test;
* NTT outputs as the natural-order DFT X[k] = sum_j x_j w^(jk), w = g^((p-1)/n),
  which is the property src/ntt/cpu.rs:170-187 asserts, evaluated here as an
  O(n^2) sum with Python integers;
* INTT outputs as n^-1 sum_k X_k w^(-jk) (src/ntt/cpu.rs:48-60);
* the generic-modulus case p = 257 of src/ntt/cpu.rs:112-121;
* ZODA column encodes as polynomial evaluation: interpolate the (zero padded)
  data over the size-n_k subgroup, evaluate on the size-n_kn subgroup
  (src/benchmarks/utils.rs:206-210), again by O(n^2) sums.

Run:  python tests/golden/make_golden.py   (a few seconds, pure Python).
"""
import json
import os

P = 2013265921


def prime_factors(m):
    out, q = [], 2
    while q * q <= m:
        if m % q == 0:
            out.append(q)
            while m % q == 0:
                m //= q
        q += 1
    if m > 1:
        out.append(m)
    return out


def smallest_generator(p):
    fs = prime_factors(p - 1)
    g = 2
    while any(pow(g, (p - 1) // q, p) == 1 for q in fs):
        g += 1
    return g


def dft(x, p, inverse=False):
    n = len(x)
    g = smallest_generator(p)
    w = pow(g, (p - 1) // n, p)
    if inverse:
        w = pow(w, p - 2, p)
    out = [sum(x[j] * pow(w, j * k, p) for j in range(n)) % p for k in range(n)]
    if inverse:
        ninv = pow(n, p - 2, p)
        out = [(v * ninv) % p for v in out]
    return out


def next_pow2(x):
    r = 1
    while r < x:
        r <<= 1
    return r


def encode_column(col, n):
    k = len(col)
    n_k, n_kn = next_pow2(k), next_pow2(k + n)
    coeffs = dft(list(col) + [0] * (n_k - k), P, inverse=True)
    g = smallest_generator(P)
    w = pow(g, (P - 1) // n_kn, P)
    return [sum(c * pow(w, i * r, P) for i, c in enumerate(coeffs)) % P for r in range(n_kn)]


def lcg_values(count, seed):
    """Deterministic values spread over [0, p) (64-bit LCG, top 31 bits, mod p)."""
    s, out = seed, []
    for _ in range(count):
        s = (s * 6364136223846793005 + 1442695040888963407) % (1 << 64)
        out.append((s >> 33) % P)
    return out


def main():
    g = {}
    g["p"] = P
    g["generator"] = smallest_generator(P)
    g["generator_257"] = smallest_generator(257)
    # src/field/babybear.rs:84-105
    g["field_kat"] = {
        "add_100_200": (100 + 200) % P,
        "sub_200_100": (200 - 100) % P,
        "sub_100_200": (100 - 200) % P,
        "mul_100_200": (100 * 200) % P,
        "new_p_plus_5": (P + 5) % P,
        "pm1_plus_1": (P - 1 + 1) % P,
        "neg1_squared": ((P - 1) * (P - 1)) % P,
    }
    g["roots_8"] = [pow(pow(g["generator"], (P - 1) // 8, P), i, P) for i in range(8)]
    ntt_cases = []
    for log_n in range(0, 8):
        n = 1 << log_n
        x = [(i * 7 + 3) % P for i in range(n)]          # src/ntt/cuda.rs:138-140
        ntt_cases.append({"n": n, "input": x, "ntt": dft(x, P), "intt": dft(x, P, inverse=True)})
    for n, seed in ((8, 1), (32, 2), (128, 3)):
        x = lcg_values(n, seed)
        ntt_cases.append({"n": n, "input": x, "ntt": dft(x, P), "intt": dft(x, P, inverse=True)})
    g["ntt"] = ntt_cases
    # src/ntt/cpu.rs:170-187 input
    g["naive_eval_8"] = {"input": [(i * 13 + 5) % P for i in range(8)]}
    g["naive_eval_8"]["ntt"] = dft(g["naive_eval_8"]["input"], P)
    # src/ntt/cpu.rs:112-121 (p = 257)
    poly = [1, 3, 4, 5, 6, 7, 9, 10]
    w257 = pow(g["generator_257"], 256 // 8, 257)
    g["p257"] = {"input": poly, "roots": [pow(w257, i, 257) for i in range(8)],
                 "ntt": dft(poly, 257)}
    enc = []
    for k, n, cols, seed in ((1, 1, 2, 5), (2, 2, 3, 6), (4, 4, 3, 7), (8, 8, 5, 8),
                             (16, 16, 2, 9), (3, 3, 2, 10), (5, 3, 2, 11), (6, 10, 2, 12),
                             (32, 32, 1, 13)):
        data = [lcg_values(cols, seed * 100 + r) for r in range(k)]
        cols_out = [encode_column([data[r][c] for r in range(k)], n) for c in range(cols)]
        n_kn = len(cols_out[0])
        out_rows = [[cols_out[c][r] for c in range(cols)] for r in range(n_kn)]
        enc.append({"k": k, "n": n, "cols": cols, "data": data, "out": out_rows})
    # the reference's ramp input (src/benchmarks/utils.rs:81-84) on a 4 x 4 square of 512 B shares
    k = 4
    cols = 4 * 128
    ramp = [[(r * cols + c) % P for c in range(cols)] for r in range(k)]
    sample_cols = [0, 1, 127, 255, 511]
    outs = {str(c): encode_column([ramp[r][c] for r in range(k)], k) for c in sample_cols}
    g["encode"] = enc
    g["ramp_4"] = {"k": k, "cols": cols, "sample_cols": sample_cols, "out_cols": outs}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.json")
    with open(path, "w") as f:
        json.dump(g, f, separators=(",", ":"))
    print("wrote", path)


if __name__ == "__main__":
    main()
