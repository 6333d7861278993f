"""CPU check of the erasure decoder's schedule (zoda-cu_amd/csrc/decode.hip):
the same steps in the same order and with the same index conventions as the
device kernels -- the vanishing polynomial on the known rows, its inverse on
the coset in bit-reversed order, radix-2 decimation in time (bit-reversed in,
natural out) and in frequency (natural in, bit-reversed out) with the
tw[h + j] = w_(2h)^j tables, the scalings by h^i / n and h^-i / n -- run on
small squares encoded by the oracle (src/ntt/cpu.rs transforms, zero-padded
to 2k as in src/zoda/babybear.rs:195-211) must give the data back for every
erasure pattern.  A wrong twiddle direction, a missing bit reversal or a
scaling on the wrong half fails here.  That is the schedule of the sizes
without an FP64 plan (dec_fused); the FP64 path's identity (erased rows from
the formal derivative of g = f Z, one transform pair) is checked the same way
with the oracle's transforms."""
import numpy as np
import pytest

import oracle as O

P = O.P
H = 31


def pw(a, e):
    return pow(int(a), int(e), P)


def brev(i, log_n):
    return int(format(i, f"0{log_n}b")[::-1], 2) if log_n else 0


def tables(log_n, inverse):
    n = 1 << log_n
    w = pw(31, (P - 1) >> log_n)
    if inverse:
        w = pw(w, P - 2)
    t = [0] * max(n, 2)
    h = 1
    while h < n:
        for j in range(h):
            t[h + j] = pw(w, j * (n // (2 * h)))
        h <<= 1
    return t


def dit(v, log_n, tw):
    for lh in range(log_n):
        h = 1 << lh
        for b in range(len(v) // 2):
            j = b & (h - 1)
            base = ((b >> lh) << (lh + 1)) + j
            u, w = v[base], v[base + h] * tw[h + j] % P
            v[base], v[base + h] = (u + w) % P, (u - w) % P


def dif(v, log_n, tw):
    for lh in reversed(range(log_n)):
        h = 1 << lh
        for b in range(len(v) // 2):
            j = b & (h - 1)
            base = ((b >> lh) << (lh + 1)) + j
            u, w = v[base], v[base + h]
            v[base], v[base + h] = (u + w) % P, (u - w) * tw[h + j] % P


def decode_column(col, known, k):
    n = 2 * k
    log_n = n.bit_length() - 1
    w = pw(31, (P - 1) // n)
    x = [pw(w, r) for r in range(n)]
    erased = [x[r] for r in range(n) if r not in known]
    zk = [0] * n
    zinv_br = [0] * n
    for t in range(2 * n):                      # dec_prep
        coset = t >= n
        r = brev(t - n, log_n) if coset else t
        if not coset and r not in known:
            continue
        pt = H * x[r] % P if coset else x[r]
        acc = 1
        for e in erased:
            acc = acc * ((pt - e) % P) % P
        if coset:
            zinv_br[t - n] = pw(acc, P - 2)
        else:
            zk[r] = acc
    y = [int(col[r]) * zk[r] % P for r in range(n)]   # dec_gather
    n_inv, h_inv = pw(n, P - 2), pw(H, P - 2)
    v = [0] * n
    for i in range(n):                          # dec_fused
        v[brev(i, log_n)] = y[i]
    dit(v, log_n, tables(log_n, True))
    v = [v[i] * pw(H, i) * n_inv % P for i in range(n)]
    dif(v, log_n, tables(log_n, False))
    v = [v[i] * zinv_br[i] % P for i in range(n)]
    dit(v, log_n, tables(log_n, True))
    assert all(val == 0 for val in v[k:]), "f has degree >= k"
    low = [v[i] * pw(h_inv, i) * n_inv % P for i in range(k)]
    dif(low, log_n - 1, tables(log_n - 1, False))
    return [low[brev(j, log_n - 1)] for j in range(k)]


@pytest.mark.parametrize("k", [1, 2, 4, 16])
def test_schedule_recovers_data(k):
    n = 2 * k
    rng = np.random.default_rng(k)
    data = rng.integers(0, P, size=(k, 2), dtype=np.uint32)
    data[0, 0] = P - 1
    enc = O.encode(data, threads=1)
    patterns = [set(range(1, n, 2)), set(range(k)), set(range(k, n))]
    patterns += [set(rng.choice(n, size=k, replace=False).tolist()) for _ in range(3)]
    for known in patterns:
        for c in range(data.shape[1]):
            assert decode_column(enc[:, c], known, k) == [int(d) for d in data[:, c]], (k, sorted(known), c)


def derivative_decode_column(col, known, k):
    """The FP64 path's identity on one column, with the oracle's transforms
    (src/ntt/cpu.rs: ntt, intt with its 1/n): g = f Z has the values y, so
    n w^r g'(w^r) = NTT_n(i * n G_i) and, where Z(w^e) = 0,
    f(w^e) = w^e g'(w^e) / (w^e Z'(w^e)); known data rows are copied."""
    n = 2 * k
    roots = O.roots_of_unity(n)
    x = [int(roots[r]) for r in range(n)]
    erased = [r for r in range(n) if r not in known]

    def z_at(pt):
        acc = 1
        for e in erased:
            acc = acc * (pt - x[e]) % P
        return acc

    y = [int(col[r]) * z_at(x[r]) % P if r in known else 0 for r in range(n)]
    g = O.intt(np.array(y, dtype=np.uint32), roots, P)        # coefficients of g
    dg = O.ntt(np.array([i * int(g[i]) % P for i in range(n)], dtype=np.uint32), roots, P)  # w^r g'(w^r)
    out = []
    for j in range(k):
        r = 2 * j
        if r in known:
            out.append(int(col[r]))
            continue
        zd = 1
        for e in erased:
            if e != r:
                zd = zd * (x[r] - x[e]) % P
        out.append(int(dg[r]) * pw(x[r] * zd % P, P - 2) % P)
    return out


@pytest.mark.parametrize("k", [1, 2, 4, 16, 64])
def test_derivative_identity_recovers_data(k):
    # the decoder's FP64 path: one transform pair and the formal derivative
    n = 2 * k
    rng = np.random.default_rng(100 + k)
    data = rng.integers(0, P, size=(k, 2), dtype=np.uint32)
    data[0, 0] = P - 1
    enc = O.encode(data, threads=1)
    patterns = [set(range(1, n, 2)), set(range(k)), set(range(k, n))]
    patterns += [set(rng.choice(n, size=k, replace=False).tolist()) for _ in range(3)]
    for known in patterns:
        for c in range(data.shape[1]):
            assert derivative_decode_column(enc[:, c], known, k) == [int(d) for d in data[:, c]], (k, sorted(known), c)
