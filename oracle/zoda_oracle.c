/*
 * zoda_oracle.c -- CPU restatement of the reference crate's NTT and ZODA column
 * encode, used ONLY as test infrastructure: the parity tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it as the checker.  Nothing in
 * the product library links or calls this file.
 *
 * Every function follows the reference algorithm step by step (same loop order,
 * same u64 `%` reductions) so that its output is the reference's output:
 *
 *   zo_pow_mod          src/ntt/cpu.rs:62-71
 *   zo_prime_factors    src/ntt/cpu.rs:74-86
 *   zo_find_generator   src/ntt/cpu.rs:89-94
 *   zo_roots_of_unity   src/ntt/cpu.rs:97-110
 *   zo_reverse          src/ntt/cpu.rs:34-46   (recursive even/odd split)
 *   zo_ntt              src/ntt/cpu.rs:1-32    (radix-2 DIT, u64 %)
 *   zo_intt             src/ntt/cpu.rs:48-60   (reversed roots, n^(p-2) scale)
 *   zo_ntt_babybear     src/ntt/cpu.rs:133-141
 *   zo_intt_babybear    src/ntt/cpu.rs:144-152
 *   zo_encode_column    src/benchmarks/utils.rs:206-210 and
 *                       src/zoda/babybear.rs:41-44 (resize n_k, INTT,
 *                       resize n_kn, NTT)
 *   zo_encode           src/benchmarks/utils.rs:100-146 (row-major k x cols
 *                       input; output in row-major or the column-major
 *                       PinnedSquare layout of src/ntt/cuda.rs:31-67)
 *   zo_rlc_row          src/zoda/rlc.rs:32-40
 *
 * Pinning (what fixes each function independently of this restatement) is
 * listed in tests/test_oracle.py: brute-force O(n^2) DFT goldens, the reference
 * test vectors of src/ntt/cpu.rs:112-121,158-187 and src/field/babybear.rs:
 * 84-105, round trips and the ZODA codeword invariants.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ZO_P 2013265921u

/* src/ntt/cpu.rs:62-71 */
uint32_t zo_pow_mod(uint32_t b, uint32_t e, uint32_t p) {
    uint32_t r = 1u;
    b %= p;
    while (e > 0) {
        if (e & 1u) r = (uint32_t)(((uint64_t)r * b) % p);
        b = (uint32_t)(((uint64_t)b * b) % p);
        e >>= 1;
    }
    return r;
}

/* src/ntt/cpu.rs:74-86: distinct prime factors by trial division. Writes at
 * most `cap` factors into out and returns how many there are. */
uint32_t zo_prime_factors(uint32_t m, uint32_t* out, uint32_t cap) {
    uint32_t count = 0;
    uint32_t q = 2;
    while ((uint64_t)q * q <= m) {
        if (m % q == 0) {
            if (count < cap) out[count] = q;
            count++;
            while (m % q == 0) m /= q;
        }
        q++;
    }
    if (m > 1) {
        if (count < cap) out[count] = m;
        count++;
    }
    return count;
}

/* src/ntt/cpu.rs:89-94: smallest g whose order is p-1. */
uint32_t zo_find_generator(uint32_t p) {
    uint32_t f[32];
    uint32_t nf = zo_prime_factors(p - 1, f, 32);
    for (uint32_t g = 2; g < p; g++) {
        int ok = 1;
        for (uint32_t i = 0; i < nf; i++)
            if (zo_pow_mod(g, (p - 1) / f[i], p) == 1) { ok = 0; break; }
        if (ok) return g;
    }
    return 0;
}

/* src/ntt/cpu.rs:97-110: roots[i] = omega^i, omega = g^((p-1)/n).
 * Returns 0 on success, -1 when n does not divide p-1 (the reference asserts). */
int zo_roots_of_unity(uint32_t n, uint32_t p, uint32_t* roots) {
    if (n < 1 || (p - 1) % n != 0) return -1;
    uint32_t g = zo_find_generator(p);
    uint32_t omega = zo_pow_mod(g, (p - 1) / n, p);
    uint32_t cur = 1;
    for (uint32_t i = 0; i < n; i++) {
        roots[i] = cur;
        cur = (uint32_t)(((uint64_t)cur * omega) % p);
    }
    return 0;
}

/* src/ntt/cpu.rs:34-46: evens first, odds second, recursively. `tmp` holds n
 * scratch words. */
static void zo_reverse_rec(uint32_t* v, uint32_t n, uint32_t* tmp) {
    if (n <= 1) return;
    uint32_t h = n / 2;
    for (uint32_t i = 0; i < h; i++) {
        tmp[i] = v[2 * i];
        tmp[h + i] = v[2 * i + 1];
    }
    memcpy(v, tmp, (size_t)n * sizeof(uint32_t));
    zo_reverse_rec(v, h, tmp);
    zo_reverse_rec(v + h, h, tmp);
}

void zo_reverse(uint32_t* values, uint32_t n) {
    if (n <= 1) return;
    uint32_t* tmp = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    zo_reverse_rec(values, n, tmp);
    free(tmp);
}

/* src/ntt/cpu.rs:1-32: in place. roots has n entries (roots[i] = omega^i). */
void zo_ntt(uint32_t* coeffs, const uint32_t* roots, uint32_t n, uint32_t modulus) {
    zo_reverse(coeffs, n);
    for (uint32_t len = 2; len <= n; len *= 2) {
        uint32_t half = len / 2;
        uint32_t stride = n / len;
        for (uint32_t b = 0; b < n; b += len) {
            uint32_t* lo = coeffs + b;
            uint32_t* hi = coeffs + b + half;
            for (uint32_t j = 0; j < half; j++) {
                uint64_t w = roots[(size_t)j * stride];
                uint32_t a = lo[j];
                uint64_t bb = ((uint64_t)hi[j] * w) % modulus;
                lo[j] = (uint32_t)((a + (uint32_t)bb) % modulus);
                hi[j] = (uint32_t)((a + modulus - (uint32_t)bb) % modulus);
            }
        }
    }
}

/* src/ntt/cpu.rs:48-60 */
void zo_intt(uint32_t* values, const uint32_t* roots, uint32_t n, uint32_t modulus) {
    uint32_t* inv = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    inv[0] = roots[0];
    for (uint32_t i = 1; i < n; i++) inv[i] = roots[n - i];
    zo_ntt(values, inv, n, modulus);
    free(inv);
    uint64_t n_inv = zo_pow_mod(n, modulus - 2, modulus);
    for (uint32_t i = 0; i < n; i++) values[i] = (uint32_t)(((uint64_t)values[i] * n_inv) % modulus);
}

/* src/ntt/cpu.rs:133-141: returns -1 if n is not a power of two. */
int zo_ntt_babybear(uint32_t* values, uint32_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    uint32_t* roots = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    zo_roots_of_unity(n, ZO_P, roots);
    zo_ntt(values, roots, n, ZO_P);
    free(roots);
    return 0;
}

/* src/ntt/cpu.rs:144-152 */
int zo_intt_babybear(uint32_t* values, uint32_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    uint32_t* roots = (uint32_t*)malloc((size_t)n * sizeof(uint32_t));
    zo_roots_of_unity(n, ZO_P, roots);
    zo_intt(values, roots, n, ZO_P);
    free(roots);
    return 0;
}

static uint32_t next_pow2(uint32_t x) {
    uint32_t r = 1;
    while (r < x) r <<= 1;
    return r;
}

/* One column: src/benchmarks/utils.rs:206-210.  col_in holds k values; col_out
 * receives n_kn = next_pow2(k+n) values.  Roots for both sizes are passed in
 * so a caller encoding many columns computes them once. */
static void encode_column_with_roots(const uint32_t* col_in, uint32_t k, uint32_t n_k,
                                     uint32_t n_kn, const uint32_t* roots_k,
                                     const uint32_t* roots_kn, uint32_t* col_out) {
    memset(col_out, 0, (size_t)n_kn * sizeof(uint32_t));
    memcpy(col_out, col_in, (size_t)k * sizeof(uint32_t));
    zo_intt(col_out, roots_k, n_k, ZO_P);
    zo_ntt(col_out, roots_kn, n_kn, ZO_P);
}

int zo_encode_column(const uint32_t* col_in, uint32_t k, uint32_t n, uint32_t* col_out) {
    if (k == 0) return -1;
    uint32_t n_k = next_pow2(k), n_kn = next_pow2(k + n);
    uint32_t* rk = (uint32_t*)malloc((size_t)n_k * sizeof(uint32_t));
    uint32_t* rkn = (uint32_t*)malloc((size_t)n_kn * sizeof(uint32_t));
    zo_roots_of_unity(n_k, ZO_P, rk);
    zo_roots_of_unity(n_kn, ZO_P, rkn);
    encode_column_with_roots(col_in, k, n_k, n_kn, rk, rkn, col_out);
    free(rk);
    free(rkn);
    return 0;
}

typedef struct {
    const uint32_t* input;
    uint32_t k, cols, n_k, n_kn;
    const uint32_t *rk, *rkn;
    uint32_t* out;
    int layout;
    uint32_t c0, c1;
} enc_job;

static void* encode_range(void* arg) {
    enc_job* j = (enc_job*)arg;
    uint32_t* cin = (uint32_t*)malloc((size_t)j->k * sizeof(uint32_t));
    uint32_t* cout = (uint32_t*)malloc((size_t)j->n_kn * sizeof(uint32_t));
    for (uint32_t c = j->c0; c < j->c1; c++) {
        for (uint32_t r = 0; r < j->k; r++) cin[r] = j->input[(size_t)r * j->cols + c];
        encode_column_with_roots(cin, j->k, j->n_k, j->n_kn, j->rk, j->rkn, cout);
        if (j->layout == 0) {
            for (uint32_t r = 0; r < j->n_kn; r++) j->out[(size_t)r * j->cols + c] = cout[r];
        } else {
            memcpy(j->out + (size_t)c * j->n_kn, cout, (size_t)j->n_kn * sizeof(uint32_t));
        }
    }
    free(cin);
    free(cout);
    return NULL;
}

/* src/benchmarks/utils.rs:100-146: input is k rows of `cols` values, row-major.
 * Output has n_kn = next_pow2(k+n) rows: layout 0 writes it row-major
 * (out[r*cols+c]), layout 1 column-major (out[c*n_kn+r], the PinnedSquare
 * layout).  `threads` > 1 splits the columns over POSIX threads. */
int zo_encode(const uint32_t* input, uint32_t k, uint32_t cols, uint32_t n, uint32_t* out,
              int layout, int threads) {
    if (k == 0 || cols == 0) return -1;
    if (threads < 1) threads = 1;
    uint32_t n_k = next_pow2(k), n_kn = next_pow2(k + n);
    uint32_t* rk = (uint32_t*)malloc((size_t)n_k * sizeof(uint32_t));
    uint32_t* rkn = (uint32_t*)malloc((size_t)n_kn * sizeof(uint32_t));
    zo_roots_of_unity(n_k, ZO_P, rk);
    zo_roots_of_unity(n_kn, ZO_P, rkn);
    if ((uint32_t)threads > cols) threads = (int)cols;
    enc_job* jobs = (enc_job*)calloc((size_t)threads, sizeof(enc_job));
    pthread_t* tids = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    for (int t = 0; t < threads; t++) {
        enc_job* j = &jobs[t];
        j->input = input; j->k = k; j->cols = cols; j->n_k = n_k; j->n_kn = n_kn;
        j->rk = rk; j->rkn = rkn; j->out = out; j->layout = layout;
        j->c0 = (uint32_t)((uint64_t)cols * t / threads);
        j->c1 = (uint32_t)((uint64_t)cols * (t + 1) / threads);
        if (threads == 1) encode_range(j);
        else pthread_create(&tids[t], NULL, encode_range, j);
    }
    if (threads > 1)
        for (int t = 0; t < threads; t++) pthread_join(tids[t], NULL);
    free(jobs);
    free(tids);
    free(rk);
    free(rkn);
    return 0;
}

/* src/zoda/rlc.rs:32-40: acc[l] = sum_c row[c] * coeff[c][l], BabyBear ops of
 * src/field/babybear.rs:41-78.  coeffs is cols x limbs, row-major. */
void zo_rlc_row(const uint32_t* row, const uint32_t* coeffs, uint32_t cols, uint32_t limbs,
                uint32_t* acc) {
    for (uint32_t l = 0; l < limbs; l++) acc[l] = 0;
    for (uint32_t c = 0; c < cols; c++)
        for (uint32_t l = 0; l < limbs; l++) {
            uint32_t prod = (uint32_t)(((uint64_t)row[c] * coeffs[(size_t)c * limbs + l]) % ZO_P);
            uint32_t s = acc[l] + prod;
            acc[l] = s >= ZO_P ? s - ZO_P : s;
        }
}
