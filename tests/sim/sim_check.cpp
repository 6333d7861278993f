// Runs every kernel emulation of encode_sim.cpp against the oracle, as one
// program, so that it can be built with AddressSanitizer and
// UndefinedBehaviorSanitizer (see the Makefile's `sanitize` target): an LDS
// or global index outside its buffer in any plan's per-thread code then stops
// the run with a report.  Test infrastructure only.
//
// Exit status 0 when every emulated encode equals the oracle's and every
// emulated decode returns the data; 1 (after a message) otherwise.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

extern "C" {
int zoda_sim_encode_variant(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t n, uint32_t* out, int layout,
                            int variant);
int zoda_sim_encode_persistent(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out);
int zoda_sim_encode_slice(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out);
int zoda_sim_decode(const uint32_t* enc, uint32_t k, uint32_t cols, const uint32_t* known_rows, uint32_t nknown,
                    uint32_t* out);
int zo_encode(const uint32_t* input, uint32_t k, uint32_t cols, uint32_t n, uint32_t* out, int layout, int threads);
int zoda_sim_ntt_fp64(uint32_t* values, uint32_t log_n, uint32_t polys, uint32_t stride, int inverse);
int zo_ntt_babybear(uint32_t* values, uint32_t n);
int zo_intt_babybear(uint32_t* values, uint32_t n);
int zoda_sim_encode_cols(const uint32_t* in, uint32_t k, uint32_t cols, uint32_t* out, uint32_t chunk);
}

namespace {

constexpr uint32_t P = 2013265921u;

uint64_t g_state = 0x9e3779b97f4a7c15ull;

uint32_t next_value() {  // xorshift64*, reduced below p
  g_state ^= g_state >> 12;
  g_state ^= g_state << 25;
  g_state ^= g_state >> 27;
  return (uint32_t)((g_state * 0x2545f4914f6cdd1dull) >> 32) % P;
}

std::vector<uint32_t> square(uint32_t k, uint32_t cols, bool extreme) {
  std::vector<uint32_t> v((size_t)k * cols);
  for (size_t i = 0; i < v.size(); i++) v[i] = extreme ? (i % 3 == 0 ? P - 1 : i % 3 == 1 ? 0 : P - 2) : next_value();
  return v;
}

int g_failures = 0, g_runs = 0;

void compare(const char* what, uint32_t k, uint32_t cols, int layout, int variant,
             const std::vector<uint32_t>& got, const std::vector<uint32_t>& want) {
  g_runs++;
  for (size_t i = 0; i < want.size(); i++)
    if (got[i] != want[i]) {
      printf("MISMATCH %s k=%u cols=%u layout=%d variant=%d at %zu: %u != %u\n", what, k, cols, layout, variant, i,
             got[i], want[i]);
      g_failures++;
      return;
    }
}

// The fused single-square plans (variant -1) and the batched plans (0..3),
// both layouts, ragged and whole-tile widths, systematic extension n = k.
void check_plans() {
  const uint32_t widths[] = {1, 13, 64};
  for (uint32_t L = 0; L <= 12; L++) {
    const uint32_t k = 1u << L;
    for (uint32_t cols : widths) {
      if (L >= 11 && cols > 16) continue;  // keep the largest sides quick
      for (int extreme = 0; extreme < 2; extreme++) {
        if (extreme && cols != 13) continue;
        const std::vector<uint32_t> in = square(k, cols, extreme);
        for (int layout = 0; layout < 2; layout++) {
          std::vector<uint32_t> want((size_t)2 * k * cols), got(want.size());
          zo_encode(in.data(), k, cols, k, want.data(), layout, 1);
          for (int variant = -1; variant <= 3; variant++) {
            std::fill(got.begin(), got.end(), 0xffffffffu);
            if (zoda_sim_encode_variant(in.data(), k, cols, k, got.data(), layout, variant) != 0) continue;
            compare("plan", k, cols, layout, variant, got, want);
          }
        }
      }
    }
  }
}

// Non-systematic extensions n < k (rows n_kn = 2 n_k still).
void check_uneven() {
  const uint32_t cases[][2] = {{5, 3}, {100, 60}, {1000, 24}, {3000, 1096}};
  for (const auto& c : cases) {
    const uint32_t k = c[0], n = c[1], cols = 7;
    uint32_t nkn = 1;
    while (nkn < k + n) nkn <<= 1;
    const std::vector<uint32_t> in = square(k, cols, false);
    for (int layout = 0; layout < 2; layout++) {
      std::vector<uint32_t> want((size_t)nkn * cols), got(want.size());
      zo_encode(in.data(), k, cols, n, want.data(), layout, 1);
      if (zoda_sim_encode_variant(in.data(), k, cols, n, got.data(), layout, -1) != 0) continue;
      compare("uneven", k, cols, layout, -1, got, want);
    }
  }
}

void check_persistent_and_slice() {
  for (uint32_t L = 0; L <= 12; L++) {
    const uint32_t k = 1u << L, cols = 48;
    const std::vector<uint32_t> in = square(k, cols, false);
    std::vector<uint32_t> want((size_t)2 * k * cols), got(want.size());
    if (zoda_sim_encode_persistent(in.data(), k, cols, got.data()) != 0) continue;
    zo_encode(in.data(), k, cols, k, want.data(), 0, 1);
    compare("persistent", k, cols, 0, -1, got, want);
  }
  for (uint32_t kc : {2048u * 1000 + 32, 2048u * 1000 + 96, 4096u * 1000 + 16, 4096u * 1000 + 48}) {
    const uint32_t k = kc / 1000, cols = kc % 1000;
    for (int extreme = 0; extreme < 2; extreme++) {
      const std::vector<uint32_t> in = square(k, cols, extreme);
      std::vector<uint32_t> want((size_t)2 * k * cols), got(want.size());
      zo_encode(in.data(), k, cols, k, want.data(), 0, 1);
      if (zoda_sim_encode_slice(in.data(), k, cols, got.data()) != 0) {
        printf("slice emulation refused k=%u cols=%u\n", k, cols);
        g_failures++;
        continue;
      }
      compare("slice", k, cols, 0, -1, got, want);
    }
  }
}

// Decoder: any k of the 2k rows give the data back (even rows, odd rows, a
// random half).
void check_decoder() {
  for (uint32_t L = 1; L <= 11; L++) {
    const uint32_t k = 1u << L, cols = L >= 10 ? 3 : 9;
    const std::vector<uint32_t> in = square(k, cols, false);
    std::vector<uint32_t> enc((size_t)2 * k * cols);
    zo_encode(in.data(), k, cols, k, enc.data(), 0, 1);
    for (int pattern = 0; pattern < 3; pattern++) {
      std::vector<uint32_t> rows;
      if (pattern < 2) {
        for (uint32_t r = 0; r < k; r++) rows.push_back(2 * r + (uint32_t)pattern);
      } else {
        std::vector<uint32_t> all(2 * k);
        for (uint32_t r = 0; r < 2 * k; r++) all[r] = r;
        for (uint32_t r = 2 * k - 1; r > 0; r--) std::swap(all[r], all[next_value() % (r + 1)]);
        rows.assign(all.begin(), all.begin() + k);
      }
      std::vector<uint32_t> got((size_t)k * cols, 0xffffffffu);
      const int rc = zoda_sim_decode(enc.data(), k, cols, rows.data(), (uint32_t)rows.size(), got.data());
      if (rc == -1) break;  // no FP64 decoder plan for this side
      if (rc != 0) {
        printf("decode emulation failed k=%u pattern=%d (%d)\n", k, pattern, rc);
        g_failures++;
        continue;
      }
      compare("decode", k, cols, 0, pattern, got, in);
    }
  }
}

}  // namespace

// The FP64 poly transform for every size 2^2..2^12, both directions, a
// ragged last tile and a stride past n, against the oracle's ntt / intt.
void check_poly_transforms() {
  for (uint32_t L = 2; L <= 12; L++) {
    const uint32_t n = 1u << L, polys = 4096 / n + 1, stride = n + 4;
    for (int inverse = 0; inverse < 2; inverse++) {
      std::vector<uint32_t> v = square(polys, stride, false), want = v;
      for (uint32_t c = 0; c < polys; c++)
        (inverse ? zo_intt_babybear : zo_ntt_babybear)(want.data() + (size_t)c * stride, n);
      if (zoda_sim_ntt_fp64(v.data(), L, polys, stride, inverse) != 0) {
        printf("FAILED ntt_fp64 L=%u inverse=%d\n", L, inverse);
        g_failures++;
        continue;
      }
      compare(inverse ? "intt_fp64" : "ntt_fp64", L, polys, 0, inverse, v, want);
    }
  }
}

// The column-major encode (encode_cols_core.hpp): chunked transposes and the
// column kernel, sides 2^6 .. 2^12, 64 columns in chunks of 64 and 192
// columns in chunks of 128 (a partial last chunk), random and extreme values.
void check_col_encode() {
  for (uint32_t L = 6; L <= 12; L++) {
    const uint32_t k = 1u << L;
    for (int extreme = 0; extreme < 2; extreme++)
      for (uint32_t cols : {64u, 192u}) {
        const std::vector<uint32_t> in = square(k, cols, extreme != 0);
        std::vector<uint32_t> got((size_t)2 * k * cols), want(got.size());
        zo_encode(in.data(), k, cols, k, want.data(), 1, 1);
        if (zoda_sim_encode_cols(in.data(), k, cols, got.data(), cols == 64 ? 64 : 128) != 0) {
          printf("FAILED encode_cols k=%u\n", k);
          g_failures++;
          continue;
        }
        compare("encode_cols", k, cols, 1, extreme, got, want);
      }
  }
}

int main() {
  check_poly_transforms();
  check_col_encode();
  check_plans();
  check_uneven();
  check_persistent_and_slice();
  check_decoder();
  printf("%d emulated runs, %d mismatches\n", g_runs, g_failures);
  return g_failures == 0 && g_runs > 0 ? 0 : 1;
}
