// C++ host-API tests -- the reference crate's GPU tests (src/ntt/cuda.rs:
// 150-237) and its encode validation (src/benchmarks/utils.rs:186-219),
// restated against zoda/zoda.hpp, with the CPU oracle (oracle/liboracle.so)
// as the checker.  Built by __graft_entry__.build(); run by
// tests/test_gpu_cpp.py on a machine with a GPU.  Prints one line per test and exits
// non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../zoda-cu_amd/include/zoda/zoda.hpp"

extern "C" {
int zo_ntt_babybear(uint32_t* values, uint32_t n);
int zo_intt_babybear(uint32_t* values, uint32_t n);
int zo_encode_column(const uint32_t* col_in, uint32_t k, uint32_t n, uint32_t* col_out);
}

#define EXPECT(cond, ...)                                   \
  do {                                                      \
    if (!(cond)) {                                          \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__);      \
      std::printf(__VA_ARGS__);                             \
      std::printf("\n");                                    \
      std::exit(1);                                         \
    }                                                       \
  } while (0)

static std::vector<uint32_t> sample(size_t n) {
  std::vector<uint32_t> v(n);
  for (size_t i = 0; i < n; i++) v[i] = (uint32_t)((i * 7 + 3) % zoda::BABYBEAR_PRIME);
  return v;
}

static void test_ntt_intt_vs_cpu() {
  for (int inverse = 0; inverse < 2; inverse++)
    for (int log_n : {1, 4, 8, 12}) {
      const size_t n = (size_t)1 << log_n;
      std::vector<uint32_t> cpu = sample(n);
      zoda::PinnedSquare gpu(n, 1);
      for (size_t i = 0; i < n; i++) gpu[i] = cpu[i];
      if (inverse) {
        zo_intt_babybear(cpu.data(), (uint32_t)n);
        zoda::intt_gpu(gpu.data(), gpu.size(), n, n);
      } else {
        zo_ntt_babybear(cpu.data(), (uint32_t)n);
        zoda::ntt_gpu(gpu.data(), gpu.size(), n, n);
      }
      for (size_t i = 0; i < n; i++) EXPECT(cpu[i] == gpu[i], "inverse=%d n=%zu: mismatch at %zu", inverse, n, i);
    }
  std::printf("ok test_ntt_intt_vs_cpu\n");
}

static void test_roundtrip() {
  const size_t n = 1 << 10;
  std::vector<uint32_t> orig = sample(n);
  zoda::PinnedSquare v(n, 1);
  for (size_t i = 0; i < n; i++) v[i] = orig[i];
  zoda::ntt_gpu(v.data(), v.size(), n, n);
  zoda::intt_gpu(v.data(), v.size(), n, n);
  for (size_t i = 0; i < n; i++) EXPECT(v[i] == orig[i], "roundtrip failed at %zu", i);
  std::printf("ok test_roundtrip\n");
}

static void test_many_columns_and_strided() {
  {
    const size_t n = 256, cols = 37;
    zoda::PinnedSquare sq(n, cols);
    for (size_t r = 0; r < n; r++)
      for (size_t c = 0; c < cols; c++) sq.set(r, c, (uint32_t)((r * 31 + c * 7 + 1) % zoda::BABYBEAR_PRIME));
    std::vector<std::vector<uint32_t>> cpu(cols);
    for (size_t c = 0; c < cols; c++) cpu[c].assign(sq.column(c), sq.column(c) + n);
    zoda::ntt_gpu(sq.data(), sq.size(), n, n);
    for (size_t c = 0; c < cols; c++) {
      zo_ntt_babybear(cpu[c].data(), (uint32_t)n);
      for (size_t r = 0; r < n; r++) EXPECT(cpu[c][r] == sq.get(r, c), "col=%zu row=%zu", c, r);
    }
  }
  {
    const size_t n = 64, stride = 256, cols = 5;
    zoda::PinnedSquare sq(stride, cols);
    for (size_t c = 0; c < cols; c++)
      for (size_t r = 0; r < stride; r++) sq.set(r, c, (uint32_t)((r * 13 + c * 3 + 2) % zoda::BABYBEAR_PRIME));
    std::vector<uint32_t> before(sq.data(), sq.data() + sq.size());
    zoda::ntt_gpu(sq.data(), sq.size(), n, stride);
    for (size_t c = 0; c < cols; c++) {
      std::vector<uint32_t> cpu(before.begin() + c * stride, before.begin() + c * stride + n);
      zo_ntt_babybear(cpu.data(), (uint32_t)n);
      for (size_t r = 0; r < stride; r++) {
        const uint32_t expect = r < n ? cpu[r] : before[c * stride + r];
        EXPECT(expect == sq.get(r, c), "strided col=%zu row=%zu", c, r);
      }
    }
  }
  std::printf("ok test_many_columns_and_strided\n");
}

static void test_bad_arguments() {
  zoda::PinnedSquare sq(64, 1);
  bool threw = false;
  try { zoda::ntt_gpu(sq.data(), sq.size(), 6, 6); } catch (const std::invalid_argument&) { threw = true; }
  EXPECT(threw, "n = 6 accepted");
  threw = false;
  try { zoda::ntt_gpu(sq.data(), sq.size(), 64, 32); } catch (const std::invalid_argument&) { threw = true; }
  EXPECT(threw, "stride < n accepted");
  EXPECT(zoda_hip_ntt(sq.data(), nullptr, 8, 1, 8) == ZODA_EINVAL, "NULL roots accepted");
  std::printf("ok test_bad_arguments\n");
}

// encode_gpu + the column check of validate_zoda_encoding (utils.rs:186-219),
// on the reference's ramp input (utils.rs:81-84).
static void test_encode_validated(size_t side) {
  zoda::EncodingConfig cfg{side, side, side * 512};
  const size_t cols = cfg.num_positions();
  std::vector<uint32_t> input(cfg.k * cols);
  for (size_t i = 0; i < input.size(); i++) input[i] = (uint32_t)(i % zoda::BABYBEAR_PRIME);
  zoda::PinnedSquare square(cfg.ntt_size_kn(), cols);
  zoda::encode_gpu(cfg, input.data(), square);
  const size_t checks = cols < 64 ? cols : 64;
  std::vector<uint32_t> col_in(cfg.k), col_out(cfg.ntt_size_kn());
  for (size_t i = 0; i < checks; i++) {
    const size_t c = i * cols / checks;
    for (size_t r = 0; r < cfg.k; r++) col_in[r] = (uint32_t)((r * cols + c) % zoda::BABYBEAR_PRIME);
    zo_encode_column(col_in.data(), (uint32_t)cfg.k, (uint32_t)cfg.n, col_out.data());
    for (size_t r = 0; r < cfg.k + cfg.n; r++)
      EXPECT(col_out[r] == square.get(r, c), "side=%zu col=%zu row=%zu", side, c, r);
  }
  std::printf("ok test_encode_validated side=%zu\n", side);
}

// The dense data square against the reference's cell-list semantics: filled
// column by column (as the encoder fills it), in any order of growth, its rows,
// columns and root equal those of the reference square; the root is printed so
// the Python side can check it against hashlib.
static void test_data_square() {
  const size_t cols = 7, rows = 5;
  zoda::DataSquare sq;
  for (size_t c = 0; c < cols; c++)
    for (size_t r = 0; r < rows; r++) sq.set_cell(c, r, (uint32_t)((r * cols + c) * 2654435761u % zoda::BABYBEAR_PRIME));
  EXPECT(sq.columns() == cols && sq.rows() == rows, "shape");
  for (size_t r = 0; r < rows; r++) {
    std::vector<uint32_t> row = sq.get_row(r);
    for (size_t c = 0; c < cols; c++) EXPECT(row[c] == sq.get_cell(c, r), "row %zu col %zu", r, c);
  }
  std::vector<uint32_t> col3 = sq.get_column(3);
  for (size_t r = 0; r < rows; r++) EXPECT(col3[r] == (uint32_t)((r * cols + 3) * 2654435761u % zoda::BABYBEAR_PRIME), "col");
  // built in another order (rows first, growing both ways): same cells, same root
  zoda::DataSquare sq2(1, 1);
  for (size_t r = rows; r-- > 0;)
    for (size_t c = cols; c-- > 0;) sq2.set_cell(c, r, sq.get_cell(c, r));
  EXPECT(sq2.hash_root() == sq.hash_root(), "root depends on the build order");
  std::printf("data_square_root %s\n", sq.hash_root().c_str());
  std::printf("ok test_data_square\n");
}

int main() {
  EXPECT(zoda::gpu_available(), "no HIP device");
  test_ntt_intt_vs_cpu();
  test_roundtrip();
  test_many_columns_and_strided();
  test_bad_arguments();
  test_data_square();
  for (size_t side : {4, 64, 256, 1024}) test_encode_validated(side);
  std::printf("all ok\n");
  return 0;
}
