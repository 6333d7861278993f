// C++ host API of the MI355X ZODA encoder, over the C ABI of zoda_hip.h.
//
// Mirrors the reference crate's GPU-facing Rust API with the same argument
// meaning and checks:
//   gpu_available()            src/ntt/cuda.rs:25-29   (cuda_available)
//   roots_of_unity(n)          src/ntt/cpu.rs:97-110
//   ntt_gpu / intt_gpu         src/ntt/cuda.rs:99-131  (ntt_cuda / intt_cuda)
//   PinnedSquare               src/ntt/cuda.rs:31-97
//   EncodingConfig             src/benchmarks/utils.rs:10-49
//   encode_gpu                 src/benchmarks/utils.rs:100-146
// Where the Rust code panics on a bad argument (assert!), these throw
// std::invalid_argument; where it returns Err(String), they throw zoda::Error.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <algorithm>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/zoda_hip.h"

namespace zoda {

constexpr uint32_t BABYBEAR_PRIME = ZODA_P;

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, const char* name) {
  if (rc != 0)
    throw Error(rc, std::string(name) + " failed with status " + std::to_string(rc) + ": " +
                        zoda_hip_last_error());
}

inline bool gpu_available() {
  int count = 0;
  return zoda_hip_device_count(&count) == 0 && count > 0;
}

inline std::vector<uint32_t> roots_of_unity(uint32_t n) {
  std::vector<uint32_t> r(n ? n : 1);
  check(zoda_hip_roots_of_unity(n, r.data()), "zoda_hip_roots_of_unity");
  r.resize(n);
  return r;
}

namespace detail {
inline bool is_pow2(size_t x) { return x && !(x & (x - 1)); }

inline void run(uint32_t* values, size_t len, size_t n, size_t stride, bool inverse) {
  if (!is_pow2(n)) throw std::invalid_argument("NTT size must be power of 2");
  if (stride < n) throw std::invalid_argument("stride must be at least n");
  if (stride == 0 || len % stride != 0)
    throw std::invalid_argument("values must be whole polynomials of stride elements");
  const size_t polys = len / stride;
  std::vector<uint32_t> roots = roots_of_unity((uint32_t)n);
  if (inverse)
    check(zoda_hip_intt(values, roots.data(), (uint32_t)n, (uint32_t)polys, (uint32_t)stride),
          "zoda_hip_intt");
  else
    check(zoda_hip_ntt(values, roots.data(), (uint32_t)n, (uint32_t)polys, (uint32_t)stride),
          "zoda_hip_ntt");
}
}  // namespace detail

// Forward NTT in place on `len` host values: polynomials `stride` apart, the
// first n of each transformed.
inline void ntt_gpu(uint32_t* values, size_t len, size_t n, size_t stride) {
  detail::run(values, len, n, stride, false);
}

inline void intt_gpu(uint32_t* values, size_t len, size_t n, size_t stride) {
  detail::run(values, len, n, stride, true);
}

// `cols` polynomials of `rows` values in pinned host memory, zero-filled;
// element (row, col) at col * rows + row.
class PinnedSquare {
 public:
  PinnedSquare(size_t rows, size_t cols) : rows_(rows), cols_(cols) {
    const size_t count = rows * cols > 0 ? rows * cols : 1;
    ptr_ = zoda_hip_host_alloc(count);
    if (!ptr_) throw Error(ZODA_ENOMEM, "pinned allocation of " + std::to_string(count) + " u32 failed");
    for (size_t i = 0; i < count; i++) ptr_[i] = 0;
  }
  PinnedSquare(const PinnedSquare&) = delete;
  PinnedSquare& operator=(const PinnedSquare&) = delete;
  PinnedSquare(PinnedSquare&& o) noexcept : ptr_(o.ptr_), rows_(o.rows_), cols_(o.cols_) { o.ptr_ = nullptr; }
  ~PinnedSquare() { zoda_hip_host_free(ptr_); }

  size_t rows() const { return rows_; }
  size_t cols() const { return cols_; }
  size_t size() const { return rows_ * cols_; }
  uint32_t* data() { return ptr_; }
  const uint32_t* data() const { return ptr_; }
  uint32_t& operator[](size_t i) { return ptr_[i]; }
  void set(size_t row, size_t col, uint32_t v) { ptr_[col * rows_ + row] = v; }
  uint32_t get(size_t row, size_t col) const { return ptr_[col * rows_ + row] % BABYBEAR_PRIME; }
  const uint32_t* column(size_t col) const { return ptr_ + col * rows_; }
  std::vector<uint32_t> row(size_t r) const {
    std::vector<uint32_t> out(cols_);
    for (size_t c = 0; c < cols_; c++) out[c] = get(r, c);
    return out;
  }

 private:
  uint32_t* ptr_ = nullptr;
  size_t rows_, cols_;
};

struct EncodingConfig {
  size_t k, n, row_size;
  size_t data_size_bytes() const { return k * row_size; }
  size_t total_rows() const { return k + n; }
  size_t elements_per_row() const { return row_size / 4; }
  size_t num_positions() const { return elements_per_row(); }
  static size_t next_pow2(size_t x) {
    size_t r = 1;
    while (r < x) r <<= 1;
    return r;
  }
  size_t ntt_size_k() const { return next_pow2(k); }
  size_t ntt_size_kn() const { return next_pow2(k + n); }
};

// encode_gpu: row-major k x cols input in host memory -> `square`
// (ntt_size_kn rows x cols, column-major PinnedSquare layout).  The layout
// change happens on the device; there is no host transpose.
inline void encode_gpu(const EncodingConfig& cfg, const uint32_t* input, PinnedSquare& square) {
  const size_t cols = cfg.num_positions();
  if (square.rows() != cfg.ntt_size_kn() || square.cols() != cols)
    throw std::invalid_argument("square has the wrong shape");
  check(zoda_hip_encode_host(input, (uint32_t)cfg.k, (uint32_t)cols, (uint32_t)cfg.n, square.data(),
                             ZODA_LAYOUT_COL),
        "zoda_hip_encode_host");
}

// Dense replacement of BabyBearDataSquare (src/zoda/babybear.rs:49-112):
// the same accessors over one column-major array instead of a list of cells
// searched linearly (which makes building an s x s square O(s^4)).  Cells
// never set read as 0; the reference's square has them missing, which only
// differs for squares that are not filled.  hash_root() hashes the cells
// column by column -- the reference hashes in insertion order, and the
// encoder fills its square column by column (babybear.rs:156-160), so the
// roots agree for it.
class DataSquare {
 public:
  DataSquare() = default;
  DataSquare(size_t columns, size_t rows) : columns_(columns), rows_(rows), v_(columns * rows, 0u) {}

  void set_cell(size_t column, size_t row, uint32_t value) {
    if (column >= columns_ || row >= rows_) grow(std::max(columns_, column + 1), std::max(rows_, row + 1));
    v_[column * rows_ + row] = value;
  }
  uint32_t get_cell(size_t column, size_t row) const {
    return column < columns_ && row < rows_ ? v_[column * rows_ + row] : 0u;
  }
  std::vector<uint32_t> get_row(size_t row) const {
    std::vector<uint32_t> out(columns_);
    for (size_t c = 0; c < columns_; c++) out[c] = get_cell(c, row);
    return out;
  }
  std::vector<uint32_t> get_column(size_t column) const {
    std::vector<uint32_t> out(rows_);
    for (size_t r = 0; r < rows_; r++) out[r] = get_cell(column, r);
    return out;
  }
  // SHA-256 of the cells' little-endian bytes, column by column, as hex.
  std::string hash_root() const {
    char hex[65];
    check(zoda_data_root(v_.data(), v_.size(), hex), "zoda_data_root");
    return std::string(hex);
  }
  size_t columns() const { return columns_; }
  size_t rows() const { return rows_; }
  const uint32_t* data() const { return v_.data(); }

 private:
  void grow(size_t columns, size_t rows) {
    std::vector<uint32_t> v(columns * rows, 0u);
    for (size_t c = 0; c < columns_; c++)
      for (size_t r = 0; r < rows_; r++) v[c * rows + r] = v_[c * rows_ + r];
    v_.swap(v);
    columns_ = columns;
    rows_ = rows;
  }
  size_t columns_ = 0, rows_ = 0;
  std::vector<uint32_t> v_;
};

// Device-resident encode (asynchronous on `stream`, a hipStream_t or nullptr).
inline void encode_device(const uint32_t* d_in, uint32_t k, uint32_t cols, uint32_t n, uint32_t* d_out,
                          int layout = ZODA_LAYOUT_ROW, void* stream = nullptr) {
  check(zoda_hip_encode_device(d_in, k, cols, n, d_out, layout, stream), "zoda_hip_encode_device");
}

}  // namespace zoda
