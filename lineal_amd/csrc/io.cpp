// Matrix and vector files of the reference (io/tensor.hpp:25-263), host side.
//
//   SMF  u64 rows, u64 columns, u64 nonzeros; u64 row offsets[rows + 1];
//        (u64 column, f64 value) pairs, row by row
//   DVF  u64 size; f64 values[size]
//
// All words little-endian, as the reference writes them on x86 hosts.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "internal.hpp"

namespace lgpu::io {
namespace {

struct File {
  std::FILE* f = nullptr;
  std::string path;
  File(const char* p, const char* mode) : path(p ? p : "") {
    LGPU_REQUIRE(p, "NULL path");
    f = std::fopen(p, mode);
    LGPU_REQUIRE(f, "cannot open " + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
  File(const File&) = delete;
  File& operator=(const File&) = delete;

  template <typename W>
  W get() {
    W w;
    LGPU_REQUIRE(std::fread(&w, sizeof(W), 1, f) == 1, "truncated file " + path);
    return w;
  }
  template <typename W>
  void get_n(W* out, size_t n) {
    if (n) LGPU_REQUIRE(std::fread(out, sizeof(W), n, f) == n, "truncated file " + path);
  }
  template <typename W>
  void put(W w) {
    LGPU_REQUIRE(std::fwrite(&w, sizeof(W), 1, f) == 1, "cannot write " + path);
  }
  void seek(long off) { LGPU_REQUIRE(std::fseek(f, off, SEEK_SET) == 0, "cannot seek in " + path); }
};

}  // namespace

SmfHeader smf_header(const char* path) {
  File in(path, "rb");
  SmfHeader h;
  h.rows = in.get<uint64_t>();
  h.cols = in.get<uint64_t>();
  h.nnz = in.get<uint64_t>();
  return h;
}

// SmfReader::read (io/tensor.hpp:47-75) into RowWiseBuilder semantics
// (tensor/dynamic/csr/shared.hpp:730-764): within a row, an entry whose column
// was seen before is added to the first one in f64, in file order; the row is
// then sorted by column.
void smf_read(const char* path, SmfHeader& h, std::vector<int64_t>& rowptr,
              std::vector<int64_t>& cols, std::vector<double>& vals) {
  File in(path, "rb");
  h.rows = in.get<uint64_t>();
  h.cols = in.get<uint64_t>();
  h.nnz = in.get<uint64_t>();
  LGPU_REQUIRE(h.rows < (uint64_t(1) << 40) && h.cols < (uint64_t(1) << 40),
               "implausible SMF dimensions in " + std::string(path));
  std::vector<uint64_t> off(size_t(h.rows + 1));
  in.get_n(off.data(), off.size());
  LGPU_REQUIRE(off[0] == 0, "SMF row offsets must start at 0");
  for (size_t r = 0; r < h.rows; ++r)
    LGPU_REQUIRE(off[r + 1] >= off[r], "SMF row offsets must not decrease");
  const uint64_t entries = off[size_t(h.rows)];
  std::vector<unsigned char> raw(size_t(entries) * 16);
  in.get_n(raw.data(), raw.size());
  rowptr.assign(1, 0);
  cols.clear();
  vals.clear();
  std::vector<std::pair<int64_t, double>> row;
  for (size_t r = 0; r < h.rows; ++r) {
    row.clear();
    for (uint64_t k = off[r]; k < off[r + 1]; ++k) {
      uint64_t c;
      double v;
      std::memcpy(&c, &raw[size_t(k) * 16], 8);
      std::memcpy(&v, &raw[size_t(k) * 16 + 8], 8);
      LGPU_REQUIRE(c < h.cols, "SMF column index out of range in " + std::string(path));
      auto it = std::find_if(row.begin(), row.end(),
                             [&](const auto& e) { return e.first == int64_t(c); });
      if (it != row.end()) it->second += v;
      else row.emplace_back(int64_t(c), v);
    }
    std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    for (const auto& [c, v] : row) {
      cols.push_back(c);
      vals.push_back(v);
    }
    rowptr.push_back(int64_t(cols.size()));
  }
}

// SmfWriter (io/tensor.hpp:158-213): the header's nonzero count is the
// matrix's stored-entry count when it has one (CsrMatrix::non_zero_num), else
// the number of entries written; offsets and pairs cover the nonzero entries
// only, in column order.
void smf_write(const char* path, uint64_t rows, uint64_t cols_num, const int64_t* rowptr,
               const int64_t* cols, const double* vals, bool stored_count) {
  File out(path, "wb");
  out.put<uint64_t>(rows);
  out.put<uint64_t>(cols_num);
  out.put<uint64_t>(stored_count ? uint64_t(rowptr[rows]) : 0);
  uint64_t written = 0;
  out.put<uint64_t>(0);
  for (uint64_t r = 0; r < rows; ++r) {
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) written += vals[k] != 0.0;
    out.put<uint64_t>(written);
  }
  for (uint64_t r = 0; r < rows; ++r)
    for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k)
      if (vals[k] != 0.0) {
        out.put<uint64_t>(uint64_t(cols[k]));
        out.put<double>(vals[k]);
      }
  if (!stored_count) {
    out.seek(16);
    out.put<uint64_t>(written);
  }
}

uint64_t dvf_size(const char* path) {
  File in(path, "rb");
  return in.get<uint64_t>();
}

// DvfReader::read (io/tensor.hpp:226-232)
void dvf_read(const char* path, std::vector<double>& v) {
  File in(path, "rb");
  const uint64_t n = in.get<uint64_t>();
  LGPU_REQUIRE(n < (uint64_t(1) << 40), "implausible DVF size in " + std::string(path));
  v.resize(size_t(n));
  in.get_n(v.data(), v.size());
}

// DvfWriter (io/tensor.hpp:255-263)
void dvf_write(const char* path, const double* v, uint64_t n) {
  File out(path, "wb");
  out.put<uint64_t>(n);
  for (uint64_t i = 0; i < n; ++i) out.put<double>(v[i]);
}

}  // namespace lgpu::io
