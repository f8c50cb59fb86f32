// Shared pieces of the per-format generators and printers.
#pragma once

#include <tbsla/cpp/Matrix.hpp>

#include <cmath>
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "gen.hpp"

namespace tbsla {
namespace host {

// One rank's block [f_row, f_row+ln_row) x [f_col, f_col+ln_col) of a cqmat
// matrix, plus the stream geometry and value overlay.
struct CqmatBlock {
  int n_row, n_col, c;
  double q;
  unsigned seed;
  std::uint64_t value_seed;  // 0: reference value 1.0
  long f_row, ln_row, f_col, ln_col;
  gen::CqmatStream s;
  CqmatBlock(int n_row_, int n_col_, int c_, double q_, unsigned seed_, std::uint64_t vseed, long fr, long lr,
             long fc, long lc)
      : n_row(n_row_), n_col(n_col_), c(c_), q(q_), seed(seed_), value_seed(vseed), f_row(fr), ln_row(lr),
        f_col(fc), ln_col(lc), s(n_row_, n_col_, c_) {}
  bool in_block(long ii, long jj) const {
    return ii >= f_row && ii < f_row + ln_row && jj >= f_col && jj < f_col + ln_col;
  }
  double value(long gidx) const {
    return value_seed ? 1.0 - gen::u01(value_seed, static_cast<std::uint64_t>(gidx)) : 1.0;
  }
  void entry(long i, long &ii, long &jj) const {
    std::uint64_t r, cl;
    gen::cqmat_value(static_cast<std::size_t>(i), static_cast<std::size_t>(n_row), static_cast<std::size_t>(n_col),
                     static_cast<std::size_t>(c), q, seed, r, cl);
    ii = static_cast<long>(r);
    jj = static_cast<long>(cl);
  }
  // rows with candidates inside the row block (MatrixCSR.cpp:347 bound)
  long lim3() const { return std::min<long>(std::min<long>(n_row, n_col), f_row + ln_row); }
  // candidate count of the row block (the reference's `nv`, MatrixCSR.cpp:319-341)
  long candidates_in_block() const {
    long nv = 0;
    for (long r = f_row; r < f_row + ln_row; r++) nv += s.candidates(r);
    return nv;
  }
};

// In-block entries of rows [r0, r1), generated row-parallel, concatenated in
// row order.  counts[r - r0] = entries kept in row r.
struct RowEntries {
  std::vector<int> counts;
  std::vector<int> rows, cols;
  std::vector<double> vals;
};
RowEntries generate_rows(const CqmatBlock &b, long r0, long r1, bool want_rows);

// The CSR/SCOO `nnz` pass (MatrixCSR.cpp:343-358): for rows past the full
// rows it walks the stream with a stride of c per row, so it can count
// entries the fill never emits.
long overcount_nnz(const CqmatBlock &b, long kept_full_rows);

// Number of entries in the two cdiag diagonals inside the block.
long cdiag_block_count(long f_row, long ln_row, long f_col, long ln_col, int cdiag);

// utils/vector.hpp:11-19 (streamvector) and utils/vector.cpp:59-75
template <class T>
void streamvector(std::ostream &os, const std::string &name, const std::vector<T> &v) {
  os << name << " : ";
  for (std::size_t i = 0; i < v.size(); i++) {
    os << v[i];
    if (i + 1 < v.size()) os << ", ";
  }
}
void print_dense_matrix(int nr, int nc, const std::vector<double> &m, std::ostream &os);
std::ostream &print_header(std::ostream &os, const char *title, int n_row, int n_col, int ln_row, int ln_col,
                           int f_row, int f_col, long nnz, int pr, int pc, int NR, int NC);

// stable (row, col) ordering permutation of a COO (the order of
// csr.cpp:3-8; ties keep their input order)
std::vector<int> row_col_order(int n_row, const std::vector<int> &row, const std::vector<int> &col);

// binary header (MatrixCSR.cpp:180-190)
void write_header(std::ostream &os, int n_row, int n_col, int ln_row, int ln_col, int f_row, int f_col, long nnz,
                  int pr, int pc, int NR, int NC);
void read_header(std::istream &is, int &n_row, int &n_col, int &ln_row, int &ln_col, int &f_row, int &f_col,
                 long &nnz, int &pr, int &pc, int &NR, int &NC);
template <class T>
void write_array(std::ostream &os, const std::vector<T> &v) {
  std::size_t n = v.size();
  os.write(reinterpret_cast<const char *>(&n), sizeof(n));
  os.write(reinterpret_cast<const char *>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
}
// A size that the rest of a seekable stream cannot hold (a truncated or
// corrupt file) throws std::runtime_error before anything is allocated.
template <class T>
void read_array(std::istream &is, std::vector<T> &v) {
  std::size_t n = 0;
  is.read(reinterpret_cast<char *>(&n), sizeof(n));
  if (!is) throw std::runtime_error("tbsla: truncated binary file (array size missing)");
  const std::streampos here = is.tellg();
  if (here != std::streampos(-1)) {
    is.seekg(0, std::ios::end);
    const std::streamoff left = is.tellg() - here;
    is.seekg(here);
    if (left < 0 || n > static_cast<std::size_t>(left) / sizeof(T))
      throw std::runtime_error("tbsla: binary file holds " + std::to_string(left) + " bytes after an array size of " +
                               std::to_string(n) + " elements");
  }
  v.resize(n);
  is.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
  if (!is) throw std::runtime_error("tbsla: truncated binary file");
}

}  // namespace host
}  // namespace tbsla
