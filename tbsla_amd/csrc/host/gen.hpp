// Host-side generators and partition math, bit-identical to the reference.
//
//   lnv / pflv        src/cpp/utils/range.cpp:7-26
//   cdiag_value       src/cpp/utils/values_generation.cpp:4-24
//   cqmat_value       src/cpp/utils/values_generation.cpp:26-45 (glibc rand_r)
//   column counts     src/cpp/utils/values_generation.cpp:47-62
//
// The reference walks the cqmat entry stream serially.  Here the stream is
// addressed in closed form: CqmatStream::base(r) is the global generation
// index of row r's first candidate, so rows are generated in parallel and
// every entry still comes from cqmat_value(i) with the reference's i.
#pragma once

#include <stdlib.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

namespace tbsla {
namespace gen {

std::size_t lnv(std::size_t size, int local_pos, int number_pos);
std::size_t pflv(std::size_t size, int local_pos, int number_pos);

// values_generation.cpp:26-45.  Returns the (row, column) of stream entry i;
// the value is always 1.0 in the reference.
inline void cqmat_value(std::size_t i, std::size_t nr, std::size_t nc, std::size_t c_, double q,
                        unsigned seed_mult, std::uint64_t &row, std::uint64_t &col) {
  unsigned int seedp = static_cast<unsigned int>(i);
  if (seed_mult > 0) seedp = static_cast<unsigned int>(static_cast<std::size_t>(seed_mult) * i);
  const std::size_t c = std::min(nc, c_);
  const std::size_t nfull = std::min(nc - c + 1, nr);
  if (i < c * nfull) {
    row = i / c;
    if ((static_cast<double>(rand_r(&seedp)) / RAND_MAX) < q)
      col = static_cast<std::size_t>(rand_r(&seedp)) % nc;
    else
      col = i / c + i % c;
    return;
  }
  std::size_t it = i - c * nfull;
  std::size_t curr_row = 1;
  while (curr_row < nr - nfull && it >= c - curr_row) {
    it -= (c - curr_row);
    curr_row++;
  }
  curr_row += nfull - 1;
  row = curr_row;
  // (double)(int / int): the integer-division quirk of values_generation.cpp:43
  if (static_cast<double>(rand_r(&seedp) / RAND_MAX) < q)
    col = static_cast<std::size_t>(rand_r(&seedp)) % nc;
  else
    col = curr_row + it;
}

// values_generation.cpp:4-24 (size_t / long arithmetic wraps modulo 2^64).
void cdiag_value(std::size_t i, std::size_t nr, std::size_t nc, std::size_t cdiag, std::uint64_t &row,
                 std::uint64_t &col);

// Project-defined counter-based uniform generator (no reference counterpart):
// u01(seed, i) = (splitmix64(splitmix64(seed) ^ i) >> 11) * 2^-53, in [0, 1).
std::uint64_t splitmix64(std::uint64_t z);
inline double u01(std::uint64_t seed, std::uint64_t idx) {
  return static_cast<double>(splitmix64(splitmix64(seed) ^ idx) >> 11) * 0x1.0p-53;
}

// Geometry of the cqmat entry stream of an (n_row x n_col, c) matrix, in the
// int arithmetic of MatrixCSR.cpp:317-341.
struct CqmatStream {
  long n_row, n_col, cc, min_;
  CqmatStream(int n_row_, int n_col_, int c) : n_row(n_row_), n_col(n_col_) {
    cc = std::min(c, n_col_);
    min_ = std::min<long>(n_col - cc + 1, n_row);
  }
  // candidates generated for row r (0 for rows past min(n_row, n_col))
  long candidates(long r) const {
    if (r < min_) return cc;
    if (r >= std::min(n_row, n_col)) return 0;
    return std::max<long>(0, cc - (r - min_) - 1);
  }
  // global generation index of row r's first candidate
  long base(long r) const {
    if (r <= min_) return r * cc;
    long k = std::min(r, std::min(n_row, n_col)) - min_;  // partial rows before r
    // sum_{t<k} (cc - t - 1) with the reference loop's clipping at zero length
    long s = 0;
    if (k > 0) {
      long kk = std::min(k, cc > 0 ? cc - 1 : 0);
      s = kk * (cc - 1) - kk * (kk - 1) / 2;
    }
    return min_ * cc + s;
  }
  long total() const { return base(std::min(n_row, n_col)); }
};

// Column counts of the whole cqmat matrix (the divisor of
// fill_cqmat_stochastic, values_generation.cpp:47-62).  Counts are exact
// integers, so double(count) equals the reference's running sum of 1.0.
std::vector<double> cqmat_column_sums(std::size_t nr, std::size_t nc, std::size_t c, double q, unsigned seed_mult);

// Run f(begin, end) over [0, n) on the host's cores.
void parallel_for(long n, const std::function<void(long, long)> &f, long min_chunk = 4096);
int host_threads();

}  // namespace gen
}  // namespace tbsla
