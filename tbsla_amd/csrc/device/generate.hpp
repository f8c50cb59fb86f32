// GPU generation of fill_cqmat / fill_cqmat_stochastic CSR blocks
// (declarations usable from host C++).
#pragma once

#include <functional>
#include <vector>

#include "../host/fill_common.hpp"

namespace tbsla {
namespace hip {

struct GeneratedCsr {
  std::vector<int> rowptr;  // nrows + 1 entries, rowptr[0] = 0
  std::vector<int> colidx;
  std::vector<double> values;
};

// The in-block entries of rows [b.f_row, b.f_row + nrows) of a cqmat matrix,
// generated on the current GPU and copied to the host: the same arrays as
// host::generate_rows (entry order, columns, values of the overlay or 1.0).
// With `stochastic`, every value is 1 / (count of its column over the whole
// n_row x n_col matrix), as fill_cqmat_stochastic computes it
// (values_generation.cpp:47-62, MatrixCSR.cpp:411-421).
void generate_cqmat_csr(const host::CqmatBlock &b, long nrows, bool stochastic, GeneratedCsr &out);

// The same rows as SCOO arrays: global row ids (row), global columns (col),
// values, in entry order (MatrixSCOO.cpp:288-399); rowptr as above.
void generate_cqmat_scoo(const host::CqmatBlock &b, long nrows, bool stochastic, std::vector<int> &rowptr,
                         std::vector<int> &row, std::vector<int> &col, std::vector<double> &values);

// The same rows as a row-major ELL block of fill_rows >= nrows rows
// (MatrixELL.cpp:262-404): K = width(rowptr) slots per row, a row's entries
// first, the rest column 0, value 0.0 (rows past nrows: padding only); no
// arrays when fill is false.
void generate_cqmat_ell(const host::CqmatBlock &b, long nrows, long fill_rows, bool stochastic,
                        const std::function<int(const std::vector<int> &)> &width, int &K, std::vector<int> &columns,
                        std::vector<double> &values, bool fill);

// COO: entries [first, first + n) of the global stream of b's matrix (b
// covering the whole matrix), global ids (MatrixCOO.cpp:332-383).
void generate_cqmat_coo(const host::CqmatBlock &b, long first, long n, bool stochastic, std::vector<int> &row,
                        std::vector<int> &col, std::vector<double> &values);

}  // namespace hip
}  // namespace tbsla
