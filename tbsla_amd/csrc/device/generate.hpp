// GPU generation of fill_cqmat / fill_cqmat_stochastic CSR blocks
// (declarations usable from host C++).
#pragma once

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

}  // namespace hip
}  // namespace tbsla
