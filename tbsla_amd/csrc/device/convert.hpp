// GPU COO -> CSR / ELL conversion (declarations usable from host C++).
#pragma once

#include <vector>

namespace tbsla {
namespace hip {

// The entries of a COO (row[k], col[k], val[k]) in (row, column) order, ties
// in input order -- the order of host::row_col_order -- sorted on the
// current GPU.  CSR: rowptr (n_row + 1), colidx, values.  Rows must lie in
// [0, n_row) (std::invalid_argument otherwise).
void coo_to_csr_gpu(int n_row, const std::vector<int> &row, const std::vector<int> &col,
                    const std::vector<double> &val, std::vector<int> &rowptr, std::vector<int> &colidx,
                    std::vector<double> &values);

// ELL, row-major n_row x K with K the longest row: slot j of row i holds the
// row's j-th entry in that order, the rest (column 0, value 0).
void coo_to_ell_gpu(int n_row, const std::vector<int> &row, const std::vector<int> &col,
                    const std::vector<double> &val, int &K, std::vector<int> &columns, std::vector<double> &values);

}  // namespace hip
}  // namespace tbsla
