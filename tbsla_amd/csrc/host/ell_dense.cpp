// tbsla::cpp::MatrixELL and MatrixDENSE (reference: src/cpp/matrix/MatrixELL.cpp,
// src/cpp/matrix/MatrixDENSE.cpp).
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>

#include <algorithm>

#include "../device/products.hpp"
#include "fill_common.hpp"
#include "../device/generate.hpp"
#include "../device/convert.hpp"

namespace tbsla {
namespace cpp {

// ------------------------------------------------------------------ ELL ----

// MatrixELL.cpp:12-54: (row, col) order, row-major padding to max_col
MatrixELL::MatrixELL(const MatrixCOO &m) { from_coo(m, false); }

MatrixELL::MatrixELL(const MatrixCOO &m, bool gpu) { from_coo(m, gpu); }

// MatrixELL.cpp:12-54; the (row, column) order of the entries is computed on
// the host or, with `gpu`, on the current GPU (the same order)
void MatrixELL::from_coo(const MatrixCOO &m, bool gpu) {
  n_row = m.get_n_row();
  n_col = m.get_n_col();
  ln_row = m.get_n_row();
  ln_col = m.get_n_col();
  f_row = m.get_f_row();
  f_col = m.get_f_col();
  pr = m.get_pr();
  pc = m.get_pc();
  NR = m.get_NR();
  NC = m.get_NC();
  nnz = m.get_nnz();
  const auto &r = m.get_row();
  const auto &c = m.get_col();
  const auto &v = m.get_values();
  if (v.empty()) {
    max_col = 0;
    nnz = 0;
    return;
  }
  if (gpu) {
    if (r.size() != v.size() || c.size() != v.size())
      throw std::invalid_argument("tbsla: GPU conversion needs row, col and values of one length");
    int K = 0;
    hip::coo_to_ell_gpu(n_row, r, c, v, K, columns, values);
    max_col = K;
    return;
  }
  std::vector<int> p = host::row_col_order(n_row, r, c);
  std::vector<int> nvrow(static_cast<std::size_t>(n_row), 0);
  for (std::size_t k = 0; k < v.size(); k++) nvrow[static_cast<std::size_t>(r[k])]++;
  max_col = *std::max_element(nvrow.begin(), nvrow.end());
  values.assign(static_cast<std::size_t>(n_row) * max_col, 0);
  columns.assign(static_cast<std::size_t>(n_row) * max_col, 0);
  std::size_t incr = 0;
  for (int i = 0; i < n_row; i++)
    for (int j = 0; j < nvrow[static_cast<std::size_t>(i)]; j++) {
      values[static_cast<std::size_t>(i) * max_col + j] = v[static_cast<std::size_t>(p[incr])];
      columns[static_cast<std::size_t>(i) * max_col + j] = c[static_cast<std::size_t>(p[incr])];
      incr++;
    }
}

std::ostream &operator<<(std::ostream &os, const MatrixELL &m) { return m.print(os); }

std::ostream &MatrixELL::print(std::ostream &os) const {
  host::print_header(os, "------ ELL ------", n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  os << "maximum number of columns  -  max_col : " << max_col << std::endl;
  host::streamvector<double>(os, "val", values);
  os << std::endl;
  host::streamvector<int>(os, "col", columns);
  os << std::endl;
  os << "-----------------" << std::endl << std::flush;
  return os;
}

std::ostream &MatrixELL::print_as_dense(std::ostream &os) {
  std::vector<double> d(static_cast<std::size_t>(n_row) * n_col, 0);
  if (max_col > 0)
    for (std::size_t i = 0; i < std::min(static_cast<std::size_t>(n_row), values.size() / max_col); i++)
      for (int j = 0; j < max_col; j++)
        d[i * n_col + static_cast<std::size_t>(columns[i * max_col + j])] += values[i * max_col + j];
  host::print_dense_matrix(n_row, n_col, d, os);
  return os;
}

std::vector<double> MatrixELL::spmv(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r(static_cast<std::size_t>(n_row), 0);
  this->Ax(r, v, vect_incr);
  return r;
}

// MatrixELL.cpp:101-114 on the GPU (ell_kernel, sliced column-major mirror)
void MatrixELL::Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr) const {
  hip::Ax_host(*this, r, v, vect_incr);
}

std::ostream &MatrixELL::print_infos(std::ostream &os) {
  os << "-----------------" << std::endl;
  os << "------ ELL ------" << std::endl;
  os << "--- general   ---" << std::endl;
  os << "n_row : " << n_row << std::endl;
  os << "n_col : " << n_col << std::endl;
  os << "nnz : " << nnz << std::endl;
  os << "max_col : " << max_col << std::endl;
  os << "--- capacity  ---" << std::endl;
  os << "values : " << values.capacity() << std::endl;
  os << "columns : " << columns.capacity() << std::endl;
  os << "--- size      ---" << std::endl;
  os << "values : " << values.size() << std::endl;
  os << "columns : " << columns.size() << std::endl;
  os << "-----------------" << std::endl;
  return os;
}

std::ostream &MatrixELL::print_stats(std::ostream &os) {
  os << "upper values : " << 0 << std::endl;
  os << "lower values : " << 0 << std::endl;
  os << "diag  values : " << 0 << std::endl;
  return os;
}

std::ostream &MatrixELL::write(std::ostream &os) {
  host::write_header(os, n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  os.write(reinterpret_cast<const char *>(&max_col), sizeof(max_col));
  host::write_array(os, values);
  host::write_array(os, columns);
  return os;
}

std::istream &MatrixELL::read(std::istream &is, std::size_t, std::size_t) {
  // read into temporaries: a truncated or corrupt file throws and leaves
  // the matrix as it was
  int h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_pr, h_pc, h_NR, h_NC;
  long h_nnz;
  host::read_header(is, h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_nnz, h_pr, h_pc, h_NR, h_NC);
  int h_max_col = 0;
  is.read(reinterpret_cast<char *>(&h_max_col), sizeof(h_max_col));
  if (!is) throw std::runtime_error("tbsla: truncated binary file (ELL max_col)");
  std::vector<double> r_values;
  host::read_array(is, r_values);
  std::vector<int> r_columns;
  host::read_array(is, r_columns);
  n_row = h_nr; n_col = h_nc; ln_row = h_lnr; ln_col = h_lnc; f_row = h_fr; f_col = h_fc;
  nnz = h_nnz; pr = h_pr; pc = h_pc; NR = h_NR; NC = h_NC;
  max_col = h_max_col;
  values.swap(r_values);
  columns.swap(r_columns);
  count_stochastic_ = false;
  invalidate_device();
  return is;
}

// MatrixELL.cpp:192-260
void MatrixELL::fill_cdiag(int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  columns.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  max_col = cdiag == 0 ? 1 : 2;
  const long nv = host::cdiag_block_count(f_row, ln_row, f_col, ln_col, cdiag);
  if (nv == 0) return;
  nnz = nv;
  auto inb = [&](long ii, long jj) {
    return ii >= f_row && ii < static_cast<long>(f_row) + ln_row && jj >= f_col && jj < static_cast<long>(f_col) + ln_col;
  };
  values.reserve(static_cast<std::size_t>(max_col) * ln_row);
  columns.reserve(static_cast<std::size_t>(max_col) * ln_row);
  for (long i = f_row; i < static_cast<long>(f_row) + ln_row; i++) {
    const bool lo = inb(i, i - cdiag);
    columns.push_back(lo ? static_cast<int>(i - cdiag) : 0);
    values.push_back(lo ? 1 : 0);
    if (cdiag != 0) {
      const bool hi = inb(i, i + cdiag);
      columns.push_back(hi ? static_cast<int>(i + cdiag) : 0);
      values.push_back(hi ? 1 : 0);
    }
  }
}

void MatrixELL::fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                           int NC) {
  fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
}

// MatrixELL.cpp:262-392, including the slot-index-vs-column count of the
// partial rows (:338-342) that can raise max_col.
void MatrixELL::fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                  int NR, int NC, std::uint64_t value_seed) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  columns.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  host::RowEntries e = host::generate_rows(b, f_row, std::max<long>(lim3, f_row), false);
  nnz = 0;
  max_col = 0;
  for (long r = f_row; r < lim3; r++) {
    int nbc = e.counts[static_cast<std::size_t>(r - f_row)];
    nnz += nbc;
    if (r >= b.s.min_)
      for (long j = b.s.candidates(r); j < b.s.cc; j++)
        if (j >= f_col && j < static_cast<long>(f_col) + ln_col) nbc++;
    max_col = std::max(max_col, nbc);
  }
  if (b.candidates_in_block() == 0) return;
  const long fill_end = std::min<long>(n_row, static_cast<long>(f_row) + ln_row);
  const long nrows = std::max(0L, fill_end - f_row);
  values.assign(static_cast<std::size_t>(nrows) * max_col, 0.0);
  columns.assign(static_cast<std::size_t>(nrows) * max_col, 0);
  std::size_t src = 0;
  for (long r = f_row; r < std::max<long>(lim3, f_row); r++) {
    const int cnt = e.counts[static_cast<std::size_t>(r - f_row)];
    const std::size_t dst = static_cast<std::size_t>(r - f_row) * max_col;
    std::copy(e.vals.begin() + static_cast<long>(src), e.vals.begin() + static_cast<long>(src + cnt), values.begin() + static_cast<long>(dst));
    std::copy(e.cols.begin() + static_cast<long>(src), e.cols.begin() + static_cast<long>(src + cnt), columns.begin() + static_cast<long>(dst));
    src += static_cast<std::size_t>(cnt);
  }
}

// MatrixELL.cpp:394-404 (padding slots hold 0.0 and are left alone)
// fill_cqmat_values / fill_cqmat_stochastic with the block's entries
// generated and laid out on the GPU (device/generate.hip); metadata, nnz,
// the slot count (with the quirk of MatrixELL.cpp:338-342) and the early
// return follow the host path above.
void MatrixELL::fill_cqmat_gpu(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                               int NC, std::uint64_t value_seed, bool stochastic) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  columns.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, stochastic ? 0 : value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  const long nrows = std::max<long>(lim3, f_row) - f_row;
  const long fill_end = std::min<long>(n_row, static_cast<long>(f_row) + ln_row);
  long kept = 0;
  auto width = [&](const std::vector<int> &rp) {
    int K = 0;
    kept = 0;
    for (long r = f_row; r < lim3; r++) {
      int nbc = rp[static_cast<std::size_t>(r - f_row) + 1] - rp[static_cast<std::size_t>(r - f_row)];
      kept += nbc;
      if (r >= b.s.min_)
        for (long j = b.s.candidates(r); j < b.s.cc; j++)
          if (j >= f_col && j < static_cast<long>(f_col) + ln_col) nbc++;
      K = std::max(K, nbc);
    }
    return K;
  };
  hip::generate_cqmat_ell(b, nrows, std::max(0L, fill_end - f_row), stochastic, width, max_col, columns, values,
                          b.candidates_in_block() != 0);
  nnz = kept;
  if (stochastic && !values.empty()) count_stochastic_ = true;
}

void MatrixELL::fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                      int NR, int NC) {
  fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
  if (values.empty()) return;
  std::vector<double> sum = gen::cqmat_column_sums(n_row, n_col, c, q, seed_mult);
  const long rows = std::min<long>(ln_row, static_cast<long>(values.size() / static_cast<std::size_t>(max_col)));
  gen::parallel_for(rows, [&](long b, long e) {
    for (long i = b; i < e; i++)
      for (long j = 0; j < max_col; j++) {
        const std::size_t k = static_cast<std::size_t>(i * max_col + j);
        if (values[k] != 0) values[k] /= sum[static_cast<std::size_t>(columns[k])];
      }
  });
  count_stochastic_ = true;
}

// MatrixELL.cpp:406-422
void MatrixELL::normalize_columns() {
  std::vector<double> sum(static_cast<std::size_t>(ln_col), 0);
  const long rows = max_col ? std::min<long>(ln_row, static_cast<long>(values.size() / static_cast<std::size_t>(max_col))) : 0;
  for (long i = 0; i < rows; i++)
    for (long j = 0; j < max_col; j++) {
      long idx = columns[static_cast<std::size_t>(i * max_col + j)] - f_col;
      if (idx < 0) idx = 0;
      sum[static_cast<std::size_t>(idx)] += values[static_cast<std::size_t>(i * max_col + j)];
    }
  for (long i = 0; i < rows; i++)
    for (long j = 0; j < max_col; j++) {
      long idx = columns[static_cast<std::size_t>(i * max_col + j)] - f_col;
      if (idx < 0) idx = 0;
      values[static_cast<std::size_t>(i * max_col + j)] /= sum[static_cast<std::size_t>(idx)];
    }
  count_stochastic_ = false;
  invalidate_device();
}

// ---------------------------------------------------------------- DENSE ----

// MatrixDENSE.cpp:11-28 (duplicates accumulate)
MatrixDENSE::MatrixDENSE(const MatrixCOO &m) {
  n_row = m.get_n_row();
  n_col = m.get_n_col();
  ln_row = m.get_n_row();
  ln_col = m.get_n_col();
  f_row = m.get_f_row();
  f_col = m.get_f_col();
  pr = m.get_pr();
  pc = m.get_pc();
  NR = m.get_NR();
  NC = m.get_NC();
  nnz = m.get_nnz();
  values.assign(static_cast<std::size_t>(ln_row) * ln_col, 0);
  const auto &r = m.get_row();
  const auto &c = m.get_col();
  const auto &v = m.get_values();
  for (std::size_t i = 0; i < v.size(); i++)
    values[static_cast<std::size_t>(r[i]) * ln_col + static_cast<std::size_t>(c[i])] += v[i];
}

std::ostream &MatrixDENSE::print(std::ostream &os) const {
  host::print_header(os, "----- DENSE -----", n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::print_dense_matrix(ln_row, ln_col, values, os);
  os << "-----------------" << std::endl << std::flush;
  return os;
}

std::ostream &MatrixDENSE::print_as_dense(std::ostream &os) { return print(os); }

std::ostream &operator<<(std::ostream &os, const MatrixDENSE &m) { return m.print(os); }

std::vector<double> MatrixDENSE::spmv(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r(static_cast<std::size_t>(n_row), 0);
  this->Ax(r, v, vect_incr);
  return r;
}

// MatrixDENSE.cpp:64-73 on the GPU (dense_kernel)
void MatrixDENSE::Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr) const {
  hip::Ax_host(*this, r, v, vect_incr);
}

std::ostream &MatrixDENSE::print_infos(std::ostream &os) {
  os << "-----------------" << std::endl;
  os << "----- DENSE -----" << std::endl;
  os << "--- general   ---" << std::endl;
  os << "n_row : " << n_row << std::endl;
  os << "n_col : " << n_col << std::endl;
  os << "nnz : " << nnz << std::endl;
  os << "--- capacity  ---" << std::endl;
  os << "values : " << values.capacity() << std::endl;
  os << "--- size      ---" << std::endl;
  os << "values : " << values.size() << std::endl;
  os << "-----------------" << std::endl;
  return os;
}

std::ostream &MatrixDENSE::print_stats(std::ostream &os) {
  os << "upper values : " << 0 << std::endl;
  os << "lower values : " << 0 << std::endl;
  os << "diag  values : " << 0 << std::endl;
  return os;
}

std::ostream &MatrixDENSE::write(std::ostream &os) {
  host::write_header(os, n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::write_array(os, values);
  return os;
}

std::istream &MatrixDENSE::read(std::istream &is, std::size_t, std::size_t) {
  // read into temporaries: a truncated or corrupt file throws and leaves
  // the matrix as it was
  int h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_pr, h_pc, h_NR, h_NC;
  long h_nnz;
  host::read_header(is, h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_nnz, h_pr, h_pc, h_NR, h_NC);
  std::vector<double> r_values;
  host::read_array(is, r_values);
  n_row = h_nr; n_col = h_nc; ln_row = h_lnr; ln_col = h_lnc; f_row = h_fr; f_col = h_fc;
  nnz = h_nnz; pr = h_pr; pc = h_pc; NR = h_NR; NC = h_NC;
  values.swap(r_values);
  count_stochastic_ = false;
  invalidate_device();
  return is;
}

// MatrixDENSE.cpp:137-190
void MatrixDENSE::fill_cdiag(int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  const long nv = host::cdiag_block_count(f_row, ln_row, f_col, ln_col, cdiag);
  if (nv == 0) return;
  nnz = nv;
  values.assign(static_cast<std::size_t>(ln_col) * ln_row, 0);
  auto inb = [&](long ii, long jj) {
    return ii >= f_row && ii < static_cast<long>(f_row) + ln_row && jj >= f_col && jj < static_cast<long>(f_col) + ln_col;
  };
  for (long i = f_row; i < static_cast<long>(f_row) + ln_row; i++) {
    if (inb(i, i - cdiag)) values[static_cast<std::size_t>((i - f_row) * ln_col + (i - cdiag - f_col))] = 1;
    if (cdiag != 0 && inb(i, i + cdiag)) values[static_cast<std::size_t>((i - f_row) * ln_col + (i + cdiag - f_col))] = 1;
  }
}

void MatrixDENSE::fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                             int NC) {
  fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
}

// MatrixDENSE.cpp:192-271 (duplicates accumulate in generation order)
void MatrixDENSE::fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                    int NR, int NC, std::uint64_t value_seed) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, value_seed, f_row, ln_row, f_col, ln_col);
  if (b.candidates_in_block() == 0) return;
  const long lim3 = b.lim3();
  host::RowEntries e = host::generate_rows(b, f_row, std::max<long>(lim3, f_row), false);
  nnz = static_cast<long>(e.vals.size());
  values.assign(static_cast<std::size_t>(ln_col) * ln_row, 0);
  std::size_t src = 0;
  for (long r = f_row; r < std::max<long>(lim3, f_row); r++) {
    const int cnt = e.counts[static_cast<std::size_t>(r - f_row)];
    for (int k = 0; k < cnt; k++, src++)
      values[static_cast<std::size_t>((r - f_row) * ln_col + (e.cols[src] - f_col))] += e.vals[src];
  }
}

// MatrixDENSE.cpp:273-286 (empty columns divide by 1)
void MatrixDENSE::fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                        int NR, int NC) {
  fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
  if (values.empty()) return;
  std::vector<double> sum = gen::cqmat_column_sums(n_row, n_col, c, q, seed_mult);
  for (auto &s : sum)
    if (s == 0) s = 1;
  gen::parallel_for(ln_row, [&](long b, long e) {
    for (long i = b; i < e; i++)
      for (long j = 0; j < ln_col; j++) values[static_cast<std::size_t>(i * ln_col + j)] /= sum[static_cast<std::size_t>(f_col + j)];
  });
  count_stochastic_ = true;
}

// MatrixDENSE.cpp:288-300
void MatrixDENSE::normalize_columns() {
  std::vector<double> sum(static_cast<std::size_t>(ln_col), 0);
  for (long i = 0; i < ln_row; i++)
    for (long j = 0; j < ln_col; j++) sum[static_cast<std::size_t>(j)] += values[static_cast<std::size_t>(i * ln_col + j)];
  for (long i = 0; i < ln_row; i++)
    for (long j = 0; j < ln_col; j++) values[static_cast<std::size_t>(i * ln_col + j)] /= sum[static_cast<std::size_t>(j)];
  count_stochastic_ = false;
  invalidate_device();
}

}  // namespace cpp
}  // namespace tbsla
