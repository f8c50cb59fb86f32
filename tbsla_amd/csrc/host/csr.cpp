// tbsla::cpp::MatrixCSR (reference: src/cpp/matrix/MatrixCSR.cpp).
#include <tbsla/cpp/MatrixCSR.hpp>

#include <algorithm>
#include <numeric>

#include "../device/products.hpp"
#include "fill_common.hpp"
#include "../device/convert.hpp"
#include "../device/generate.hpp"

namespace tbsla {
namespace cpp {

// MatrixCSR.cpp:13-27 (nnz, left unset there, is the number of values here)
MatrixCSR::MatrixCSR(int n_row, int n_col, std::vector<double> &values, std::vector<int> &rowptr,
                     std::vector<int> &colidx) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->values = values;
  this->rowptr = rowptr;
  this->colidx = colidx;
  this->ln_row = n_row;
  this->ln_col = n_col;
  this->f_row = 0;
  this->f_col = 0;
  this->pr = 0;
  this->pc = 0;
  this->NR = 1;
  this->NC = 1;
  this->nnz = static_cast<long>(values.size());
}

// MatrixCSR.cpp:29-74: (row, col) order, rowptr from row counts.  Entries
// with equal (row, col) keep their COO order (the reference's std::sort
// leaves their order unspecified).
MatrixCSR::MatrixCSR(const MatrixCOO &m) { from_coo(m, false); }

MatrixCSR::MatrixCSR(const MatrixCOO &m, bool gpu) { from_coo(m, gpu); }

// MatrixCSR.cpp:29-74; the (row, column) order of the entries is computed on
// the host or, with `gpu`, on the current GPU (the same order)
void MatrixCSR::from_coo(const MatrixCOO &m, bool gpu) {
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
  if (gpu) {
    if (r.size() != v.size() || c.size() != v.size())
      throw std::invalid_argument("tbsla: GPU conversion needs row, col and values of one length");
    hip::coo_to_csr_gpu(n_row, r, c, v, rowptr, colidx, values);
    return;
  }
  const std::size_t n = v.size();
  rowptr.assign(static_cast<std::size_t>(n_row) + 1, 0);
  if (n == 0) return;
  std::vector<int> p = host::row_col_order(n_row, r, c);
  colidx.resize(n);
  values.resize(n);
  for (std::size_t k = 0; k < n; k++) {
    colidx[k] = c[static_cast<std::size_t>(p[k])];
    values[k] = v[static_cast<std::size_t>(p[k])];
    rowptr[static_cast<std::size_t>(r[static_cast<std::size_t>(p[k])]) + 1]++;
  }
  for (int i = 0; i < n_row; i++) rowptr[static_cast<std::size_t>(i) + 1] += rowptr[static_cast<std::size_t>(i)];
}

std::ostream &operator<<(std::ostream &os, const MatrixCSR &m) { return m.print(os); }

std::ostream &MatrixCSR::print(std::ostream &os) const {
  host::print_header(os, "------ CSR ------", n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::streamvector<double>(os, "values", values);
  os << std::endl;
  host::streamvector<int>(os, "rowptr", rowptr);
  os << std::endl;
  host::streamvector<int>(os, "colidx", colidx);
  os << std::endl;
  os << "-----------------" << std::endl << std::flush;
  return os;
}

std::ostream &MatrixCSR::print_as_dense(std::ostream &os) {
  std::vector<double> d(static_cast<std::size_t>(n_row) * n_col, 0);
  if (!values.empty())
    for (std::size_t i = 0; i + 1 < rowptr.size(); i++)
      for (int j = rowptr[i] - rowptr.front(); j < rowptr[i + 1] - rowptr.front(); j++)
        d[i * static_cast<std::size_t>(n_col) + static_cast<std::size_t>(colidx[static_cast<std::size_t>(j)])] +=
            values[static_cast<std::size_t>(j)];
  host::print_dense_matrix(n_row, n_col, d, os);
  return os;
}

std::vector<double> MatrixCSR::spmv(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r(static_cast<std::size_t>(n_row), 0);
  this->Ax(r, v, vect_incr);
  return r;
}

// MatrixCSR.cpp:129-140 on the GPU (csr_tile_kernel)
void MatrixCSR::Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr) const {
  hip::Ax_host(*this, r, v, vect_incr);
}

std::ostream &MatrixCSR::print_infos(std::ostream &os) {
  os << "-----------------" << std::endl;
  os << "------ CSR ------" << std::endl;
  os << "--- general   ---" << std::endl;
  os << "n_row : " << n_row << std::endl;
  os << "n_col : " << n_col << std::endl;
  os << "--- capacity  ---" << std::endl;
  os << "values : " << values.capacity() << std::endl;
  os << "rowptr : " << rowptr.capacity() << std::endl;
  os << "colidx : " << colidx.capacity() << std::endl;
  os << "--- size      ---" << std::endl;
  os << "values : " << values.size() << std::endl;
  os << "rowptr : " << rowptr.size() << std::endl;
  os << "colidx : " << colidx.size() << std::endl;
  os << "-----------------" << std::endl;
  return os;
}

std::ostream &MatrixCSR::print_stats(std::ostream &os) {
  int s = 0, u = 0, d = 0;
  for (std::size_t i = 0; i + 1 < rowptr.size(); i++)
    for (int j = rowptr[i] - rowptr.front(); j < rowptr[i + 1] - rowptr.front(); j++) {
      const long c = colidx[static_cast<std::size_t>(j)];
      if (static_cast<long>(i) < c) s++;
      else if (static_cast<long>(i) > c) u++;
      else d++;
    }
  os << "upper values : " << u << std::endl;
  os << "lower values : " << s << std::endl;
  os << "diag  values : " << d << std::endl;
  return os;
}

std::ostream &MatrixCSR::write(std::ostream &os) {
  host::write_header(os, n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::write_array(os, values);
  host::write_array(os, rowptr);
  host::write_array(os, colidx);
  return os;
}

std::istream &MatrixCSR::read(std::istream &is, std::size_t, std::size_t) {
  // read into temporaries: a truncated or corrupt file throws and leaves
  // the matrix as it was
  int h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_pr, h_pc, h_NR, h_NC;
  long h_nnz;
  host::read_header(is, h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_nnz, h_pr, h_pc, h_NR, h_NC);
  std::vector<double> r_values;
  host::read_array(is, r_values);
  std::vector<int> r_rowptr;
  host::read_array(is, r_rowptr);
  std::vector<int> r_colidx;
  host::read_array(is, r_colidx);
  n_row = h_nr; n_col = h_nc; ln_row = h_lnr; ln_col = h_lnc; f_row = h_fr; f_col = h_fc;
  nnz = h_nnz; pr = h_pr; pc = h_pc; NR = h_NR; NC = h_NC;
  values.swap(r_values);
  rowptr.swap(r_rowptr);
  colidx.swap(r_colidx);
  count_stochastic_ = false;
  invalidate_device();
  return is;
}

// MatrixCSR.cpp:234-297
void MatrixCSR::fill_cdiag(int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  colidx.clear();
  rowptr.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  const long nv = host::cdiag_block_count(f_row, ln_row, f_col, ln_col, cdiag);
  if (nv == 0) return;
  auto inb = [&](long ii, long jj) {
    return ii >= f_row && ii < static_cast<long>(f_row) + ln_row && jj >= f_col && jj < static_cast<long>(f_col) + ln_col;
  };
  values.reserve(static_cast<std::size_t>(nv));
  colidx.reserve(static_cast<std::size_t>(nv));
  rowptr.reserve(static_cast<std::size_t>(ln_row) + 1);
  int incr = 0;
  rowptr.push_back(incr);
  for (long i = f_row; i < static_cast<long>(f_row) + ln_row; i++) {
    if (inb(i, i - cdiag)) {
      colidx.push_back(static_cast<int>(i - cdiag));
      values.push_back(1);
      incr++;
    }
    if (cdiag != 0 && inb(i, i + cdiag)) {
      colidx.push_back(static_cast<int>(i + cdiag));
      values.push_back(1);
      incr++;
    }
    rowptr.push_back(incr);
  }
}

void MatrixCSR::fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                           int NC) {
  fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
}

// MatrixCSR.cpp:299-409, rows generated in parallel
void MatrixCSR::fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                  int NR, int NC, std::uint64_t value_seed) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  colidx.clear();
  rowptr.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  host::RowEntries e = host::generate_rows(b, f_row, std::max<long>(lim3, f_row), false);
  long kept_full = 0;
  const long full_end = std::min(lim3, b.s.min_);
  for (long r = f_row; r < full_end; r++) kept_full += e.counts[static_cast<std::size_t>(r - f_row)];
  nnz = host::overcount_nnz(b, kept_full);
  if (b.candidates_in_block() == 0) return;
  rowptr.resize(e.counts.size() + 1);
  rowptr[0] = 0;
  for (std::size_t k = 0; k < e.counts.size(); k++) rowptr[k + 1] = rowptr[k] + e.counts[k];
  colidx = std::move(e.cols);
  values = std::move(e.vals);
}

// fill_cqmat_values / fill_cqmat_stochastic with the block's entries
// generated on the GPU (device/generate.hip); metadata, nnz and the early
// returns follow the host path above.
void MatrixCSR::fill_cqmat_gpu(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                               int NC, std::uint64_t value_seed, bool stochastic) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  colidx.clear();
  rowptr.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  // fill_cqmat_stochastic fills with value 1.0 (no overlay), then divides
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, stochastic ? 0 : value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  const long nrows = std::max<long>(lim3, f_row) - f_row;
  hip::GeneratedCsr g;
  hip::generate_cqmat_csr(b, nrows, stochastic, g);
  const long full_end = std::min(lim3, b.s.min_);
  const long kept_full = full_end > f_row ? g.rowptr[static_cast<std::size_t>(full_end - f_row)] : 0;
  nnz = host::overcount_nnz(b, kept_full);
  if (b.candidates_in_block() == 0) return;
  rowptr = std::move(g.rowptr);
  colidx = std::move(g.colidx);
  values = std::move(g.values);
  if (stochastic && !values.empty()) count_stochastic_ = true;
}

// MatrixCSR.cpp:411-421
void MatrixCSR::fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                      int NR, int NC) {
  fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
  if (values.empty()) return;
  std::vector<double> sum = gen::cqmat_column_sums(n_row, n_col, c, q, seed_mult);
  gen::parallel_for(static_cast<long>(rowptr.size()) - 1, [&](long b, long e) {
    for (long i = b; i < e; i++)
      for (long j = rowptr[static_cast<std::size_t>(i)]; j < rowptr[static_cast<std::size_t>(i) + 1]; j++)
        values[static_cast<std::size_t>(j)] /= sum[static_cast<std::size_t>(colidx[static_cast<std::size_t>(j)])];
  });
  count_stochastic_ = true;
}

// MatrixCSR.cpp:423-435 (sequential column sums, entry order)
void MatrixCSR::normalize_columns() {
  std::vector<double> sum(static_cast<std::size_t>(ln_col), 0);
  for (std::size_t i = 0; i + 1 < rowptr.size(); i++)
    for (long j = rowptr[i]; j < rowptr[i + 1]; j++)
      sum[static_cast<std::size_t>(colidx[static_cast<std::size_t>(j)] - f_col)] += values[static_cast<std::size_t>(j)];
  for (std::size_t i = 0; i + 1 < rowptr.size(); i++)
    for (long j = rowptr[i]; j < rowptr[i + 1]; j++)
      values[static_cast<std::size_t>(j)] /= sum[static_cast<std::size_t>(colidx[static_cast<std::size_t>(j)] - f_col)];
  count_stochastic_ = false;
  invalidate_device();
}

}  // namespace cpp
}  // namespace tbsla
