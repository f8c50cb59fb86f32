// tbsla::cpp::MatrixCOO and MatrixSCOO (reference: src/cpp/matrix/MatrixCOO.cpp,
// src/cpp/matrix/MatrixSCOO.cpp).
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>

#include <algorithm>
#include <sstream>

#include "../device/products.hpp"
#include "fill_common.hpp"
#include "../device/generate.hpp"

namespace tbsla {
namespace cpp {

// ------------------------------------------------------------------ COO ----

void MatrixCOO::init(int n_row, int n_col, long int n_values) {
  this->n_row = n_row;
  this->n_col = n_col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
  values.reserve(static_cast<std::size_t>(n_values));
  row.reserve(static_cast<std::size_t>(n_values));
  col.reserve(static_cast<std::size_t>(n_values));
}

MatrixCOO::MatrixCOO(int n_row, int n_col, std::vector<double> &values, std::vector<int> &row, std::vector<int> &col) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->values = values;
  this->row = row;
  this->col = col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
  nnz = static_cast<long>(values.size());
}

MatrixCOO::MatrixCOO(int n_row, int n_col, long int n_values) { init(n_row, n_col, n_values); }

MatrixCOO::MatrixCOO(int n_row, int n_col) {
  this->n_row = n_row;
  this->n_col = n_col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
}

std::ostream &operator<<(std::ostream &os, const MatrixCOO &m) { return m.print(os); }

std::ostream &MatrixCOO::print_as_dense(std::ostream &os) {
  std::vector<double> d(static_cast<std::size_t>(n_row) * n_col, 0);
  for (std::size_t i = 0; i < row.size(); i++)
    d[static_cast<std::size_t>(row[i]) * n_col + static_cast<std::size_t>(col[i])] += values[i];
  host::print_dense_matrix(n_row, n_col, d, os);
  return os;
}

std::ostream &MatrixCOO::print(std::ostream &os) const {
  host::print_header(os, "------ COO ------", n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::streamvector<double>(os, "val", values);
  os << std::endl;
  host::streamvector<int>(os, "row", row);
  os << std::endl;
  host::streamvector<int>(os, "col", col);
  os << std::endl;
  os << "-----------------" << std::endl << std::flush;
  return os;
}

std::vector<double> MatrixCOO::spmv(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r(static_cast<std::size_t>(n_row), 0);
  this->Ax(r, v, vect_incr);
  return r;
}

// MatrixCOO.cpp:107-116 on the GPU (seg_kernel, each row in entry order)
void MatrixCOO::Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr) const {
  hip::Ax_host(*this, r, v, vect_incr);
}

// MatrixCOO.cpp:118-126
void MatrixCOO::push_back(int r, int c, double v) {
  if (r >= n_row || r < 0) return;
  if (c >= n_col || c < 0) return;
  values.push_back(v);
  row.push_back(r);
  col.push_back(c);
  invalidate_device();
}

void MatrixCOO::update_nnz() {
  values.shrink_to_fit();
  col.shrink_to_fit();
  row.shrink_to_fit();
  nnz = static_cast<long>(values.size());
}

std::ostream &MatrixCOO::print_infos(std::ostream &os) {
  os << "-----------------" << std::endl;
  os << "------ COO ------" << std::endl;
  os << "--- general   ---" << std::endl;
  os << "n_row : " << n_row << std::endl;
  os << "n_col : " << n_col << std::endl;
  os << "--- capacity  ---" << std::endl;
  os << "values : " << values.capacity() << std::endl;
  os << "row : " << row.capacity() << std::endl;
  os << "col : " << col.capacity() << std::endl;
  os << "--- size      ---" << std::endl;
  os << "values : " << values.size() << std::endl;
  os << "row : " << row.size() << std::endl;
  os << "col : " << col.size() << std::endl;
  os << "-----------------" << std::endl;
  return os;
}

static std::ostream &coo_stats(std::ostream &os, const std::vector<int> &row, const std::vector<int> &col) {
  int s = 0, u = 0, d = 0;
  if (row.size() != col.size()) {
    std::cerr << "Err \n";
    return os;
  }
  for (std::size_t i = 0; i < row.size(); i++) {
    if (row[i] < col[i]) s++;
    else if (row[i] > col[i]) u++;
    else d++;
  }
  os << "upper values : " << u << std::endl;
  os << "lower values : " << s << std::endl;
  os << "diag  values : " << d << std::endl;
  return os;
}

std::ostream &MatrixCOO::print_stats(std::ostream &os) { return coo_stats(os, row, col); }

std::ostream &MatrixCOO::write(std::ostream &os) {
  host::write_header(os, n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::write_array(os, values);
  host::write_array(os, row);
  host::write_array(os, col);
  return os;
}

std::istream &MatrixCOO::read(std::istream &is, std::size_t, std::size_t) {
  // read into temporaries: a truncated or corrupt file throws and leaves
  // the matrix as it was
  int h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_pr, h_pc, h_NR, h_NC;
  long h_nnz;
  host::read_header(is, h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_nnz, h_pr, h_pc, h_NR, h_NC);
  std::vector<double> r_values;
  host::read_array(is, r_values);
  std::vector<int> r_row;
  host::read_array(is, r_row);
  std::vector<int> r_col;
  host::read_array(is, r_col);
  n_row = h_nr; n_col = h_nc; ln_row = h_lnr; ln_col = h_lnc; f_row = h_fr; f_col = h_fc;
  nnz = h_nnz; pr = h_pr; pc = h_pc; NR = h_NR; NC = h_NC;
  values.swap(r_values);
  row.swap(r_row);
  col.swap(r_col);
  count_stochastic_ = false;
  invalidate_device();
  return is;
}

// MatrixCOO.cpp:230-292 (MatrixMarket coordinate real general/symmetric)
void MatrixCOO::readMM(std::string fname) {
  std::ifstream is(fname);
  std::string line;
  if (!is.is_open()) return;
  std::getline(is, line);
  std::istringstream hs(line);
  std::vector<std::string> sp;
  for (std::string w; hs >> w;) sp.push_back(w);
  std::cout << line << "\n";
  if (sp.size() < 5 || sp[0] != "%%MatrixMarket" || sp[1] != "matrix" || sp[2] != "coordinate" || sp[3] != "real")
    throw MatrixFormatReadException();
  const bool symmetric = sp[4] == "symmetric";
  if (!symmetric && sp[4] != "general") throw MatrixFormatReadException();
  while (line.rfind("%", 0) == 0) std::getline(is, line);
  std::cout << line << "\n";
  std::istringstream ss(line);
  int nr = 0, nc = 0, nv = 0;
  ss >> nr >> nc >> nv;
  values.clear();
  row.clear();
  col.clear();
  init(nr, nc, symmetric ? 2L * nv : nv);
  for (int i = 0; i < nv && !is.eof(); i++) {
    int r, c;
    double v;
    is >> r >> c >> v;
    push_back(r - 1, c - 1, v);
    if (symmetric && r != c) push_back(c - 1, r - 1, v);
  }
  update_nnz();
  count_stochastic_ = false;
  invalidate_device();
}

// MatrixCOO.cpp:294-330: 1-D split of the cdiag_value stream over NR*NC ranks
void MatrixCOO::fill_cdiag(int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = NR;
  ln_col = NC;
  f_row = 0;
  f_col = 0;
  long gnv = std::max(std::min<long>(n_row, static_cast<long>(n_col) - cdiag), 0L) +
             std::max(std::min<long>(static_cast<long>(n_row) - cdiag, n_col), 0L);
  if (cdiag == 0) gnv /= 2;
  nnz = 0;
  if (gnv == 0) return;
  const long s = static_cast<long>(gen::pflv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  const long n = static_cast<long>(gen::lnv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  nnz = n;
  values.reserve(static_cast<std::size_t>(n));
  col.reserve(static_cast<std::size_t>(n));
  row.reserve(static_cast<std::size_t>(n));
  for (int i = static_cast<int>(s); i < s + n; i++) {
    std::uint64_t r, c;
    gen::cdiag_value(static_cast<std::size_t>(i), static_cast<std::size_t>(n_row), static_cast<std::size_t>(n_col),
                     static_cast<std::size_t>(cdiag), r, c);
    push_back(static_cast<int>(r), static_cast<int>(c), 1);
  }
}

void MatrixCOO::fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                           int NC) {
  fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
}

// MatrixCOO.cpp:332-375: entries [s, s+n) of the global stream, in parallel
void MatrixCOO::fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                  int NR, int NC, std::uint64_t value_seed) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = NR;
  ln_col = NC;
  f_row = 0;
  f_col = 0;
  const gen::CqmatStream st(n_row, n_col, c);
  long gnv = 0;
  for (long i = 0; i < std::min<long>(static_cast<long>(n_col) - st.cc + 1, n_row); i++) gnv += st.cc;
  for (long i = 0; i < std::min<long>(n_row, n_col) - std::min<long>(static_cast<long>(n_col) - st.cc + 1, n_row); i++)
    gnv += st.cc - i - 1;
  if (gnv == 0) return;
  const long s = static_cast<long>(gen::pflv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  const long n = static_cast<long>(gen::lnv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  nnz = n;
  values.resize(static_cast<std::size_t>(n));
  row.resize(static_cast<std::size_t>(n));
  col.resize(static_cast<std::size_t>(n));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, value_seed, 0, n_row, 0, n_col);
  gen::parallel_for(n, [&](long lo, long hi) {
    for (long k = lo; k < hi; k++) {
      long ii, jj;
      b.entry(s + k, ii, jj);
      row[static_cast<std::size_t>(k)] = static_cast<int>(ii);
      col[static_cast<std::size_t>(k)] = static_cast<int>(jj);
      values[static_cast<std::size_t>(k)] = b.value(s + k);
    }
  }, 1 << 14);
}

// fill_cqmat_values / fill_cqmat_stochastic with the slice's entries
// generated on the GPU (device/generate.hip), one thread per stream entry;
// metadata and nnz follow the host path above.
void MatrixCOO::fill_cqmat_gpu(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                               int NC, std::uint64_t value_seed, bool stochastic) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = stochastic;  // as the host path, which sets it whatever the slice holds
  invalidate_device();
  ln_row = NR;
  ln_col = NC;
  f_row = 0;
  f_col = 0;
  const gen::CqmatStream st(n_row, n_col, c);
  long gnv = 0;
  for (long i = 0; i < std::min<long>(static_cast<long>(n_col) - st.cc + 1, n_row); i++) gnv += st.cc;
  for (long i = 0; i < std::min<long>(n_row, n_col) - std::min<long>(static_cast<long>(n_col) - st.cc + 1, n_row); i++)
    gnv += st.cc - i - 1;
  if (gnv == 0) return;
  const long s = static_cast<long>(gen::pflv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  const long n = static_cast<long>(gen::lnv(static_cast<std::size_t>(gnv), pr * NC + pc, NR * NC));
  nnz = n;
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, stochastic ? 0 : value_seed, 0, n_row, 0, n_col);
  hip::generate_cqmat_coo(b, s, n, stochastic, row, col, values);
}

// MatrixCOO.cpp:377-383
void MatrixCOO::fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                      int NR, int NC) {
  fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
  std::vector<double> sum = gen::cqmat_column_sums(n_row, n_col, c, q, seed_mult);
  gen::parallel_for(static_cast<long>(values.size()), [&](long b, long e) {
    for (long i = b; i < e; i++) values[static_cast<std::size_t>(i)] /= sum[static_cast<std::size_t>(col[static_cast<std::size_t>(i)])];
  });
  count_stochastic_ = true;
}

// MatrixCOO.cpp:385-393
void MatrixCOO::normalize_columns() {
  std::vector<double> sum(static_cast<std::size_t>(n_col), 0);
  for (std::size_t i = 0; i < values.size(); i++) sum[static_cast<std::size_t>(col[i])] += values[i];
  for (std::size_t i = 0; i < values.size(); i++) values[i] /= sum[static_cast<std::size_t>(col[i])];
  count_stochastic_ = false;
  invalidate_device();
}

// ----------------------------------------------------------------- SCOO ----

MatrixSCOO::MatrixSCOO(int n_row, int n_col, std::vector<double> &values, std::vector<int> &row,
                       std::vector<int> &col) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->values = values;
  this->row = row;
  this->col = col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
  nnz = static_cast<long>(values.size());
}

// MatrixSCOO.cpp:28-40 (ln_row/ln_col are left unset there; set to the full size)
MatrixSCOO::MatrixSCOO(int n_row, int n_col, int n_values) {
  this->n_row = n_row;
  this->n_col = n_col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
  values.reserve(static_cast<std::size_t>(n_values));
  row.reserve(static_cast<std::size_t>(n_values));
  col.reserve(static_cast<std::size_t>(n_values));
}

MatrixSCOO::MatrixSCOO(int n_row, int n_col) {
  this->n_row = n_row;
  this->n_col = n_col;
  ln_row = n_row;
  ln_col = n_col;
  f_row = f_col = pr = pc = 0;
  NR = NC = 1;
}

// MatrixSCOO.cpp:53-68
MatrixSCOO::MatrixSCOO(const MatrixCOO &m) {
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
  row = m.get_row();
  col = m.get_col();
  values = m.get_values();
  count_stochastic_ = m.count_stochastic();
}

std::ostream &operator<<(std::ostream &os, const MatrixSCOO &m) { return m.print(os); }

std::ostream &MatrixSCOO::print_as_dense(std::ostream &os) {
  std::vector<double> d(static_cast<std::size_t>(n_row) * n_col, 0);
  for (std::size_t i = 0; i < row.size(); i++)
    d[static_cast<std::size_t>(row[i]) * n_col + static_cast<std::size_t>(col[i])] += values[i];
  host::print_dense_matrix(n_row, n_col, d, os);
  return os;
}

std::ostream &MatrixSCOO::print(std::ostream &os) const {
  host::print_header(os, "----- SCOO ------", n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::streamvector<double>(os, "val", values);
  os << std::endl;
  host::streamvector<int>(os, "row", row);
  os << std::endl;
  host::streamvector<int>(os, "col", col);
  os << std::endl;
  os << "-----------------" << std::endl << std::flush;
  return os;
}

std::vector<double> MatrixSCOO::spmv(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r(static_cast<std::size_t>(n_row), 0);
  this->Ax(r, v, vect_incr);
  return r;
}

// MatrixSCOO.cpp:114-123 on the GPU (seg_kernel)
void MatrixSCOO::Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr) const {
  hip::Ax_host(*this, r, v, vect_incr);
}

void MatrixSCOO::push_back(int r, int c, double v) {
  if (r >= n_row || r < 0) return;
  if (c >= n_col || c < 0) return;
  values.push_back(v);
  row.push_back(r);
  col.push_back(c);
  invalidate_device();
}

std::ostream &MatrixSCOO::print_infos(std::ostream &os) {
  os << "-----------------" << std::endl;
  os << "----- SCOO ------" << std::endl;
  os << "--- general   ---" << std::endl;
  os << "n_row : " << n_row << std::endl;
  os << "n_col : " << n_col << std::endl;
  os << "--- capacity  ---" << std::endl;
  os << "values : " << values.capacity() << std::endl;
  os << "row : " << row.capacity() << std::endl;
  os << "col : " << col.capacity() << std::endl;
  os << "--- size      ---" << std::endl;
  os << "values : " << values.size() << std::endl;
  os << "row : " << row.size() << std::endl;
  os << "col : " << col.size() << std::endl;
  os << "-----------------" << std::endl;
  return os;
}

std::ostream &MatrixSCOO::print_stats(std::ostream &os) { return coo_stats(os, row, col); }

std::ostream &MatrixSCOO::write(std::ostream &os) {
  host::write_header(os, n_row, n_col, ln_row, ln_col, f_row, f_col, nnz, pr, pc, NR, NC);
  host::write_array(os, values);
  host::write_array(os, row);
  host::write_array(os, col);
  return os;
}

std::istream &MatrixSCOO::read(std::istream &is, std::size_t, std::size_t) {
  // read into temporaries: a truncated or corrupt file throws and leaves
  // the matrix as it was
  int h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_pr, h_pc, h_NR, h_NC;
  long h_nnz;
  host::read_header(is, h_nr, h_nc, h_lnr, h_lnc, h_fr, h_fc, h_nnz, h_pr, h_pc, h_NR, h_NC);
  std::vector<double> r_values;
  host::read_array(is, r_values);
  std::vector<int> r_row;
  host::read_array(is, r_row);
  std::vector<int> r_col;
  host::read_array(is, r_col);
  n_row = h_nr; n_col = h_nc; ln_row = h_lnr; ln_col = h_lnc; f_row = h_fr; f_col = h_fc;
  nnz = h_nnz; pr = h_pr; pc = h_pc; NR = h_NR; NC = h_NC;
  values.swap(r_values);
  row.swap(r_row);
  col.swap(r_col);
  count_stochastic_ = false;
  invalidate_device();
  return is;
}

// MatrixSCOO.cpp:230-286 (nnz is never set there; it is set to the entry count)
void MatrixSCOO::fill_cdiag(int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  const long nv = host::cdiag_block_count(f_row, ln_row, f_col, ln_col, cdiag);
  nnz = nv;
  if (nv == 0) return;
  auto inb = [&](long ii, long jj) {
    return ii >= f_row && ii < static_cast<long>(f_row) + ln_row && jj >= f_col && jj < static_cast<long>(f_col) + ln_col;
  };
  for (long i = f_row; i < static_cast<long>(f_row) + ln_row; i++) {
    if (inb(i, i - cdiag)) push_back(static_cast<int>(i), static_cast<int>(i - cdiag), 1);
    if (cdiag != 0 && inb(i, i + cdiag)) push_back(static_cast<int>(i), static_cast<int>(i + cdiag), 1);
  }
}

void MatrixSCOO::fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                            int NC) {
  fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
}

// MatrixSCOO.cpp:288-391
void MatrixSCOO::fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                   int NR, int NC, std::uint64_t value_seed) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  host::RowEntries e = host::generate_rows(b, f_row, std::max<long>(lim3, f_row), true);
  long kept_full = 0;
  const long full_end = std::min(lim3, b.s.min_);
  for (long r = f_row; r < full_end; r++) kept_full += e.counts[static_cast<std::size_t>(r - f_row)];
  nnz = host::overcount_nnz(b, kept_full);
  if (b.candidates_in_block() == 0) return;
  row = std::move(e.rows);
  col = std::move(e.cols);
  values = std::move(e.vals);
}

// fill_cqmat_values / fill_cqmat_stochastic with the block's entries
// generated on the GPU (device/generate.hip); metadata, nnz and the early
// return follow the host paths.
void MatrixSCOO::fill_cqmat_gpu(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                                int NC, std::uint64_t value_seed, bool stochastic) {
  this->n_row = n_row;
  this->n_col = n_col;
  this->pr = pr;
  this->pc = pc;
  this->NR = NR;
  this->NC = NC;
  values.clear();
  col.clear();
  row.clear();
  count_stochastic_ = false;
  invalidate_device();
  ln_row = static_cast<int>(gen::lnv(n_row, pr, NR));
  f_row = static_cast<int>(gen::pflv(n_row, pr, NR));
  ln_col = static_cast<int>(gen::lnv(n_col, pc, NC));
  f_col = static_cast<int>(gen::pflv(n_col, pc, NC));
  host::CqmatBlock b(n_row, n_col, c, q, seed_mult, stochastic ? 0 : value_seed, f_row, ln_row, f_col, ln_col);
  const long lim3 = b.lim3();
  std::vector<int> rp;
  hip::generate_cqmat_scoo(b, std::max<long>(lim3, f_row) - f_row, stochastic, rp, row, col, values);
  const long full_end = std::min(lim3, b.s.min_);
  nnz = host::overcount_nnz(b, full_end > f_row ? rp[static_cast<std::size_t>(full_end - f_row)] : 0);
  if (b.candidates_in_block() == 0) {
    row.clear();
    col.clear();
    values.clear();
  }
  count_stochastic_ = stochastic;  // as the host path, which sets it whatever the block holds
}

// MatrixSCOO.cpp:393-399
void MatrixSCOO::fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc,
                                       int NR, int NC) {
  fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC);
  std::vector<double> sum = gen::cqmat_column_sums(n_row, n_col, c, q, seed_mult);
  gen::parallel_for(static_cast<long>(values.size()), [&](long b, long e) {
    for (long i = b; i < e; i++) values[static_cast<std::size_t>(i)] /= sum[static_cast<std::size_t>(col[static_cast<std::size_t>(i)])];
  });
  count_stochastic_ = true;
}

// MatrixSCOO.cpp:401-409
void MatrixSCOO::normalize_columns() {
  std::vector<double> sum(static_cast<std::size_t>(ln_col), 0);
  for (std::size_t i = 0; i < values.size(); i++) sum[static_cast<std::size_t>(col[i] - f_col)] += values[i];
  for (std::size_t i = 0; i < values.size(); i++) values[i] /= sum[static_cast<std::size_t>(col[i] - f_col)];
  count_stochastic_ = false;
  invalidate_device();
}

}  // namespace cpp
}  // namespace tbsla
