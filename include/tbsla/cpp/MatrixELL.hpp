// tbsla::cpp::MatrixELL -- reference API: src/cpp/include/tbsla/cpp/MatrixELL.hpp:12-44.
// Host storage is row-major (i * max_col + j) like the reference; the device
// mirror is column-major.
#ifndef TBSLA_CPP_MatrixELL
#define TBSLA_CPP_MatrixELL

#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/cpp/MatrixCOO.hpp>

#include <iostream>
#include <vector>

namespace tbsla {
namespace cpp {

class MatrixELL : public virtual Matrix {
 public:
  friend std::ostream &operator<<(std::ostream &os, const MatrixELL &m);
  MatrixELL() : values(0), columns(0), max_col(0) {}
  MatrixELL(const tbsla::cpp::MatrixCOO &m);
  // Extension, no reference counterpart: the same conversion with the
  // (row, column) sort on the current GPU when `gpu` is true.
  MatrixELL(const tbsla::cpp::MatrixCOO &m, bool gpu);
  const std::vector<double> &get_values() const { return values; }
  const std::vector<int> &get_columns() const { return columns; }
  int get_max_col() const { return max_col; }
  std::vector<double> spmv(const std::vector<double> &v, int vect_incr = 0) const override;
  void Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr = 0) const override;
  using tbsla::cpp::Matrix::a_axpx_;
  std::ostream &print_stats(std::ostream &os) override;
  std::ostream &print_infos(std::ostream &os) override;
  std::ostream &print_as_dense(std::ostream &os) override;
  std::ostream &write(std::ostream &os) override;
  std::istream &read(std::istream &is, std::size_t pos = 0, std::size_t n = 1) override;
  std::ostream &print(std::ostream &os) const override;
  void fill_cdiag(int n_row, int n_col, int cdiag, int pr = 0, int pc = 0, int NR = 1, int NC = 1) override;
  void fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1, int pr = 0, int pc = 0,
                  int NR = 1, int NC = 1) override;
  void fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1, int pr = 0,
                             int pc = 0, int NR = 1, int NC = 1) override;
  void fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR,
                         int NC, std::uint64_t value_seed) override;
  void normalize_columns() override;
  // Extension, no reference counterpart: fill_cqmat_values (stochastic:
  // fill_cqmat_stochastic) with the entries generated on the current GPU.
  // The host arrays and metadata are identical to the host generators'.
  void fill_cqmat_gpu(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr, int pc, int NR, int NC,
                      std::uint64_t value_seed, bool stochastic);
  Format format() const override { return FORMAT_ELL; }

 protected:
  void from_coo(const tbsla::cpp::MatrixCOO &m, bool gpu);
  std::vector<double> values;
  std::vector<int> columns;
  int max_col;
};

}  // namespace cpp
}  // namespace tbsla

#endif
