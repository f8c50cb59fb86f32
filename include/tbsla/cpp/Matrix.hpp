// tbsla::cpp::Matrix -- MI355X build of the TBSLA matrix API.
//
// Same class, method names, argument order and semantics as the reference
// header (src/cpp/include/tbsla/cpp/Matrix.hpp:9-47).  Products (Ax, spmv,
// a_axpx_, page_rank, personalized_page_rank) run on the GPU through the
// HIP kernels of libtbsla_hip.so; generators, conversions and file I/O run
// on the host and are bit-identical to the reference.
//
// Additions to the reference API (marked below): a virtual destructor, host
// array accessors, format(), and the device mirror used by the kernels.
#ifndef TBSLA_CPP_Matrix
#define TBSLA_CPP_Matrix

#include <cstddef>
#include <cstdint>
#include <exception>
#include <fstream>
#include <iostream>
#include <memory>
#include <vector>

namespace tbsla {
namespace hip {
struct DeviceCache;  // device mirrors of one host matrix (one per value type)
}
namespace cpp {

enum Format : int { FORMAT_COO = 0, FORMAT_SCOO = 1, FORMAT_CSR = 2, FORMAT_ELL = 3, FORMAT_DENSE = 4 };

class Matrix {
 public:
  Matrix();
  Matrix(const Matrix &o);
  Matrix &operator=(const Matrix &o);
  virtual ~Matrix();  // addition: the reference has none

  friend std::ostream &operator<<(std::ostream &os, const Matrix &m) { return m.print(os); }

  // Returns a zeroed vector of length n_row; CSR/ELL assign rows, COO/SCOO/DENSE
  // accumulate (Matrix.hpp:12, MatrixCSR.cpp:123-127).
  virtual std::vector<double> spmv(const std::vector<double> &v, int vect_incr = 0) const = 0;
  virtual void Ax(std::vector<double> &r, const std::vector<double> &v, int vect_incr = 0) const = 0;
  std::vector<double> a_axpx_(const std::vector<double> &v, int vect_incr = 0) const;
  std::vector<double> page_rank(double epsilon, double beta, int max_iteration, int &nb_iterations_done);
  std::vector<double> personalized_page_rank(double epsilon, double beta, int max_iteration,
                                             std::vector<int> personalized_nodes,
                                             int &nb_iterations_done);
  int get_n_row() const { return n_row; }
  int get_n_col() const { return n_col; }
  int get_f_row() const { return f_row; }
  int get_f_col() const { return f_col; }
  int get_ln_row() const { return ln_row; }
  int get_ln_col() const { return ln_col; }
  int get_pr() const { return pr; }
  int get_pc() const { return pc; }
  int get_NR() const { return NR; }
  int get_NC() const { return NC; }
  long int get_nnz() const { return nnz; }

  virtual std::ostream &print_stats(std::ostream &os) = 0;
  virtual std::ostream &print_infos(std::ostream &os) = 0;
  virtual std::ostream &write(std::ostream &os) = 0;
  virtual std::istream &read(std::istream &is, std::size_t pos = 0, std::size_t n = 1) = 0;
  virtual std::ostream &print(std::ostream &os) const = 0;
  virtual std::ostream &print_as_dense(std::ostream &os) = 0;

  virtual void fill_cdiag(int n_row, int n_col, int cdiag, int pr = 0, int pc = 0, int NR = 1, int NC = 1) = 0;
  virtual void fill_cqmat(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1, int pr = 0,
                          int pc = 0, int NR = 1, int NC = 1) = 0;
  virtual void fill_cqmat_stochastic(int n_row, int n_col, int c, double q, unsigned int seed_mult = 1,
                                     int pr = 0, int pc = 0, int NR = 1, int NC = 1) = 0;
  virtual void normalize_columns() = 0;

  // ---- additions -------------------------------------------------------
  virtual Format format() const = 0;
  // fill_cqmat with the project value overlay: value = 1 - u01(value_seed, i),
  // i the entry's global generation index (value_seed == 0: reference 1.0).
  virtual void fill_cqmat_values(int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr,
                                 int pc, int NR, int NC, std::uint64_t value_seed) = 0;
  // Lazily built device mirrors; every mutator drops them.
  hip::DeviceCache &device_cache() const;
  void invalidate_device() const;
  // True when the values were produced by fill_cqmat_stochastic, i.e. every
  // value equals fl64(1 / global column count) of its column.
  bool count_stochastic() const { return count_stochastic_; }

 protected:
  int n_row = 0, n_col = 0, f_row = 0, f_col = 0, ln_row = 0, ln_col = 0, pr = 0, pc = 0, NR = 1, NC = 1;
  long int nnz = 0;
  bool count_stochastic_ = false;
  mutable std::shared_ptr<hip::DeviceCache> dev_;
};

struct MatrixFormatReadException : public std::exception {
  const char *what() const throw() { return "This Matrix format import is not implemented!"; }
};

}  // namespace cpp
}  // namespace tbsla

#endif
