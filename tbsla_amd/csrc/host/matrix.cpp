// tbsla::cpp::Matrix base: a_axpx_ (Matrix.cpp:6-11), page_rank
// (Matrix.cpp:13-61), personalized_page_rank (Matrix.cpp:63-114), all on the
// GPU; device-mirror lifetime.
#include <tbsla/cpp/Matrix.hpp>

#include <algorithm>
#include <stdexcept>

#include "../device/mirror.hpp"
#include "../device/pagerank.hpp"

namespace tbsla {
namespace cpp {

Matrix::Matrix() = default;
Matrix::Matrix(const Matrix &o)
    : n_row(o.n_row), n_col(o.n_col), f_row(o.f_row), f_col(o.f_col), ln_row(o.ln_row), ln_col(o.ln_col),
      pr(o.pr), pc(o.pc), NR(o.NR), NC(o.NC), nnz(o.nnz), count_stochastic_(o.count_stochastic_) {}
Matrix &Matrix::operator=(const Matrix &o) {
  if (this != &o) {
    n_row = o.n_row; n_col = o.n_col; f_row = o.f_row; f_col = o.f_col; ln_row = o.ln_row; ln_col = o.ln_col;
    pr = o.pr; pc = o.pc; NR = o.NR; NC = o.NC; nnz = o.nnz; count_stochastic_ = o.count_stochastic_;
    dev_.reset();
  }
  return *this;
}
Matrix::~Matrix() = default;

hip::DeviceCache &Matrix::device_cache() const {
  if (!dev_) dev_ = std::make_shared<hip::DeviceCache>();
  return *dev_;
}

void Matrix::invalidate_device() const { dev_.reset(); }

std::vector<double> Matrix::a_axpx_(const std::vector<double> &v, int vect_incr) const {
  std::vector<double> r = this->spmv(v, vect_incr);
  if (v.size() < r.size()) throw std::invalid_argument("tbsla: a_axpx_ needs a square matrix");
  std::transform(r.begin(), r.end(), v.begin(), r.begin(), std::plus<double>());
  r = this->spmv(r, vect_incr);
  return r;
}

std::vector<double> Matrix::page_rank(double epsilon, double beta, int max_iteration, int &nb_iterations_done) {
  hip::PageRankParams p;
  p.epsilon = epsilon;
  p.beta = beta;
  p.max_iteration = max_iteration;
  std::vector<double> out(static_cast<std::size_t>(n_col));
  hip::PageRankReport rep = hip::page_rank(*this, p, out.data());
  nb_iterations_done = rep.iterations;
  return out;
}

std::vector<double> Matrix::personalized_page_rank(double epsilon, double beta, int max_iteration,
                                                   std::vector<int> personalized_nodes, int &nb_iterations_done) {
  hip::PageRankParams p;
  p.epsilon = epsilon;
  p.beta = beta;
  p.max_iteration = max_iteration;
  p.nodes = personalized_nodes.data();
  p.n_nodes = static_cast<int>(personalized_nodes.size());
  p.personalized = true;
  std::vector<double> out(static_cast<std::size_t>(n_col));
  hip::PageRankReport rep = hip::page_rank(*this, p, out.data());
  nb_iterations_done = rep.iterations;
  return out;
}

}  // namespace cpp
}  // namespace tbsla
