// ref_shim.cpp -- flat extern "C" driver over the compiled TBSLA reference.
//
// TEST INFRASTRUCTURE ONLY.  Linked with the reference's own sources
// (src/cpp/matrix, src/cpp/utils, src/cpp/interface/c) into
// oracle/_ref/libtbsla_ref.so by oracle/Makefile.  It only calls the
// reference's public API (src/cpp/include/tbsla/cpp/*.hpp); matrix contents
// are read back through the reference's own binary write() (e.g.
// MatrixCSR.cpp:179-204) so no private member is touched.
#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>

#include <chrono>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

using namespace tbsla::cpp;

namespace {
Matrix *make(int fmt) {
  switch (fmt) {
    case 0: return new MatrixCOO();
    case 1: return new MatrixSCOO();
    case 2: return new MatrixCSR();
    case 3: return new MatrixELL();
    case 4: return new MatrixDENSE();
  }
  return nullptr;
}
struct Blob {
  std::string bytes;
};
}  // namespace

extern "C" {

void *ref_new(int fmt) { return make(fmt); }
// tbsla::cpp::Matrix has no virtual destructor and is a virtual base, so the
// object is deleted through its most-derived type.
void ref_free(void *p) {
  Matrix *m = static_cast<Matrix *>(p);
  if (auto *a = dynamic_cast<MatrixCOO *>(m)) delete a;
  else if (auto *b = dynamic_cast<MatrixSCOO *>(m)) delete b;
  else if (auto *c = dynamic_cast<MatrixCSR *>(m)) delete c;
  else if (auto *d = dynamic_cast<MatrixELL *>(m)) delete d;
  else if (auto *e = dynamic_cast<MatrixDENSE *>(m)) delete e;
}

void ref_fill_cdiag(void *m, int n_row, int n_col, int c, int pr, int pc, int NR, int NC) {
  static_cast<Matrix *>(m)->fill_cdiag(n_row, n_col, c, pr, pc, NR, NC);
}
void ref_fill_cqmat(void *m, int n_row, int n_col, int c, double q, unsigned s, int pr,
                    int pc, int NR, int NC) {
  static_cast<Matrix *>(m)->fill_cqmat(n_row, n_col, c, q, s, pr, pc, NR, NC);
}
void ref_fill_cqmat_stochastic(void *m, int n_row, int n_col, int c, double q, unsigned s,
                               int pr, int pc, int NR, int NC) {
  static_cast<Matrix *>(m)->fill_cqmat_stochastic(n_row, n_col, c, q, s, pr, pc, NR, NC);
}
void ref_normalize_columns(void *m) { static_cast<Matrix *>(m)->normalize_columns(); }

// X(const MatrixCOO&) conversions; `coo` must come from ref_new(0).
void *ref_convert(void *coo, int fmt) {
  const MatrixCOO &c = *dynamic_cast<MatrixCOO *>(static_cast<Matrix *>(coo));
  switch (fmt) {
    case 1: return static_cast<Matrix *>(new MatrixSCOO(c));
    case 2: return static_cast<Matrix *>(new MatrixCSR(c));
    case 3: return static_cast<Matrix *>(new MatrixELL(c));
    case 4: return static_cast<Matrix *>(new MatrixDENSE(c));
  }
  return nullptr;
}

// Array constructors (MatrixCSR.cpp:13, MatrixCOO.cpp:30, MatrixSCOO.cpp:12).
void *ref_from_arrays(int fmt, int n_row, int n_col, const double *v, size_t nv,
                      const int *a, size_t na, const int *b, size_t nb) {
  std::vector<double> vals(v, v + nv);
  std::vector<int> va(a, a + na), vb(b, b + nb);
  switch (fmt) {
    case 0: return static_cast<Matrix *>(new MatrixCOO(n_row, n_col, vals, va, vb));
    case 1: return static_cast<Matrix *>(new MatrixSCOO(n_row, n_col, vals, va, vb));
    case 2: return static_cast<Matrix *>(new MatrixCSR(n_row, n_col, vals, va, vb));
  }
  return nullptr;
}

// Serialise with the reference's write(); the caller frees with ref_blob_free.
void *ref_serialize(void *m, const char **data, size_t *size) {
  std::ostringstream os(std::ios::binary);
  static_cast<Matrix *>(m)->write(os);
  Blob *b = new Blob{os.str()};
  *data = b->bytes.data();
  *size = b->bytes.size();
  return b;
}
void ref_blob_free(void *b) { delete static_cast<Blob *>(b); }

// Read one matrix with the reference's read() (MatrixCSR.cpp:206-232 and
// siblings) from `size` bytes at `data`; returns the new matrix.
void *ref_deserialize(int fmt, const char *data, size_t size) {
  Matrix *m = make(fmt);
  std::istringstream is(std::string(data, size), std::ios::binary);
  m->read(is);
  return m;
}

long ref_get_nnz(void *m) { return static_cast<Matrix *>(m)->get_nnz(); }
int ref_get_n_row(void *m) { return static_cast<Matrix *>(m)->get_n_row(); }
int ref_get_n_col(void *m) { return static_cast<Matrix *>(m)->get_n_col(); }

// spmv(v, vect_incr): out has n_row entries.
void ref_spmv(void *m, const double *x, size_t nx, int vect_incr, double *out) {
  std::vector<double> v(x, x + nx);
  std::vector<double> r = static_cast<Matrix *>(m)->spmv(v, vect_incr);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
}
// Ax(r, v): r is read and written in place (nr entries).
void ref_Ax(void *m, double *r, size_t nr, const double *x, size_t nx, int vect_incr) {
  std::vector<double> v(x, x + nx), rr(r, r + nr);
  static_cast<Matrix *>(m)->Ax(rr, v, vect_incr);
  std::memcpy(r, rr.data(), nr * sizeof(double));
}
// Repeated Ax on preallocated vectors; returns nothing, used for CPU timing.
void ref_Ax_repeat(void *m, double *r, size_t nr, const double *x, size_t nx, int reps) {
  std::vector<double> v(x, x + nx), rr(nr, 0.0);
  for (int k = 0; k < reps; k++) static_cast<Matrix *>(m)->Ax(rr, v, 0);
  std::memcpy(r, rr.data(), nr * sizeof(double));
}
// Timings for the CPU baseline: `reps` calls of Ax on preallocated vectors
// (the kernel) and of spmv() (the API: a new zeroed n_row result per call),
// each timed alone; r receives the last Ax result.
void ref_Ax_times(void *m, double *r, size_t nr, const double *x, size_t nx, int reps, double *seconds) {
  std::vector<double> v(x, x + nx), rr(nr, 0.0);
  for (int k = 0; k < reps; k++) {
    const auto t0 = std::chrono::steady_clock::now();
    static_cast<Matrix *>(m)->Ax(rr, v, 0);
    seconds[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  std::memcpy(r, rr.data(), nr * sizeof(double));
}
void ref_spmv_times(void *m, const double *x, size_t nx, int reps, double *seconds) {
  std::vector<double> v(x, x + nx);
  for (int k = 0; k < reps; k++) {
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<double> r = static_cast<Matrix *>(m)->spmv(v, 0);
    seconds[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (r.empty()) seconds[k] = -1;
  }
}
void ref_a_axpx(void *m, const double *x, size_t nx, double *out) {
  std::vector<double> v(x, x + nx);
  std::vector<double> r = static_cast<Matrix *>(m)->a_axpx_(v, 0);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
}
int ref_page_rank(void *m, double epsilon, double beta, int max_it, double *out) {
  int it = 0;
  std::vector<double> r = static_cast<Matrix *>(m)->page_rank(epsilon, beta, max_it, it);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
  return it;
}
int ref_personalized_page_rank(void *m, double epsilon, double beta, int max_it,
                               const int *nodes, int n_nodes, double *out) {
  int it = 0;
  std::vector<int> nd(nodes, nodes + n_nodes);
  std::vector<double> r =
      static_cast<Matrix *>(m)->personalized_page_rank(epsilon, beta, max_it, nd, it);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
  return it;
}

// MatrixCOO::readMM (MatrixCOO.cpp:230-292); returns 0 when the reference
// throws MatrixFormatReadException
int ref_read_mm(void *m, const char *path) {
  try {
    dynamic_cast<MatrixCOO *>(static_cast<Matrix *>(m))->readMM(std::string(path));
  } catch (const MatrixFormatReadException &) {
    return 0;
  }
  return 1;
}

}  // extern "C"
