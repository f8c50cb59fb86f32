// C ABIs of libtbsla_hip.so: the reference's C_CPP_* interface
// (src/cpp/interface/c/matrix/*.cpp) and the tbsla_* / tbsla_hip_* interface
// declared in include/tbsla_hip.h.
#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>

#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <numeric>
#include <string>

#include "../../../include/tbsla_cpp_c.h"
#include "../../../include/tbsla_hip.h"
#include "../device/adjoint.hpp"
#include "../device/batch.hpp"
#include "../device/kernels.hpp"
#include "../device/mirror.hpp"
#include "../device/pagerank.hpp"
#include "../device/products.hpp"
#include "../host/fill_common.hpp"

using namespace tbsla;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

template <class F>
int guarded(F &&f) {
  try {
    f();
    return TBSLA_OK;
  } catch (const hip::HipError &e) {
    return fail(TBSLA_EHIP, e.what());
  } catch (const std::invalid_argument &e) {
    return fail(TBSLA_EINVAL, e.what());
  } catch (const std::exception &e) {
    return fail(TBSLA_EOTHER, e.what());
  }
}

cpp::Matrix *make_matrix(int format) {
  switch (format) {
    case TBSLA_COO: return new cpp::MatrixCOO();
    case TBSLA_SCOO: return new cpp::MatrixSCOO();
    case TBSLA_CSR: return new cpp::MatrixCSR();
    case TBSLA_ELL: return new cpp::MatrixELL();
    case TBSLA_DENSE: return new cpp::MatrixDENSE();
  }
  return nullptr;
}

hip::Dtype dtype_of(int d) {
  if (d < 0 || d > 3) throw std::invalid_argument("tbsla: unknown dtype " + std::to_string(d));
  return static_cast<hip::Dtype>(d);
}

}  // namespace

struct tbsla_matrix {
  std::unique_ptr<cpp::Matrix> m;
};

struct tbsla_batch {
  std::unique_ptr<tbsla::hip::PageRankBatch> b;
};

// ===================================================================== C_CPP

namespace {
std::vector<double> *vec(C_CPP_Vector_t *v) { return static_cast<std::vector<double> *>(v->obj); }
}  // namespace

extern "C" {

// Vector.cpp:13-22
C_CPP_Vector_t *C_CPP_Vector_create(void) {
  C_CPP_Vector_t *v = static_cast<C_CPP_Vector_t *>(std::malloc(sizeof(C_CPP_Vector_t)));
  v->obj = new std::vector<double>();
  return v;
}
// Vector.cpp:24-29
void C_CPP_Vector_destroy(C_CPP_Vector_t *v) {
  if (v == NULL) return;
  delete vec(v);
  std::free(v);
}
// Vector.cpp:31-40
void C_CPP_Vector_print(C_CPP_Vector_t *v) {
  if (v == NULL) {
    std::printf("C_CPP_Vector_print : input vector is NULL\n");
    return;
  }
  host::streamvector<double>(std::cout, "v", *vec(v));
  std::cout << std::endl << std::flush;
}
// Vector.cpp:42-50: takes the object, the previous one is not freed
void C_CPP_Vector_set(C_CPP_Vector_t *v, void *obj) {
  if (v == NULL) return;
  v->obj = obj;
}
// Vector.cpp:52-61
void C_CPP_Vector_copy(C_CPP_Vector_t *v, void *obj) {
  if (v == NULL) return;
  const auto *src = static_cast<std::vector<double> *>(obj);
  vec(v)->assign(src->begin(), src->end());
}
// Vector.cpp:63-71
void C_CPP_Vector_fill(C_CPP_Vector_t *v, int n, int s) {
  if (v == NULL) return;
  auto *o = vec(v);
  o->clear();
  o->resize(static_cast<std::size_t>(n));
  std::iota(o->begin(), o->end(), s);
}
// Vector.cpp:73-88
bool C_CPP_Vector_read(C_CPP_Vector_t *v, char *filename, int seek) {
  if (v == NULL) return false;
  std::ifstream is(std::string(filename), std::ifstream::binary);
  if (!is.is_open()) return false;
  is.seekg(seek, is.beg);
  return guarded([&] { host::read_array(is, *vec(v)); }) == TBSLA_OK;
}
// Vector.cpp:90-103 (appends)
bool C_CPP_Vector_write(C_CPP_Vector_t *v, char *filename) {
  if (v == NULL) return false;
  std::ofstream os(std::string(filename), std::ofstream::binary | std::ofstream::app);
  if (!os.is_open()) return false;
  host::write_array(os, *vec(v));
  return static_cast<bool>(os);
}
// Vector.cpp:105-114
bool C_CPP_Vector_add(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2) {
  if (v1 == NULL || v2 == NULL) return false;
  auto *a = vec(v1), *b = vec(v2);
  if (b->size() < a->size()) return false;
  std::transform(a->begin(), a->end(), b->begin(), a->begin(), std::plus<double>());
  return true;
}
// Vector.cpp:116-125
bool C_CPP_Vector_add_incr(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2, int incr) {
  if (v1 == NULL || v2 == NULL) return false;
  auto *a = vec(v1), *b = vec(v2);
  if (incr < 0 || b->size() < a->size() + static_cast<std::size_t>(incr)) return false;
  std::transform(a->begin(), a->end(), b->begin() + incr, a->begin(), std::plus<double>());
  return true;
}
// Vector.cpp:127-136
bool C_CPP_Vector_gather(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2) {
  if (v1 == NULL || v2 == NULL) return false;
  auto *a = vec(v1), *b = vec(v2);
  a->insert(a->end(), b->begin(), b->end());
  return true;
}

}  // extern "C"

// interface/c/matrix/MatrixCSR.cpp:4-93, the same for every format
#define TBSLA_C_MATRIX_DEF(F)                                                                                  \
  struct C_CPP_Matrix##F {                                                                                     \
    void *obj;                                                                                                 \
  };                                                                                                           \
  extern "C" C_CPP_Matrix##F##_t *C_CPP_Matrix##F##_create(void) {                                             \
    C_CPP_Matrix##F##_t *m = static_cast<C_CPP_Matrix##F##_t *>(std::malloc(sizeof(C_CPP_Matrix##F##_t)));     \
    m->obj = new cpp::Matrix##F();                                                                             \
    return m;                                                                                                  \
  }                                                                                                            \
  extern "C" void C_CPP_Matrix##F##_destroy(C_CPP_Matrix##F##_t *m) {                                          \
    if (m == NULL) return;                                                                                     \
    delete static_cast<cpp::Matrix##F *>(m->obj);                                                             \
    std::free(m);                                                                                              \
  }                                                                                                            \
  extern "C" void C_CPP_Matrix##F##_fill_cdiag(C_CPP_Matrix##F##_t *m, int n_row, int n_col, int cdiag, int pr,  \
                                               int pc, int NR, int NC) {                                       \
    if (m == NULL) return;                                                                                     \
    guarded([&] { static_cast<cpp::Matrix##F *>(m->obj)->fill_cdiag(n_row, n_col, cdiag, pr, pc, NR, NC); });  \
  }                                                                                                            \
  extern "C" void C_CPP_Matrix##F##_fill_cqmat(C_CPP_Matrix##F##_t *m, int n_row, int n_col, int c, double q,    \
                                               unsigned int seed_mult, int pr, int pc, int NR, int NC) {       \
    if (m == NULL) return;                                                                                     \
    guarded([&] { static_cast<cpp::Matrix##F *>(m->obj)->fill_cqmat(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC); }); \
  }                                                                                                            \
  extern "C" bool C_CPP_Matrix##F##_read(C_CPP_Matrix##F##_t *m, char *filename, int seek) {                   \
    if (m == NULL) return false;                                                                               \
    std::ifstream is(std::string(filename), std::ifstream::binary);                                            \
    if (!is.is_open()) return false;                                                                           \
    is.seekg(seek, is.beg);                                                                                    \
    return guarded([&] { static_cast<cpp::Matrix##F *>(m->obj)->read(is); }) == TBSLA_OK;                      \
  }                                                                                                            \
  extern "C" bool C_CPP_Matrix##F##_write(C_CPP_Matrix##F##_t *m, char *filename) {                            \
    if (m == NULL) return false;                                                                               \
    std::ofstream os(std::string(filename), std::ofstream::binary | std::ofstream::app);                       \
    if (!os.is_open()) return false;                                                                           \
    return guarded([&] {                                                                                       \
             static_cast<cpp::Matrix##F *>(m->obj)->write(os);                                                 \
             if (!os) throw std::runtime_error(std::string("tbsla: write to ") + filename + " failed");        \
           }) == TBSLA_OK;                                                                                     \
  }                                                                                                            \
  extern "C" void C_CPP_Matrix##F##_print(C_CPP_Matrix##F##_t *m) {                                            \
    if (m == NULL) return;                                                                                     \
    auto *o = static_cast<cpp::Matrix##F *>(m->obj);                                                           \
    o->print_infos(std::cerr);                                                                                 \
    std::cerr << *o << std::endl;                                                                              \
  }                                                                                                            \
  extern "C" C_CPP_Vector_t *C_CPP_Matrix##F##_spmv(C_CPP_Matrix##F##_t *m, C_CPP_Vector_t *v) {               \
    C_CPP_Vector_t *r = C_CPP_Vector_create();                                                                 \
    g_last_error.clear();                                                                                      \
    if (m == NULL || v == NULL) return r;                                                                      \
    guarded([&] {                                                                                              \
      std::vector<double> y = static_cast<cpp::Matrix##F *>(m->obj)->spmv(*vec(v), 0);                         \
      C_CPP_Vector_copy(r, &y);                                                                                \
    });                                                                                                        \
    return r;                                                                                                  \
  }

TBSLA_C_MATRIX_DEF(COO)
TBSLA_C_MATRIX_DEF(SCOO)
TBSLA_C_MATRIX_DEF(CSR)
TBSLA_C_MATRIX_DEF(ELL)
TBSLA_C_MATRIX_DEF(DENSE)

// ================================================================ tbsla_hip

namespace {
struct Arrays {
  const std::vector<double> *v = nullptr;
  const std::vector<int> *i0 = nullptr, *i1 = nullptr;
  int max_col = 0;
};
Arrays arrays_of(const cpp::Matrix &m) {
  Arrays a;
  switch (m.format()) {
    case cpp::FORMAT_COO: {
      const auto &x = dynamic_cast<const cpp::MatrixCOO &>(m);
      a.v = &x.get_values(); a.i0 = &x.get_row(); a.i1 = &x.get_col();
      break;
    }
    case cpp::FORMAT_SCOO: {
      const auto &x = dynamic_cast<const cpp::MatrixSCOO &>(m);
      a.v = &x.get_values(); a.i0 = &x.get_row(); a.i1 = &x.get_col();
      break;
    }
    case cpp::FORMAT_CSR: {
      const auto &x = dynamic_cast<const cpp::MatrixCSR &>(m);
      a.v = &x.get_values(); a.i0 = &x.get_rowptr(); a.i1 = &x.get_colidx();
      break;
    }
    case cpp::FORMAT_ELL: {
      const auto &x = dynamic_cast<const cpp::MatrixELL &>(m);
      a.v = &x.get_values(); a.i1 = &x.get_columns(); a.max_col = x.get_max_col();
      break;
    }
    case cpp::FORMAT_DENSE: {
      const auto &x = dynamic_cast<const cpp::MatrixDENSE &>(m);
      a.v = &x.get_values();
      break;
    }
  }
  return a;
}
}  // namespace


extern "C" {

int tbsla_abi_version(void) { return 2; }
const char *tbsla_last_error(void) { return g_last_error.c_str(); }

tbsla_matrix *tbsla_matrix_create(int format) {
  cpp::Matrix *m = make_matrix(format);
  if (!m) {
    fail(TBSLA_EINVAL, "tbsla: unknown format " + std::to_string(format));
    return nullptr;
  }
  return new tbsla_matrix{std::unique_ptr<cpp::Matrix>(m)};
}

void tbsla_matrix_destroy(tbsla_matrix *m) { delete m; }

int tbsla_matrix_fill_cdiag(tbsla_matrix *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  if (NR <= 0 || NC <= 0 || pr < 0 || pc < 0 || pr >= NR || pc >= NC || n_row < 0 || n_col < 0)
    return fail(TBSLA_EINVAL, "tbsla: bad block layout");
  return guarded([&] { m->m->fill_cdiag(n_row, n_col, cdiag, pr, pc, NR, NC); });
}

int tbsla_matrix_fill_cqmat(tbsla_matrix *m, int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr,
                            int pc, int NR, int NC, uint64_t value_seed) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  if (NR <= 0 || NC <= 0 || pr < 0 || pc < 0 || pr >= NR || pc >= NC || n_row < 0 || n_col < 0 || c < 0)
    return fail(TBSLA_EINVAL, "tbsla: bad block layout");
  return guarded([&] { m->m->fill_cqmat_values(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed); });
}

int tbsla_matrix_fill_cqmat_gpu(tbsla_matrix *m, int n_row, int n_col, int c, double q, unsigned int seed_mult,
                                int pr, int pc, int NR, int NC, uint64_t value_seed, int stochastic) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  if (NR <= 0 || NC <= 0 || pr < 0 || pc < 0 || pr >= NR || pc >= NC || n_row < 0 || n_col < 0 || c < 0 ||
      (stochastic && c <= 0))
    return fail(TBSLA_EINVAL, "tbsla: bad block layout");
  cpp::Matrix *x = m->m.get();
  const bool st = stochastic != 0;
  switch (x->format()) {
    case cpp::FORMAT_CSR:
      return guarded([&] {
        dynamic_cast<cpp::MatrixCSR *>(x)->fill_cqmat_gpu(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed, st);
      });
    case cpp::FORMAT_ELL:
      return guarded([&] {
        dynamic_cast<cpp::MatrixELL *>(x)->fill_cqmat_gpu(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed, st);
      });
    case cpp::FORMAT_SCOO:
      return guarded([&] {
        dynamic_cast<cpp::MatrixSCOO *>(x)->fill_cqmat_gpu(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed,
                                                           st);
      });
    case cpp::FORMAT_COO:
      return guarded([&] {
        dynamic_cast<cpp::MatrixCOO *>(x)->fill_cqmat_gpu(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, value_seed, st);
      });
    default:
      return fail(TBSLA_EINVAL, "tbsla: GPU generation builds COO, SCOO, CSR or ELL matrices");
  }
}

int tbsla_matrix_fill_cqmat_stochastic(tbsla_matrix *m, int n_row, int n_col, int c, double q,
                                       unsigned int seed_mult, int pr, int pc, int NR, int NC) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  if (NR <= 0 || NC <= 0 || pr < 0 || pc < 0 || pr >= NR || pc >= NC || n_row < 0 || n_col < 0 || c <= 0)
    return fail(TBSLA_EINVAL, "tbsla: bad block layout");
  return guarded([&] { m->m->fill_cqmat_stochastic(n_row, n_col, c, q, seed_mult, pr, pc, NR, NC); });
}

int tbsla_matrix_normalize_columns(tbsla_matrix *m) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  return guarded([&] { m->m->normalize_columns(); });
}

tbsla_matrix *tbsla_matrix_convert(const tbsla_matrix *coo, int format) {
  if (!coo || coo->m->format() != cpp::FORMAT_COO) {
    fail(TBSLA_EINVAL, "tbsla: convert needs a COO matrix");
    return nullptr;
  }
  const auto &c = dynamic_cast<const cpp::MatrixCOO &>(*coo->m);
  cpp::Matrix *out = nullptr;
  int rc = guarded([&] {
    switch (format) {
      case TBSLA_COO: out = new cpp::MatrixCOO(c); break;
      case TBSLA_SCOO: out = new cpp::MatrixSCOO(c); break;
      case TBSLA_CSR: out = new cpp::MatrixCSR(c); break;
      case TBSLA_ELL: out = new cpp::MatrixELL(c); break;
      case TBSLA_DENSE: out = new cpp::MatrixDENSE(c); break;
      default: throw std::invalid_argument("tbsla: unknown format");
    }
  });
  if (rc) return nullptr;
  return new tbsla_matrix{std::unique_ptr<cpp::Matrix>(out)};
}

tbsla_matrix *tbsla_matrix_convert_gpu(const tbsla_matrix *coo, int format) {
  if (!coo || coo->m->format() != cpp::FORMAT_COO) {
    fail(TBSLA_EINVAL, "tbsla: convert needs a COO matrix");
    return nullptr;
  }
  if (format != TBSLA_CSR && format != TBSLA_ELL) {
    fail(TBSLA_EINVAL, "tbsla: GPU conversion produces CSR or ELL");
    return nullptr;
  }
  const auto &c = dynamic_cast<const cpp::MatrixCOO &>(*coo->m);
  cpp::Matrix *out = nullptr;
  int rc = guarded([&] {
    if (format == TBSLA_CSR)
      out = new cpp::MatrixCSR(c, true);
    else
      out = new cpp::MatrixELL(c, true);
  });
  if (rc) return nullptr;
  return new tbsla_matrix{std::unique_ptr<cpp::Matrix>(out)};
}

tbsla_matrix *tbsla_matrix_from_arrays(int format, int n_row, int n_col, const double *values, size_t n_values,
                                       const int *a, size_t na, const int *b, size_t nb) {
  std::vector<double> v(values, values + n_values);
  std::vector<int> va(a, a + na), vb(b, b + nb);
  cpp::Matrix *out = nullptr;
  int rc = guarded([&] {
    switch (format) {
      case TBSLA_COO: out = new cpp::MatrixCOO(n_row, n_col, v, va, vb); break;
      case TBSLA_SCOO: out = new cpp::MatrixSCOO(n_row, n_col, v, va, vb); break;
      case TBSLA_CSR: out = new cpp::MatrixCSR(n_row, n_col, v, va, vb); break;
      default: throw std::invalid_argument("tbsla: array constructors exist for COO, SCOO and CSR");
    }
  });
  if (rc) return nullptr;
  return new tbsla_matrix{std::unique_ptr<cpp::Matrix>(out)};
}


int tbsla_matrix_info(const tbsla_matrix *m, long *out) {
  if (!m || !out) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  const cpp::Matrix &x = *m->m;
  const Arrays a = arrays_of(x);
  const long v[16] = {x.format(), x.get_n_row(), x.get_n_col(), x.get_f_row(), x.get_f_col(), x.get_ln_row(),
                      x.get_ln_col(), x.get_pr(), x.get_pc(), x.get_NR(), x.get_NC(), x.get_nnz(), a.max_col,
                      static_cast<long>(a.v ? a.v->size() : 0), static_cast<long>(a.i0 ? a.i0->size() : 0),
                      static_cast<long>(a.i1 ? a.i1->size() : 0)};
  std::memcpy(out, v, sizeof(v));
  return TBSLA_OK;
}

int tbsla_matrix_array(const tbsla_matrix *m, int which, const void **data, size_t *n) {
  if (!m || !data || !n) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  const Arrays a = arrays_of(*m->m);
  *data = nullptr;
  *n = 0;
  if (which == 0 && a.v) {
    *data = a.v->data();
    *n = a.v->size();
  } else if (which == 1 && a.i0) {
    *data = a.i0->data();
    *n = a.i0->size();
  } else if (which == 2 && a.i1) {
    *data = a.i1->data();
    *n = a.i1->size();
  } else if (which < 0 || which > 2) {
    return fail(TBSLA_EINVAL, "tbsla: array index must be 0, 1 or 2");
  }
  return TBSLA_OK;
}

int tbsla_matrix_write(tbsla_matrix *m, const char *path) {
  if (!m || !path) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  std::ofstream os(path, std::ofstream::binary | std::ofstream::app);
  if (!os.is_open()) return fail(TBSLA_EIO, std::string("tbsla: cannot open ") + path);
  return guarded([&] { m->m->write(os); });
}

int tbsla_matrix_read(tbsla_matrix *m, const char *path, long seek) {
  if (!m || !path) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  std::ifstream is(path, std::ifstream::binary);
  if (!is.is_open()) return fail(TBSLA_EIO, std::string("tbsla: cannot open ") + path);
  is.seekg(seek, is.beg);
  return guarded([&] {
    m->m->read(is);
    if (!is) throw std::runtime_error("tbsla: truncated matrix file");
  });
}

int tbsla_matrix_read_mm(tbsla_matrix *m, const char *path) {
  if (!m || !path) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  auto *coo = dynamic_cast<cpp::MatrixCOO *>(m->m.get());
  if (!coo) return fail(TBSLA_EINVAL, "tbsla: readMM needs a COO matrix");
  if (!std::ifstream(path).is_open()) return fail(TBSLA_EIO, std::string("tbsla: cannot open ") + path);
  try {
    coo->readMM(path);
  } catch (const cpp::MatrixFormatReadException &e) {
    return fail(TBSLA_EINVAL, std::string("tbsla: ") + e.what());
  } catch (const std::exception &e) {
    return fail(TBSLA_EOTHER, e.what());
  }
  return TBSLA_OK;
}

int tbsla_matrix_count_stochastic(const tbsla_matrix *m) { return m && m->m->count_stochastic() ? 1 : 0; }

int tbsla_matrix_spmv(tbsla_matrix *m, const double *x, size_t nx, int vect_incr, double *y, size_t ny) {
  if (!m || (!x && nx) || (!y && ny)) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  if (ny < static_cast<size_t>(m->m->get_n_row())) return fail(TBSLA_EINVAL, "tbsla: y shorter than n_row");
  return guarded([&] {
    std::vector<double> v(x, x + nx);
    std::vector<double> r = m->m->spmv(v, vect_incr);
    std::memcpy(y, r.data(), r.size() * sizeof(double));
  });
}

int tbsla_matrix_Ax(tbsla_matrix *m, double *r, size_t nr, const double *x, size_t nx, int vect_incr) {
  if (!m || (!x && nx) || (!r && nr)) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    std::vector<double> v(x, x + nx), rr(r, r + nr);
    m->m->Ax(rr, v, vect_incr);
    std::memcpy(r, rr.data(), nr * sizeof(double));
  });
}

int tbsla_matrix_a_axpx(tbsla_matrix *m, const double *x, size_t nx, double *y, size_t ny) {
  if (!m || (!x && nx) || (!y && ny)) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  if (ny < static_cast<size_t>(m->m->get_n_row())) return fail(TBSLA_EINVAL, "tbsla: y shorter than n_row");
  return guarded([&] {
    std::vector<double> v(x, x + nx);
    std::vector<double> r = m->m->a_axpx_(v, 0);
    std::memcpy(y, r.data(), r.size() * sizeof(double));
  });
}

int tbsla_matrix_page_rank(tbsla_matrix *m, double epsilon, double beta, int max_iteration, double *out, size_t n,
                           int *iterations) {
  if (!m || !out) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  if (n < static_cast<size_t>(m->m->get_n_col())) return fail(TBSLA_EINVAL, "tbsla: out shorter than n_col");
  return guarded([&] {
    int it = 0;
    std::vector<double> r = m->m->page_rank(epsilon, beta, max_iteration, it);
    std::memcpy(out, r.data(), r.size() * sizeof(double));
    if (iterations) *iterations = it;
  });
}

int tbsla_matrix_personalized_page_rank(tbsla_matrix *m, double epsilon, double beta, int max_iteration,
                                        const int *nodes, int n_nodes, double *out, size_t n, int *iterations) {
  if (!m || !out || (!nodes && n_nodes > 0) || n_nodes <= 0) return fail(TBSLA_EINVAL, "tbsla: bad argument");
  if (n < static_cast<size_t>(m->m->get_n_col())) return fail(TBSLA_EINVAL, "tbsla: out shorter than n_col");
  return guarded([&] {
    int it = 0;
    std::vector<int> nd(nodes, nodes + n_nodes);
    std::vector<double> r = m->m->personalized_page_rank(epsilon, beta, max_iteration, nd, it);
    std::memcpy(out, r.data(), r.size() * sizeof(double));
    if (iterations) *iterations = it;
  });
}

int tbsla_hip_prepare(tbsla_matrix *m, int value_dtype) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL matrix");
  return guarded([&] { (void)hip::mirror(*m->m, dtype_of(value_dtype)); });
}

int tbsla_hip_mirror_info(tbsla_matrix *m, int value_dtype, long *out) {
  if (!m || !out) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    const hip::DeviceMatrix &d = hip::mirror(*m->m, dtype_of(value_dtype));
    const std::size_t xe = d.vdt == hip::Dtype::F64 ? 8 : 4;
    std::size_t bytes = d.rowptr.bytes() + d.colidx.bytes() + d.vals.bytes() + d.tiles.bytes() + d.ecols.bytes() +
                        d.evals.bytes() + d.segrow.bytes() + d.long_rows.bytes() +
                        d.dvals.bytes();
    const long v[8] = {d.rows, d.entries, d.x_need, d.y_lo, d.y_hi,
                       d.ntiles, static_cast<long>(bytes),
                       d.noop ? 0L : static_cast<long>(d.algorithmic_bytes(xe, xe))};
    std::memcpy(out, v, sizeof(v));
  });
}

int tbsla_hip_pb_layout_check(tbsla_matrix *m) {
  if (!m) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    if (m->m->format() != cpp::FORMAT_CSR) throw std::invalid_argument("tbsla: the propagation-blocked layout needs a CSR matrix");
    const auto &c = dynamic_cast<const cpp::MatrixCSR &>(*m->m);
    auto host = hip::pb_build_host(c);  // refuses a matrix that is not square in one block
    auto gpu = hip::pb_build_gpu(hip::mirror(*m->m, hip::Dtype::F64), c.get_n_col());
    const std::string diff = hip::pb_layout_compare(*gpu, *host);
    if (!diff.empty()) throw std::invalid_argument("tbsla: propagation-blocked layouts differ in " + diff);
  });
}

int tbsla_hip_Ax(tbsla_matrix *m, int value_dtype, const void *x, size_t nx, int x_dtype, void *y, size_t ny,
                 int y_dtype, int vect_incr, void *stream) {
  if (!m || (!x && nx) || (!y && ny)) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    hip::Ax_device(*m->m, dtype_of(value_dtype), x, nx, dtype_of(x_dtype), y, ny, dtype_of(y_dtype), vect_incr,
                   static_cast<hipStream_t>(stream));
  });
}

int tbsla_hip_fill(void *y, int y_dtype, size_t n, double value, void *stream) {
  if (!y && n) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    hip::launch_fill(y, dtype_of(y_dtype), static_cast<long>(n), value, static_cast<hipStream_t>(stream));
  });
}

int tbsla_hip_page_rank(tbsla_matrix *m, const tbsla_pagerank_params *p, double *out_dev, size_t n_out,
                        void *stream, tbsla_pagerank_report *report) {
  return tbsla_hip_page_rank_ex(m, p, nullptr, 0, nullptr, 0, out_dev, n_out, stream, report);
}

namespace {
// Stored CSR entries (rowptr[rows] - rowptr[0]), the length of per-entry arrays.
size_t csr_entries(const cpp::Matrix &m) {
  if (m.format() != cpp::FORMAT_CSR) throw std::invalid_argument("tbsla: per-entry arrays need a CSR matrix");
  const auto &rp = dynamic_cast<const cpp::MatrixCSR &>(m).get_rowptr();
  return rp.size() < 2 ? 0 : static_cast<size_t>(rp.back() - rp.front());
}
}  // namespace

int tbsla_hip_page_rank_adjoint(tbsla_matrix *m, const tbsla_adjoint_params *p, const double *values_dev,
                                size_t n_values, const double *teleport_dev, size_t n_teleport,
                                const double *pr_dev, const double *w_dev, size_t n, double *grad_values_dev,
                                double *grad_teleport_dev, void *stream, tbsla_adjoint_report *report) {
  if (!m || !p || !pr_dev || !w_dev) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    const size_t ncol = static_cast<size_t>(m->m->get_n_col());
    const size_t ne = csr_entries(*m->m);
    if (n < ncol) throw std::invalid_argument("tbsla: PageRank vector / weight shorter than n_col");
    if (values_dev && n_values < ne) throw std::invalid_argument("tbsla: values shorter than the stored entries");
    if (teleport_dev && n_teleport < ncol) throw std::invalid_argument("tbsla: teleport shorter than n_col");
    if (p->precision != TBSLA_F32 && p->precision != TBSLA_F64)
      throw std::invalid_argument("tbsla: adjoint precision must be F32 or F64");
    hip::AdjointParams q;
    q.beta = p->beta;
    q.epsilon = p->epsilon;
    q.max_iteration = p->max_iteration;
    q.precision = p->precision == TBSLA_F32 ? hip::Dtype::F32 : hip::Dtype::F64;
    q.check_every = p->check_every > 0 ? p->check_every : 8;
    q.values = values_dev;
    q.teleport = teleport_dev;
    hip::AdjointReport r = hip::page_rank_adjoint(*m->m, q, pr_dev, w_dev, grad_values_dev, grad_teleport_dev,
                                                  static_cast<hipStream_t>(stream));
    if (report) {
      report->iterations = r.iterations;
      report->last_change = r.last_change;
      report->s = r.s;
      report->seconds = r.seconds;
    }
  });
}

int tbsla_hip_page_rank_ex(tbsla_matrix *m, const tbsla_pagerank_params *p, const double *values_dev,
                           size_t n_values, const double *teleport_dev, size_t n_teleport, double *out_dev,
                           size_t n_out, void *stream, tbsla_pagerank_report *report) {
  if (!m || !p || !out_dev) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    if (n_out < static_cast<size_t>(m->m->get_n_col()))
      throw std::invalid_argument("tbsla: out has " + std::to_string(n_out) + " entries, PageRank writes " +
                                  std::to_string(m->m->get_n_col()));
    if (values_dev && n_values < csr_entries(*m->m))
      throw std::invalid_argument("tbsla: values shorter than the stored entries");
    if (teleport_dev && n_teleport < static_cast<size_t>(m->m->get_n_col()))
      throw std::invalid_argument("tbsla: teleport shorter than n_col");
    hip::PageRankParams q;
    q.values = values_dev;
    q.teleport = teleport_dev;
    q.epsilon = p->epsilon;
    q.beta = p->beta;
    q.max_iteration = p->max_iteration;
    q.personalized = p->nodes != nullptr && p->n_nodes > 0;
    q.nodes = p->nodes;
    q.n_nodes = p->n_nodes;
    if (p->precision < TBSLA_F64 || p->precision > TBSLA_PR_PATTERN_F32)
      throw std::invalid_argument("tbsla: bad precision");
    q.precision = static_cast<hip::PrPrecision>(p->precision);
    if (p->rule < 0 || p->rule > 1) throw std::invalid_argument("tbsla: bad stopping rule");
    q.rule = static_cast<hip::PrRule>(p->rule);
    q.check_every = p->check_every > 0 ? p->check_every : 8;
    q.anchor_tau = p->anchor_tau;
    if (p->gather < 0 || p->gather > 2) throw std::invalid_argument("tbsla: bad gather mode");
    q.gather = p->gather;
    hip::PageRankReport r = hip::page_rank_device(*m->m, q, out_dev, static_cast<hipStream_t>(stream));
    if (report) {
      report->iterations = r.iterations;
      report->anchors = r.anchors;
      report->last_error = r.last_error;
      report->bound = r.bound;
      report->seconds = r.seconds;
    }
  });
}

int tbsla_hip_page_rank_noalloc(tbsla_matrix *m, const tbsla_pagerank_params *p, double *out_dev, size_t n_out,
                                void *stream, tbsla_pagerank_report *report) {
  if (!m || !p || !out_dev) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    if (n_out < static_cast<size_t>(m->m->get_n_col()))
      throw std::invalid_argument("tbsla: out has " + std::to_string(n_out) + " entries, PageRank writes " +
                                  std::to_string(m->m->get_n_col()));
    if (p->rule < 0 || p->rule > 1) throw std::invalid_argument("tbsla: bad stopping rule");
    if (p->gather < 0 || p->gather > 2) throw std::invalid_argument("tbsla: bad gather mode");
    hip::PageRankParams q;
    q.epsilon = p->epsilon;
    q.beta = p->beta;
    q.max_iteration = p->max_iteration;
    q.personalized = p->nodes != nullptr && p->n_nodes > 0;
    q.nodes = p->nodes;
    q.n_nodes = p->n_nodes;
    if (p->precision != TBSLA_F64) throw std::invalid_argument("tbsla: the allocation-free PageRank runs in F64");
    q.precision = hip::PrPrecision::F64;
    q.rule = static_cast<hip::PrRule>(p->rule);
    q.check_every = p->check_every > 0 ? p->check_every : 8;
    q.gather = p->gather;
    hip::PageRankReport r = hip::page_rank_device_noalloc(*m->m, q, out_dev, static_cast<hipStream_t>(stream));
    if (report) {
      report->iterations = r.iterations;
      report->anchors = r.anchors;
      report->last_error = r.last_error;
      report->bound = r.bound;
      report->seconds = r.seconds;
    }
  });
}

tbsla_batch *tbsla_batch_create(int precision) {
  tbsla_batch *b = nullptr;
  const int rc = guarded([&] {
    if (precision != TBSLA_F64 && precision != TBSLA_F32)
      throw std::invalid_argument("tbsla: batched PageRank runs in F64 or F32");
    b = new tbsla_batch{std::unique_ptr<hip::PageRankBatch>(new hip::PageRankBatch(dtype_of(precision)))};
  });
  return rc == TBSLA_OK ? b : nullptr;
}

void tbsla_batch_destroy(tbsla_batch *b) { delete b; }

int tbsla_batch_add_graph(tbsla_batch *b, const tbsla_matrix *m, const int *rhs_ptr, const int *nodes, int n_rhs) {
  if (!b || !m || !rhs_ptr || n_rhs <= 0) return fail(TBSLA_EINVAL, "tbsla: bad argument");
  return guarded([&] {
    std::vector<std::vector<int>> sets(static_cast<std::size_t>(n_rhs));
    for (int k = 0; k < n_rhs; k++) {
      if (rhs_ptr[k + 1] < rhs_ptr[k] || rhs_ptr[k] < 0) throw std::invalid_argument("tbsla: rhs_ptr is not monotone");
      if (rhs_ptr[k + 1] > rhs_ptr[k] && !nodes) throw std::invalid_argument("tbsla: NULL nodes");
      sets[static_cast<std::size_t>(k)].assign(nodes + rhs_ptr[k], nodes + rhs_ptr[k + 1]);
    }
    b->b->add_graph(*m->m, sets);
  });
}

int tbsla_batch_prepare(tbsla_batch *b) {
  if (!b) return fail(TBSLA_EINVAL, "tbsla: NULL batch");
  return guarded([&] { b->b->prepare(); });
}

int tbsla_batch_n_rhs(const tbsla_batch *b) { return b ? b->b->n_rhs() : 0; }

int tbsla_batch_layout(const tbsla_batch *b, long *offsets, size_t n) {
  if (!b || !offsets) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    const std::vector<long> o = b->b->rhs_offsets();
    if (n < o.size()) throw std::invalid_argument("tbsla: offsets shorter than n_rhs + 1");
    std::memcpy(offsets, o.data(), o.size() * sizeof(long));
  });
}

int tbsla_batch_run(tbsla_batch *b, const tbsla_batch_params *p, double *out_dev, size_t n_out, int *rhs_iterations,
                    size_t n_rhs_it, void *stream, tbsla_batch_report *report) {
  if (!b || !p || !out_dev) return fail(TBSLA_EINVAL, "tbsla: NULL argument");
  return guarded([&] {
    const std::vector<long> o = b->b->rhs_offsets();
    if (n_out < static_cast<size_t>(o.back())) throw std::invalid_argument("tbsla: output shorter than the batch");
    if (rhs_iterations && n_rhs_it < o.size() - 1) throw std::invalid_argument("tbsla: rhs_iterations too short");
    if (p->rule < 0 || p->rule > 1) throw std::invalid_argument("tbsla: bad stopping rule");
    hip::BatchParams q;
    q.epsilon = p->epsilon;
    q.beta = p->beta;
    q.max_iteration = p->max_iteration;
    q.rule = static_cast<hip::PrRule>(p->rule);
    q.check_every = p->check_every > 0 ? p->check_every : 16;
    q.graph_capture = p->graph_capture != 0;
    q.autotune = p->autotune != 0;
    const hip::BatchReport r = b->b->run(q, out_dev, static_cast<hipStream_t>(stream));
    if (rhs_iterations)
      for (std::size_t k = 0; k < r.rhs_iterations.size(); k++) rhs_iterations[k] = r.rhs_iterations[k];
    if (report) {
      report->iterations = r.iterations;
      report->seconds = r.seconds;
      report->tune_seconds = r.tune_seconds;
      report->tune_iterations = r.tune_iterations;
      report->tune_cached = r.tune_cached ? 1 : 0;
      report->bytes_per_iteration = r.bytes_per_iteration;
      report->wg_per_cu = r.config.wg_per_cu;
      report->unroll = r.config.unroll;
      report->xcd_map = r.config.xcd_map;
      report->tile_nnz = r.config.tile_nnz;
      report->spmm_rows = r.config.spmm_rows;
    }
  });
}

}  // extern "C"
