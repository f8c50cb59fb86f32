/*
 * tbsla_cpp_c.h -- the TBSLA C interface, served by libtbsla_hip.so.
 *
 * Same symbols, C types, ownership and NULL behaviour as the reference C ABI
 * (src/cpp/interface/c/include/tbsla/cpp/{MatrixCOO,MatrixSCOO,MatrixCSR,
 * MatrixELL,MatrixDENSE,Vector}.h), so a consumer such as the YML spmv
 * component (src/yml/impl/spmv.query:17-39) links against this library
 * unchanged.  C_CPP_Matrix<F>_spmv runs the product on the GPU.
 *
 *   C_CPP_Matrix<F>_create / _destroy      interface/c/matrix/MatrixCSR.cpp:9-26
 *   C_CPP_Matrix<F>_fill_cdiag / _cqmat    interface/c/matrix/MatrixCSR.cpp:30-44
 *   C_CPP_Matrix<F>_read / _write          interface/c/matrix/MatrixCSR.cpp:46-70
 *   C_CPP_Matrix<F>_print                  interface/c/matrix/MatrixCSR.cpp:73-80
 *   C_CPP_Matrix<F>_spmv                   interface/c/matrix/MatrixCSR.cpp:82-93
 *   C_CPP_Vector_*                         interface/c/matrix/Vector.cpp:13-136
 *
 * Differences: an exception thrown by the library (for instance a vector
 * too short for the matrix, which the reference would read past) is caught
 * at this boundary, and its message is available from tbsla_last_error()
 * (tbsla_hip.h).  _spmv clears that message on entry and returns an empty
 * vector on failure; _fill_* leave the matrix as the failed call left it;
 * _read and _write return false, and a truncated or corrupt file leaves the
 * matrix unchanged (the reference reads past the end of such a file).
 */
#ifndef TBSLA_CPP_C_H
#define TBSLA_CPP_C_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

struct C_CPP_Vector {
  void *obj;
};
typedef struct C_CPP_Vector C_CPP_Vector_t;

C_CPP_Vector_t *C_CPP_Vector_create(void);
void C_CPP_Vector_destroy(C_CPP_Vector_t *v);
void C_CPP_Vector_set(C_CPP_Vector_t *v, void *obj);
void C_CPP_Vector_copy(C_CPP_Vector_t *v, void *obj);
void C_CPP_Vector_print(C_CPP_Vector_t *v);
void C_CPP_Vector_fill(C_CPP_Vector_t *v, int n, int s);
bool C_CPP_Vector_read(C_CPP_Vector_t *v, char *filename, int seek);
bool C_CPP_Vector_write(C_CPP_Vector_t *v, char *filename);
bool C_CPP_Vector_add(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2);
bool C_CPP_Vector_add_incr(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2, int incr);
bool C_CPP_Vector_gather(C_CPP_Vector_t *v1, C_CPP_Vector_t *v2);

#define TBSLA_C_MATRIX_DECL(F)                                                                              \
  struct C_CPP_Matrix##F;                                                                                  \
  typedef struct C_CPP_Matrix##F C_CPP_Matrix##F##_t;                                                      \
  C_CPP_Matrix##F##_t *C_CPP_Matrix##F##_create(void);                                                     \
  void C_CPP_Matrix##F##_destroy(C_CPP_Matrix##F##_t *m);                                                  \
  void C_CPP_Matrix##F##_fill_cdiag(C_CPP_Matrix##F##_t *m, int n_row, int n_col, int cdiag, int pr, int pc, \
                                    int NR, int NC);                                                       \
  void C_CPP_Matrix##F##_fill_cqmat(C_CPP_Matrix##F##_t *m, int n_row, int n_col, int c, double q,         \
                                    unsigned int seed_mult, int pr, int pc, int NR, int NC);               \
  bool C_CPP_Matrix##F##_read(C_CPP_Matrix##F##_t *m, char *filename, int seek);                           \
  bool C_CPP_Matrix##F##_write(C_CPP_Matrix##F##_t *m, char *filename);                                    \
  void C_CPP_Matrix##F##_print(C_CPP_Matrix##F##_t *m);                                                    \
  C_CPP_Vector_t *C_CPP_Matrix##F##_spmv(C_CPP_Matrix##F##_t *m, C_CPP_Vector_t *v);

TBSLA_C_MATRIX_DECL(COO)
TBSLA_C_MATRIX_DECL(SCOO)
TBSLA_C_MATRIX_DECL(CSR)
TBSLA_C_MATRIX_DECL(ELL)
TBSLA_C_MATRIX_DECL(DENSE)

#undef TBSLA_C_MATRIX_DECL

#ifdef __cplusplus
}
#endif

#endif /* TBSLA_CPP_C_H */
