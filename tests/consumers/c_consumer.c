/*
 * A C consumer of the TBSLA C interface (include/tbsla_cpp_c.h), compiled
 * as C99 and linked to libtbsla_hip.so, making the calls the reference's
 * YML components make (yml/impl/gen_Sparse_CQMAT.query:44-59,
 * yml/impl/spmv.query:17-32) with files between the steps: generate and
 * write a matrix, read it back into a new handle, read x, multiply, write y.
 * Built by __graft_entry__.build(), run by tests/test_gpu_consumers.py.
 *
 *   c_consumer FORMAT n_row n_col C Q S pr pc GR GC matrix_file x_file y_file
 *
 * Exit status 0 on success, 1 on a usage error, 2 on a failed call.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tbsla_cpp_c.h"

#define RUN(F)                                                                            \
  do {                                                                                    \
    C_CPP_Matrix##F##_t *a = C_CPP_Matrix##F##_create();                                  \
    C_CPP_Matrix##F##_fill_cqmat(a, n_row, n_col, c, q, s, pr, pc, gr, gc);               \
    if (!C_CPP_Matrix##F##_write(a, mfile)) return 2;                                     \
    C_CPP_Matrix##F##_destroy(a);                                                         \
    C_CPP_Matrix##F##_t *b = C_CPP_Matrix##F##_create();                                  \
    if (!C_CPP_Matrix##F##_read(b, mfile, 0)) return 2;                                   \
    C_CPP_Vector_t *x = C_CPP_Vector_create();                                            \
    if (!C_CPP_Vector_read(x, xfile, 0)) return 2;                                        \
    C_CPP_Vector_t *y = C_CPP_Matrix##F##_spmv(b, x);                                     \
    if (!C_CPP_Vector_write(y, yfile)) return 2;                                          \
    C_CPP_Vector_destroy(y);                                                              \
    C_CPP_Vector_destroy(x);                                                              \
    C_CPP_Matrix##F##_destroy(b);                                                         \
    return 0;                                                                             \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 14) {
    fprintf(stderr, "usage: %s FORMAT n_row n_col C Q S pr pc GR GC matrix_file x_file y_file\n", argv[0]);
    return 1;
  }
  const char *fmt = argv[1];
  int n_row = atoi(argv[2]), n_col = atoi(argv[3]), c = atoi(argv[4]);
  double q = atof(argv[5]);
  unsigned int s = (unsigned int)strtoul(argv[6], NULL, 10);
  int pr = atoi(argv[7]), pc = atoi(argv[8]), gr = atoi(argv[9]), gc = atoi(argv[10]);
  char *mfile = argv[11], *xfile = argv[12], *yfile = argv[13];
  if (strcmp(fmt, "COO") == 0) RUN(COO);
  if (strcmp(fmt, "SCOO") == 0) RUN(SCOO);
  if (strcmp(fmt, "CSR") == 0) RUN(CSR);
  if (strcmp(fmt, "ELL") == 0) RUN(ELL);
  if (strcmp(fmt, "DENSE") == 0) RUN(DENSE);
  fprintf(stderr, "%s unrecognized!\n", fmt);
  return 1;
}
