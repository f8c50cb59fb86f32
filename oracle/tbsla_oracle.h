/*
 * tbsla_oracle.h -- CPU restatement of the TBSLA SpMV / PageRank path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle: only the test
 * suite, the smoke check and the CPU-baseline leg of bench.py may load it.
 * The product (libtbsla_hip.so) never links or calls it.
 *
 * Every function restates one reference function and cites it
 * (paths relative to the TBSLA source tree, src/cpp/...).  Loop nests and
 * summation orders follow the reference so that results are bit-identical
 * to a single-threaded reference run when compiled with -ffp-contract=off.
 *
 * Pinning (what checks each part; tests/test_oracle.py):
 *   lnv/pflv           brute-force partition invariants + the compiled reference
 *   cdiag generators   closed form compute_spmv_res_cdiag + compiled reference arrays
 *   cqmat generators   compiled reference arrays (bit-exact) + golden SHA-256 digests
 *   conversions        compiled reference arrays; invariants (sorted rows, multiset)
 *   spmv (5 formats)   5x5 known answer {20,12,10,40,85}; cdiag closed form;
 *                      exact cross-format equality; compiled reference outputs
 *   page_rank          compiled reference outputs (bit-exact) + dense Perron
 *                      vector (numpy eig) at tiny n
 *   value overlay/x    project-defined generator (SplitMix64), pinned by
 *                      published SplitMix64 test values
 *
 * Parity unpinned: the order of duplicate (row, col) entries after a
 * COO->CSR/ELL conversion (the reference uses an unstable std::sort; this
 * restatement uses a stable merge sort).  Indices are unaffected; only the
 * value order among exact duplicates can differ.
 */
#ifndef TBSLA_ORACLE_H
#define TBSLA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_COO = 0, ORC_SCOO = 1, ORC_CSR = 2, ORC_ELL = 3, ORC_DENSE = 4 };

typedef struct orc_mat {
  int fmt;
  int n_row, n_col, f_row, f_col, ln_row, ln_col, pr, pc, NR, NC;
  long nnz;
  int max_col;            /* ELL only */
  /* values: all formats (DENSE: row-major ln_row x ln_col) */
  double *val; size_t n_val, cap_val;
  /* idx0: CSR rowptr, COO/SCOO row, unused for ELL/DENSE */
  int *idx0; size_t n_idx0, cap_idx0;
  /* idx1: CSR colidx, COO/SCOO col, ELL columns */
  int *idx1; size_t n_idx1, cap_idx1;
} orc_mat;

/* ---- utils/range.cpp ---- */
size_t orc_lnv(size_t size, int local_pos, int number_pos);
size_t orc_pflv(size_t size, int local_pos, int number_pos);

/* ---- utils/values_generation.cpp ---- */
void orc_cdiag_value(size_t i, size_t nr, size_t nc, size_t cdiag,
                     uint64_t *row, uint64_t *col, double *v);
void orc_cqmat_value(size_t i, size_t nr, size_t nc, size_t c, double q,
                     unsigned seed_mult, uint64_t *row, uint64_t *col, double *v);
void orc_cqmat_sum_columns(size_t nr, size_t nc, size_t c, double q,
                           unsigned seed_mult, double *sum /* nc */);

/* ---- project-defined generators (no reference counterpart) ---- */
uint64_t orc_splitmix64(uint64_t z);
double orc_u01(uint64_t seed, uint64_t idx);
void orc_uniform_vector(uint64_t seed, uint64_t first, size_t n, double *out);

/* ---- utils/vector.cpp closed forms ---- */
void orc_spmv_res_cdiag(int nr, int nc, int c, const double *v, double *r);

/* ---- matrix lifetime ---- */
orc_mat *orc_new(int fmt);
void orc_free(orc_mat *m);
orc_mat *orc_from_arrays(int fmt, int n_row, int n_col,
                         const double *val, size_t n_val,
                         const int *a, size_t na, const int *b, size_t nb);
orc_mat *orc_convert(const orc_mat *coo, int fmt);
void orc_info(const orc_mat *m, long out[16]);

/* ---- generators per format ----
 * value_seed == 0 : reference values (1.0)
 * value_seed != 0 : project overlay, value = 1 - u01(value_seed, i) where i is
 *                   the entry's global generation index (cqmat only). */
void orc_fill_cdiag(orc_mat *m, int n_row, int n_col, int cdiag,
                    int pr, int pc, int NR, int NC);
void orc_fill_cqmat(orc_mat *m, int n_row, int n_col, int c, double q,
                    unsigned seed_mult, int pr, int pc, int NR, int NC,
                    uint64_t value_seed);
void orc_fill_cqmat_stochastic(orc_mat *m, int n_row, int n_col, int c, double q,
                               unsigned seed_mult, int pr, int pc, int NR, int NC);
void orc_normalize_columns(orc_mat *m);

/* ---- products ---- */
void orc_Ax(const orc_mat *m, double *r, const double *v, int vect_incr);
void orc_spmv(const orc_mat *m, const double *v, int vect_incr, double *out);
void orc_a_axpx(const orc_mat *m, const double *v, int vect_incr, double *out);
int orc_page_rank(const orc_mat *m, double epsilon, double beta, int max_iteration,
                  double *out, double *last_error);
int orc_personalized_page_rank(const orc_mat *m, double epsilon, double beta,
                               int max_iteration, const int *nodes, int n_nodes,
                               double *out, double *last_error);

/* ---- binary file format (append on write, like the C ABI) ---- */
long orc_write(const orc_mat *m, const char *path);
long orc_read(orc_mat *m, const char *path, long seek);

#ifdef __cplusplus
}
#endif
#endif
