/*
 * tbsla_hip.h -- C ABI of libtbsla_hip.so for host runtimes and bindings.
 *
 * Plain pointers, sizes and integer status codes (0 = success, < 0 = error;
 * tbsla_last_error() returns the message of the calling thread's last
 * failure).  A tbsla_matrix wraps one tbsla::cpp::Matrix<F> object (host
 * arrays, bit-identical to the reference generators) plus its lazily built
 * HBM mirror.  `stream` arguments are hipStream_t values (NULL = default).
 *
 * Reference interfaces each group replaces:
 *   create/destroy/fill_* /normalize   Matrix.hpp:37-41, MatrixCSR.cpp:234-435
 *   convert / from_arrays              MatrixCSR.hpp:14-15 (X(const MatrixCOO&), array ctors)
 *   read / write                       MatrixCSR.cpp:179-232 (binary format)
 *   spmv / Ax / a_axpx                 Matrix.hpp:12-14, MatrixCSR.cpp:123-140, Matrix.cpp:6-11
 *   page_rank / personalized           Matrix.hpp:16-17, Matrix.cpp:13-114
 *   tbsla_hip_*                        device-resident forms of the same products
 *                                      (no reference counterpart: the reference is CPU-only)
 */
#ifndef TBSLA_HIP_H
#define TBSLA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tbsla_format { TBSLA_COO = 0, TBSLA_SCOO = 1, TBSLA_CSR = 2, TBSLA_ELL = 3, TBSLA_DENSE = 4 };
enum tbsla_dtype { TBSLA_F64 = 0, TBSLA_F32 = 1, TBSLA_BF16 = 2, TBSLA_PATTERN = 3 };
enum tbsla_status { TBSLA_OK = 0, TBSLA_EINVAL = -1, TBSLA_EHIP = -2, TBSLA_EIO = -3, TBSLA_EOTHER = -4 };

typedef struct tbsla_matrix tbsla_matrix;

int tbsla_abi_version(void);
const char *tbsla_last_error(void);

/* ---- host matrix (reference semantics) ---- */
tbsla_matrix *tbsla_matrix_create(int format);
void tbsla_matrix_destroy(tbsla_matrix *m);
int tbsla_matrix_fill_cdiag(tbsla_matrix *m, int n_row, int n_col, int cdiag, int pr, int pc, int NR, int NC);
/* value_seed = 0: reference values (1.0); otherwise value = 1 - u01(value_seed, i),
 * i the entry's global generation index (project overlay, no reference counterpart) */
int tbsla_matrix_fill_cqmat(tbsla_matrix *m, int n_row, int n_col, int c, double q, unsigned int seed_mult, int pr,
                            int pc, int NR, int NC, uint64_t value_seed);
int tbsla_matrix_fill_cqmat_stochastic(tbsla_matrix *m, int n_row, int n_col, int c, double q,
                                       unsigned int seed_mult, int pr, int pc, int NR, int NC);
/* fill_cqmat (stochastic = 0, with the value overlay above) or
 * fill_cqmat_stochastic (stochastic = 1, value_seed ignored) of a COO, SCOO,
 * CSR or ELL matrix with the entries generated on the current GPU; the host
 * arrays are those of tbsla_matrix_fill_cqmat / _stochastic (no reference
 * counterpart: the reference generates on the host, values_generation.cpp:
 * 26-62).  TBSLA_EINVAL for DENSE. */
int tbsla_matrix_fill_cqmat_gpu(tbsla_matrix *m, int n_row, int n_col, int c, double q, unsigned int seed_mult,
                                int pr, int pc, int NR, int NC, uint64_t value_seed, int stochastic);
int tbsla_matrix_normalize_columns(tbsla_matrix *m);
tbsla_matrix *tbsla_matrix_convert(const tbsla_matrix *coo, int format);
/* tbsla_matrix_convert to CSR or ELL with the (row, column) sort on the
 * current GPU: the same arrays (no reference counterpart: the reference sorts
 * on the host, MatrixCSR.cpp:29-74, MatrixELL.cpp:12-54).  NULL and
 * TBSLA_EINVAL for another format. */
tbsla_matrix *tbsla_matrix_convert_gpu(const tbsla_matrix *coo, int format);
tbsla_matrix *tbsla_matrix_from_arrays(int format, int n_row, int n_col, const double *values, size_t n_values,
                                       const int *a, size_t na, const int *b, size_t nb);
/* out[16]: format n_row n_col f_row f_col ln_row ln_col pr pc NR NC nnz max_col
 *          n_values n_idx0 n_idx1 */
int tbsla_matrix_info(const tbsla_matrix *m, long *out16);
/* which: 0 values, 1 rowptr (CSR) / row (COO, SCOO), 2 colidx (CSR) / col (COO, SCOO) / columns (ELL) */
int tbsla_matrix_array(const tbsla_matrix *m, int which, const void **data, size_t *n);
int tbsla_matrix_write(tbsla_matrix *m, const char *path); /* appends, like C_CPP_Matrix<F>_write */
int tbsla_matrix_read(tbsla_matrix *m, const char *path, long seek);
/* MatrixMarket "coordinate real general|symmetric" into a COO handle
 * (MatrixCOO::readMM, MatrixCOO.cpp:230-292).  TBSLA_EIO when the file cannot
 * be opened (the reference returns silently), TBSLA_EINVAL for another format
 * or an unsupported header (the reference throws MatrixFormatReadException). */
int tbsla_matrix_read_mm(tbsla_matrix *m, const char *path);
int tbsla_matrix_count_stochastic(const tbsla_matrix *m);

/* ---- host-vector products on the GPU ---- */
int tbsla_matrix_spmv(tbsla_matrix *m, const double *x, size_t nx, int vect_incr, double *y, size_t ny);
int tbsla_matrix_Ax(tbsla_matrix *m, double *r, size_t nr, const double *x, size_t nx, int vect_incr);
int tbsla_matrix_a_axpx(tbsla_matrix *m, const double *x, size_t nx, double *y, size_t ny);
int tbsla_matrix_page_rank(tbsla_matrix *m, double epsilon, double beta, int max_iteration, double *out, size_t n,
                           int *iterations);
int tbsla_matrix_personalized_page_rank(tbsla_matrix *m, double epsilon, double beta, int max_iteration,
                                        const int *nodes, int n_nodes, double *out, size_t n, int *iterations);

/* ---- device-resident products ---- */
/* Build the HBM mirror with values of `value_dtype` (idempotent). */
int tbsla_hip_prepare(tbsla_matrix *m, int value_dtype);
/* out[8]: rows, entries, x_need, y_lo, y_hi, tiles, mirror bytes, algorithmic bytes
 *         (for x/y elements of the mirror's natural type) */
int tbsla_hip_mirror_info(tbsla_matrix *m, int value_dtype, long *out8);
/* Ax semantics on device vectors (y assigned for CSR/ELL, accumulated for
 * COO/SCOO/DENSE).  Supported: values F64 with x,y F64; values F32/BF16 with
 * x,y F32.  Sizes in elements; rejected when too short for the matrix. */
int tbsla_hip_Ax(tbsla_matrix *m, int value_dtype, const void *x, size_t nx, int x_dtype, void *y, size_t ny,
                 int y_dtype, int vect_incr, void *stream);
/* y[0:n) = value (a kernel, safe inside stream capture) */
int tbsla_hip_fill(void *y, int y_dtype, size_t n, double value, void *stream);

/* PageRank precision beside the tbsla_dtype values: the mixed FP32 scheme of
 * TBSLA_F32 (FP32 products on a delta to an FP64 anchor) without a value
 * array, for matrices from fill_cqmat_stochastic (value(i,j) = 1/count_j):
 * the products sum zd = fl32(colscale .* delta). */
#define TBSLA_PR_PATTERN_F32 4

typedef struct {
  double epsilon;
  double beta;
  int max_iteration;
  int precision;     /* tbsla_dtype: F64, F32, BF16 (values), PATTERN (value-free), or
                        TBSLA_PR_PATTERN_F32 (value-free FP32 deltas on an FP64 anchor) */
  int rule;          /* 0: reference residual < epsilon; 1: sum-normalised bound */
  int check_every;   /* iterations between convergence reads */
  double anchor_tau; /* mixed precisions: move the anchor when the bound falls below anchor_tau
                        times its value at the last move; 0: the default (3e-6 for F32 and
                        PATTERN_F32, 1e-2 for BF16) */
  const int *nodes;  /* personalised teleport set, or NULL */
  int n_nodes;
  int gather;        /* CSR products: 0 auto (propagation-blocked for n >= 2^16),
                        1 fixed-order entry tiles (the reference's row sums),
                        2 propagation-blocked */
} tbsla_pagerank_params;

typedef struct {
  int iterations;
  int anchors;
  double last_error;
  double bound;
  double seconds;
} tbsla_pagerank_report;

/* Sum-normalised PageRank vector (n_col doubles) into device memory;
 * TBSLA_EINVAL, before any launch, when n_out < n_col. */
int tbsla_hip_page_rank(tbsla_matrix *m, const tbsla_pagerank_params *p, double *out_dev, size_t n_out,
                        void *stream, tbsla_pagerank_report *report);

/* PageRank with explicit device inputs (CSR, FP64): values_dev (n_values =
 * the stored entry count, CSR entry order) replaces the stored edge values,
 * teleport_dev (n_teleport = n_col) is the teleport distribution v, so that
 * the iteration is b = beta*A b + (1-beta)*(1^T b)*v.  Either may be NULL
 * (stored values / the reference's uniform teleport).  No reference
 * counterpart: the forward pass of the differentiable PageRank. */
int tbsla_hip_page_rank_ex(tbsla_matrix *m, const tbsla_pagerank_params *p, const double *values_dev,
                           size_t n_values, const double *teleport_dev, size_t n_teleport, double *out_dev,
                           size_t n_out, void *stream, tbsla_pagerank_report *report);

/* tbsla_hip_page_rank with no device memory allocation.  No counterpart in
 * the reference: for callers that cannot allocate (code under stream
 * capture, serving loops, harnesses that forbid hipMalloc).  Every buffer of
 * the run (iterates, state, partial sums, node flags) is a slice of one
 * static 40 MiB scratch buffer per GPU, held until the process ends; a run
 * on n_col nodes needs about 16 * n_col bytes (+ n_col + 4 * n_nodes for a
 * node set), so n_col up to about 2.4 million.  A larger run is refused with
 * TBSLA_EINVAL before any launch.  FP64 only (precision TBSLA_F64), with the
 * fixed-order tile products (gather 0 or 1; 2 is refused); the FP64 mirror
 * must exist already (tbsla_hip_prepare(m, TBSLA_F64)), else TBSLA_EINVAL.
 * The result equals tbsla_hip_page_rank's with gather = 1 bit for bit.
 * report->seconds is host time between stream synchronisations (no events
 * are created).  Calls that overlap on one GPU share the buffer and race. */
int tbsla_hip_page_rank_noalloc(tbsla_matrix *m, const tbsla_pagerank_params *p, double *out_dev, size_t n_out,
                                void *stream, tbsla_pagerank_report *report);

/* Self-check of the propagation-blocked layout that PageRank's CSR products
 * use (gather 0/2): builds it on the GPU (device sorts, the build page_rank
 * uses) and on the host cores, and compares every index array.  0 when they
 * are equal; TBSLA_EINVAL naming the first array that differs, or when the
 * matrix is not a square single-block CSR matrix.  No reference counterpart
 * (a test hook of this library's layout). */
int tbsla_hip_pb_layout_check(tbsla_matrix *m);

typedef struct {
  double beta;
  double epsilon;    /* stop when |lambda_k - lambda_{k-1}|_1 <= epsilon * |lambda_k|_1 */
  int max_iteration;
  int precision;     /* TBSLA_F32: FP32 transposed products; TBSLA_F64 */
  int check_every;   /* iterations between convergence reads */
} tbsla_adjoint_params;

typedef struct {
  int iterations;
  double last_change;
  double s;          /* 1^T M p */
  double seconds;
} tbsla_adjoint_report;

/* Gradient of L(p) through the PageRank fixed point p = M p / 1^T M p,
 * M = beta*A + (1-beta) v 1^T, given p (pr_dev, the forward result, n_col
 * doubles) and w = dL/dp (w_dev).  Writes dL/dA per stored CSR entry
 * (grad_values_dev, n_values doubles) and dL/dv (grad_teleport_dev, n_col);
 * either output may be NULL.  values_dev / teleport_dev as in
 * tbsla_hip_page_rank_ex.  Implicit adjoint fixed point
 * lambda = M^T (lambda - (p^T lambda) 1)/s + w on the transposed CSR
 * structure.  No reference counterpart. */
int tbsla_hip_page_rank_adjoint(tbsla_matrix *m, const tbsla_adjoint_params *p, const double *values_dev,
                                size_t n_values, const double *teleport_dev, size_t n_teleport,
                                const double *pr_dev, const double *w_dev, size_t n, double *grad_values_dev,
                                double *grad_teleport_dev, void *stream, tbsla_adjoint_report *report);

/* ---- batched personalised PageRank (no reference counterpart: the
 * reference runs Matrix::personalized_page_rank once per teleport set,
 * Matrix.cpp:63-114; here every graph and teleport set advances together) */
typedef struct tbsla_batch tbsla_batch;

/* precision: TBSLA_F64 or TBSLA_F32 (values, iterates and products; the
 * residual, maxima and sums are FP64 in both) */
tbsla_batch *tbsla_batch_create(int precision);
void tbsla_batch_destroy(tbsla_batch *b);
/* A square CSR graph with n_rhs teleport node sets: set k is
 * nodes[rhs_ptr[k] .. rhs_ptr[k+1]).  The matrix must stay alive until
 * tbsla_batch_prepare (or the first tbsla_batch_run) returns. */
int tbsla_batch_add_graph(tbsla_batch *b, const tbsla_matrix *m, const int *rhs_ptr, const int *nodes, int n_rhs);
int tbsla_batch_prepare(tbsla_batch *b);
int tbsla_batch_n_rhs(const tbsla_batch *b);
/* offsets[n_rhs + 1]: start of each teleport set's vector in the output */
int tbsla_batch_layout(const tbsla_batch *b, long *offsets, size_t n);

typedef struct {
  double epsilon;
  double beta;
  int max_iteration;
  int rule;          /* 0: reference residual < epsilon; 1: sum-normalised bound */
  int check_every;   /* iterations per captured hipGraph replay */
  int graph_capture; /* 1: replay a hipGraph; 0: plain launches */
  int autotune;      /* 1: pick the launch shape (cached by batch signature, else timed on the
                        first iterations of the run) */
} tbsla_batch_params;

typedef struct {
  int iterations;              /* largest iteration count over the teleport sets */
  double seconds;              /* time of the iteration loop, the autotuner's trial iterations included */
  double tune_seconds;         /* time of those trial iterations (first run of a batch signature) */
  int tune_iterations;         /* iterations run as autotuner trials (0: shape from the cache or set) */
  int tune_cached;             /* 1: the shape came from a batch of the same signature */
  double bytes_per_iteration;  /* algorithmic bytes of one iteration, all sets active */
  int wg_per_cu, unroll, xcd_map; /* launch shape in use */
  int tile_nnz, spmm_rows;        /* entries per tile, rows per SpMM work item in use */
} tbsla_batch_report;

/* Sum-normalised vectors of every teleport set into out_dev (FP64, layout
 * from tbsla_batch_layout); rhs_iterations (n_rhs_it >= n_rhs, or NULL)
 * receives each set's iteration count. */
int tbsla_batch_run(tbsla_batch *b, const tbsla_batch_params *p, double *out_dev, size_t n_out, int *rhs_iterations,
                    size_t n_rhs_it, void *stream, tbsla_batch_report *report);

#ifdef __cplusplus
}
#endif

#endif /* TBSLA_HIP_H */
