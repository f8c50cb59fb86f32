/*
 * tbsla_oracle.c -- CPU restatement of the TBSLA SpMV / PageRank path.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  See tbsla_oracle.h for what pins
 * each function.  Build: oracle/Makefile (gcc -O2 -ffp-contract=off).
 *
 * Citations are to the reference source tree (src/cpp/...).
 */
#define _POSIX_C_SOURCE 200809L
#include "tbsla_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* growable arrays (std::vector::push_back)                                  */
/* ------------------------------------------------------------------------ */
static void *grow(void *p, size_t *cap, size_t need, size_t elt) {
  if (need <= *cap) return p;
  size_t nc = *cap ? *cap : 16;
  while (nc < need) nc *= 2;
  void *q = realloc(p, nc * elt);
  if (!q) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
  *cap = nc;
  return q;
}
static void push_val(orc_mat *m, double v) {
  m->val = (double *)grow(m->val, &m->cap_val, m->n_val + 1, sizeof(double));
  m->val[m->n_val++] = v;
}
static void push_i0(orc_mat *m, int x) {
  m->idx0 = (int *)grow(m->idx0, &m->cap_idx0, m->n_idx0 + 1, sizeof(int));
  m->idx0[m->n_idx0++] = x;
}
static void push_i1(orc_mat *m, int x) {
  m->idx1 = (int *)grow(m->idx1, &m->cap_idx1, m->n_idx1 + 1, sizeof(int));
  m->idx1[m->n_idx1++] = x;
}
static void clear_arrays(orc_mat *m) { m->n_val = m->n_idx0 = m->n_idx1 = 0; }
static void resize_val_zero(orc_mat *m, size_t n) {
  m->val = (double *)grow(m->val, &m->cap_val, n ? n : 1, sizeof(double));
  memset(m->val, 0, n * sizeof(double));
  m->n_val = n;
}
static long lmin(long a, long b) { return a < b ? a : b; }
static long lmax(long a, long b) { return a > b ? a : b; }
static size_t zmin(size_t a, size_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* utils/range.cpp:7-15 (lnv), :21-26 (pflv)                                 */
/* ------------------------------------------------------------------------ */
size_t orc_lnv(size_t size, int local_pos, int number_pos) {
  size_t n = size / (size_t)number_pos;
  size_t mod = size % (size_t)number_pos;
  if ((size_t)local_pos < mod) n++;
  return n;
}
size_t orc_pflv(size_t size, int local_pos, int number_pos) {
  size_t mod = size % (size_t)number_pos;
  size_t n = orc_lnv(size, local_pos, number_pos) * (size_t)local_pos;
  if ((size_t)local_pos >= mod) n += mod;
  return n;
}

/* ------------------------------------------------------------------------ */
/* utils/values_generation.cpp:4-24 (cdiag_value)                            */
/* All arithmetic wraps modulo 2^64 exactly like the size_t/long mix there.  */
/* ------------------------------------------------------------------------ */
void orc_cdiag_value(size_t i, size_t nr, size_t nc, size_t cdiag,
                     uint64_t *row, uint64_t *col, double *v) {
  (void)nr;
  *v = 1.0;
  if (cdiag == 0) { *row = i; *col = i; return; }
  long t1 = lmax(lmin((long)nc - (long)cdiag, (long)cdiag), 0L);
  if (i < (size_t)t1) { *row = i; *col = i + cdiag; return; }
  long t2 = lmax((long)cdiag + 2 * ((long)nc - 2 * (long)cdiag), 0L);
  if (i < (size_t)t2) {
    long it = (long)((i + cdiag) / 2);
    *row = (uint64_t)it;
    *col = (i % 2 == 0) ? (uint64_t)it - cdiag : (uint64_t)it + cdiag;
    return;
  }
  long it = (long)(i - (size_t)((long)nc - 2 * (long)cdiag));
  if (cdiag > nc) it = (long)((size_t)it - (cdiag - nc));
  *row = (uint64_t)it;
  *col = (uint64_t)it - cdiag;
}

/* ------------------------------------------------------------------------ */
/* utils/values_generation.cpp:26-45 (cqmat_value), glibc rand_r            */
/* ------------------------------------------------------------------------ */
void orc_cqmat_value(size_t i, size_t nr, size_t nc, size_t c_, double q,
                     unsigned seed_mult, uint64_t *row, uint64_t *col, double *v) {
  unsigned int seedp = (unsigned int)i;
  if (seed_mult > 0) seedp = (unsigned int)((size_t)seed_mult * i);
  size_t c = zmin(nc, c_);
  size_t nfull = zmin(nc - c + 1, nr);
  *v = 1.0;
  if (i < c * nfull) {
    *row = i / c;
    if (((double)rand_r(&seedp) / RAND_MAX) < q)
      *col = (size_t)rand_r(&seedp) % nc;
    else
      *col = i / c + i % c;
  } else {
    size_t it = i - c * nfull;
    size_t curr_row = 1;
    while (curr_row < nr - nfull && it >= c - curr_row) {
      it -= (c - curr_row);
      curr_row++;
    }
    curr_row += nfull - 1;
    *row = curr_row;
    /* (double)(int / int): integer division, the reference quirk (:43) */
    if ((double)(rand_r(&seedp) / RAND_MAX) < q)
      *col = (size_t)rand_r(&seedp) % nc;
    else
      *col = curr_row + it;
  }
}

/* utils/values_generation.cpp:47-62 */
void orc_cqmat_sum_columns(size_t nr, size_t nc, size_t c, double q,
                           unsigned seed_mult, double *sum) {
  size_t gnv = 0;
  size_t cc = zmin(c, nc);
  for (size_t i = 0; i < zmin(nc - cc + 1, nr); i++) gnv += cc;
  for (size_t i = 0; i < zmin(nr, nc) - zmin(nc - cc + 1, nr); i++) gnv += cc - i - 1;
  for (size_t j = 0; j < nc; j++) sum[j] = 0;
  for (size_t i = 0; i < gnv; i++) {
    uint64_t r, cl; double v;
    orc_cqmat_value(i, nr, nc, c, q, seed_mult, &r, &cl, &v);
    sum[cl] += v;
  }
}

/* ------------------------------------------------------------------------ */
/* Project-defined counter-based generator (no reference counterpart).       */
/* SplitMix64 finalizer (Steele, Lea & Flood 2014; public-domain reference   */
/* code by S. Vigna).  u01(seed, idx) = top 53 bits of                       */
/* splitmix64(splitmix64(seed) ^ idx) times 2^-53, in [0, 1).                */
/* ------------------------------------------------------------------------ */
uint64_t orc_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
double orc_u01(uint64_t seed, uint64_t idx) {
  uint64_t h = orc_splitmix64(orc_splitmix64(seed) ^ idx);
  return (double)(h >> 11) * 0x1.0p-53;
}
void orc_uniform_vector(uint64_t seed, uint64_t first, size_t n, double *out) {
  for (size_t k = 0; k < n; k++) out[k] = orc_u01(seed, first + k);
}

/* ------------------------------------------------------------------------ */
/* utils/vector.cpp:16-29 (compute_spmv_res_cdiag)                           */
/* ------------------------------------------------------------------------ */
void orc_spmv_res_cdiag(int nr, int nc, int c, const double *v, double *r) {
  int i;
  for (i = 0; i < (c < nr ? c : nr); i++) r[i] = i < nc - c ? v[i + c] : 0;
  for (; i < (nr < nc - c ? nr : nc - c); i++) r[i] = c == 0 ? v[i] : v[i + c] + v[i - c];
  for (; i < nr; i++) r[i] = i < nc + c ? v[i - c] : 0;
}

/* ------------------------------------------------------------------------ */
/* lifetime                                                                  */
/* ------------------------------------------------------------------------ */
orc_mat *orc_new(int fmt) {
  orc_mat *m = (orc_mat *)calloc(1, sizeof(orc_mat));
  m->fmt = fmt;
  m->NR = m->NC = 1;
  return m;
}
void orc_free(orc_mat *m) {
  if (!m) return;
  free(m->val); free(m->idx0); free(m->idx1); free(m);
}
void orc_info(const orc_mat *m, long out[16]) {
  out[0] = m->fmt; out[1] = m->n_row; out[2] = m->n_col; out[3] = m->f_row;
  out[4] = m->f_col; out[5] = m->ln_row; out[6] = m->ln_col; out[7] = m->pr;
  out[8] = m->pc; out[9] = m->NR; out[10] = m->NC; out[11] = m->nnz;
  out[12] = m->max_col; out[13] = (long)m->n_val; out[14] = (long)m->n_idx0;
  out[15] = (long)m->n_idx1;
}
static void set_full_block(orc_mat *m, int n_row, int n_col) {
  m->n_row = n_row; m->n_col = n_col; m->ln_row = n_row; m->ln_col = n_col;
  m->f_row = 0; m->f_col = 0; m->pr = 0; m->pc = 0; m->NR = 1; m->NC = 1;
}

/* Array constructors: MatrixCSR.cpp:13-27 (values, rowptr, colidx),
 * MatrixCOO.cpp:30-44 and MatrixSCOO.cpp:12-26 (values, row, col).
 * The reference leaves nnz uninitialised there; this sets it to n_val.     */
orc_mat *orc_from_arrays(int fmt, int n_row, int n_col,
                         const double *val, size_t n_val,
                         const int *a, size_t na, const int *b, size_t nb) {
  orc_mat *m = orc_new(fmt);
  set_full_block(m, n_row, n_col);
  for (size_t k = 0; k < n_val; k++) push_val(m, val[k]);
  for (size_t k = 0; k < na; k++) push_i0(m, a[k]);
  for (size_t k = 0; k < nb; k++) push_i1(m, b[k]);
  m->nnz = (long)n_val;
  return m;
}

/* stable merge sort of a permutation by (row, col): csr.cpp:3-8 order */
static const int *g_row, *g_col;
static int key_less(int i, int j) {
  if (g_row[i] == g_row[j]) return g_col[i] < g_col[j];
  return g_row[i] < g_row[j];
}
static void msort(int *p, int *tmp, size_t n) {
  if (n < 2) return;
  size_t h = n / 2;
  msort(p, tmp, h);
  msort(p + h, tmp, n - h);
  size_t i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = key_less(p[j], p[i]) ? p[j++] : p[i++];
  while (i < h) tmp[k++] = p[i++];
  while (j < n) tmp[k++] = p[j++];
  memcpy(p, tmp, n * sizeof(int));
}
static int *sorted_perm(const orc_mat *coo) {
  size_t n = coo->n_val;
  int *p = (int *)malloc((n ? n : 1) * sizeof(int));
  int *t = (int *)malloc((n ? n : 1) * sizeof(int));
  for (size_t k = 0; k < n; k++) p[k] = (int)k;
  g_row = coo->idx0; g_col = coo->idx1;
  msort(p, t, n);
  free(t);
  return p;
}

/* Conversions from COO:
 *   CSR   MatrixCSR.cpp:29-74   (sort, permute, rowptr)
 *   ELL   MatrixELL.cpp:12-54   (sort, row-major padding to max_col)
 *   SCOO  MatrixSCOO.cpp:53-68  (copy)
 *   DENSE MatrixDENSE.cpp:11-28 (duplicates accumulate)                   */
orc_mat *orc_convert(const orc_mat *coo, int fmt) {
  orc_mat *m = orc_new(fmt);
  m->n_row = coo->n_row; m->n_col = coo->n_col;
  m->ln_row = coo->n_row; m->ln_col = coo->n_col;
  m->f_row = coo->f_row; m->f_col = coo->f_col;
  m->pr = coo->pr; m->pc = coo->pc; m->NR = coo->NR; m->NC = coo->NC;
  m->nnz = coo->nnz;
  size_t n = coo->n_val;
  if (fmt == ORC_SCOO || fmt == ORC_COO) {
    for (size_t k = 0; k < n; k++) { push_val(m, coo->val[k]); push_i0(m, coo->idx0[k]); push_i1(m, coo->idx1[k]); }
    return m;
  }
  if (fmt == ORC_DENSE) {
    resize_val_zero(m, (size_t)(m->ln_row * m->ln_col));
    for (size_t k = 0; k < n; k++)
      m->val[coo->idx0[k] * m->ln_col + coo->idx1[k]] += coo->val[k];
    return m;
  }
  int *p = sorted_perm(coo);
  if (fmt == ORC_CSR) {
    for (size_t k = 0; k < n; k++) { push_i1(m, coo->idx1[p[k]]); push_val(m, coo->val[p[k]]); }
    /* rowptr: counts per row then running sum (equivalent to :50-73) */
    for (int r = 0; r <= m->n_row; r++) push_i0(m, 0);
    for (size_t k = 0; k < n; k++) m->idx0[coo->idx0[p[k]] + 1]++;
    for (int r = 0; r < m->n_row; r++) m->idx0[r + 1] += m->idx0[r];
  } else if (fmt == ORC_ELL) {
    if (n == 0) {
      m->max_col = 0; m->nnz = 0;
    } else {
      int *nvrow = (int *)calloc((size_t)m->n_row, sizeof(int));
      for (size_t k = 0; k < n; k++) nvrow[coo->idx0[p[k]]]++;
      int mx = nvrow[0];
      for (int r = 1; r < m->n_row; r++) if (nvrow[r] > mx) mx = nvrow[r];
      m->max_col = mx;
      size_t tot = (size_t)m->n_row * (size_t)mx;
      resize_val_zero(m, tot);
      for (size_t k = 0; k < tot; k++) push_i1(m, 0);
      size_t incr = 0;
      for (int r = 0; r < m->n_row; r++)
        for (int j = 0; j < nvrow[r]; j++) {
          m->val[(size_t)r * mx + j] = coo->val[p[incr]];
          m->idx1[(size_t)r * mx + j] = coo->idx1[p[incr]];
          incr++;
        }
      free(nvrow);
    }
  }
  free(p);
  return m;
}

/* ------------------------------------------------------------------------ */
/* fill_cdiag                                                                 */
/* ------------------------------------------------------------------------ */
static void set_block(orc_mat *m, int n_row, int n_col, int pr, int pc, int NR, int NC) {
  m->n_row = n_row; m->n_col = n_col; m->pr = pr; m->pc = pc; m->NR = NR; m->NC = NC;
  m->ln_row = (int)orc_lnv((size_t)n_row, pr, NR);
  m->f_row = (int)orc_pflv((size_t)n_row, pr, NR);
  m->ln_col = (int)orc_lnv((size_t)n_col, pc, NC);
  m->f_col = (int)orc_pflv((size_t)n_col, pc, NC);
}
static int in_block(const orc_mat *m, long ii, long jj) {
  return ii >= m->f_row && ii < (long)m->f_row + m->ln_row &&
         jj >= m->f_col && jj < (long)m->f_col + m->ln_col;
}
/* count of in-block entries of the two cdiag diagonals (MatrixCSR.cpp:251-266) */
static long cdiag_block_count(const orc_mat *m, int cdiag) {
  long nv = 0;
  for (long i = m->f_row; i < (long)m->f_row + m->ln_row; i++) {
    if (in_block(m, i, i - cdiag)) nv++;
    if (cdiag != 0 && in_block(m, i, i + cdiag)) nv++;
  }
  return nv;
}

void orc_fill_cdiag(orc_mat *m, int n_row, int n_col, int cdiag,
                    int pr, int pc, int NR, int NC) {
  clear_arrays(m);
  if (m->fmt == ORC_COO) {
    /* MatrixCOO.cpp:294-330: 1-D split of the cdiag_value stream */
    m->n_row = n_row; m->n_col = n_col; m->pr = pr; m->pc = pc; m->NR = NR; m->NC = NC;
    m->ln_row = NR; m->ln_col = NC; m->f_row = 0; m->f_col = 0;
    long gnv = lmax(lmin(n_row, (long)n_col - cdiag), 0) + lmax(lmin((long)n_row - cdiag, n_col), 0);
    if (cdiag == 0) gnv /= 2;
    m->nnz = 0;
    if (gnv == 0) return;
    long s = (long)orc_pflv((size_t)gnv, pr * NC + pc, NR * NC);
    long n = (long)orc_lnv((size_t)gnv, pr * NC + pc, NR * NC);
    m->nnz = n;
    for (int i = (int)s; i < s + n; i++) {
      uint64_t r, c; double v;
      orc_cdiag_value((size_t)i, (size_t)n_row, (size_t)n_col, (size_t)cdiag, &r, &c, &v);
      int ri = (int)r, ci = (int)c;  /* push_back(int, int, double), :118-126 */
      if (ri >= n_row || ri < 0) continue;
      if (ci >= n_col || ci < 0) continue;
      push_val(m, v); push_i0(m, ri); push_i1(m, ci);
    }
    return;
  }
  set_block(m, n_row, n_col, pr, pc, NR, NC);
  if (m->fmt == ORC_ELL) m->max_col = cdiag == 0 ? 1 : 2;  /* MatrixELL.cpp:208-211 */
  long nv = cdiag_block_count(m, cdiag);
  if (nv == 0) return;
  if (m->fmt == ORC_CSR) {
    /* MatrixCSR.cpp:234-297 */
    long incr = 0;
    push_i0(m, (int)incr);
    for (long i = m->f_row; i < (long)m->f_row + m->ln_row; i++) {
      long jj = i - cdiag;
      if (in_block(m, i, jj)) { push_i1(m, (int)jj); push_val(m, 1); incr++; }
      if (cdiag != 0) {
        jj = i + cdiag;
        if (in_block(m, i, jj)) { push_i1(m, (int)jj); push_val(m, 1); incr++; }
      }
      push_i0(m, (int)incr);
    }
  } else if (m->fmt == ORC_SCOO) {
    /* MatrixSCOO.cpp:230-286 (push_back filters against n_row/n_col) */
    for (long i = m->f_row; i < (long)m->f_row + m->ln_row; i++) {
      long jj = i - cdiag;
      if (in_block(m, i, jj) && i < n_row && jj < n_col) { push_val(m, 1); push_i0(m, (int)i); push_i1(m, (int)jj); }
      if (cdiag != 0) {
        jj = i + cdiag;
        if (in_block(m, i, jj) && i < n_row && jj < n_col) { push_val(m, 1); push_i0(m, (int)i); push_i1(m, (int)jj); }
      }
    }
  } else if (m->fmt == ORC_ELL) {
    /* MatrixELL.cpp:192-260 */
    m->nnz = nv;
    for (long i = m->f_row; i < (long)m->f_row + m->ln_row; i++) {
      long jj = i - cdiag;
      if (in_block(m, i, jj)) { push_i1(m, (int)jj); push_val(m, 1); }
      else { push_i1(m, 0); push_val(m, 0); }
      if (cdiag != 0) {
        jj = i + cdiag;
        if (in_block(m, i, jj)) { push_i1(m, (int)jj); push_val(m, 1); }
        else { push_i1(m, 0); push_val(m, 0); }
      }
    }
  } else if (m->fmt == ORC_DENSE) {
    /* MatrixDENSE.cpp:137-190 */
    m->nnz = nv;
    resize_val_zero(m, (size_t)(m->ln_col * m->ln_row));
    for (long i = m->f_row; i < (long)m->f_row + m->ln_row; i++) {
      long jj = i - cdiag;
      if (in_block(m, i, jj)) m->val[(i - m->f_row) * m->ln_col + (jj - m->f_col)] = 1;
      if (cdiag != 0) {
        jj = i + cdiag;
        if (in_block(m, i, jj)) m->val[(i - m->f_row) * m->ln_col + (jj - m->f_col)] = 1;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* fill_cqmat                                                                 */
/* ------------------------------------------------------------------------ */
static double entry_value(uint64_t value_seed, long gidx, double v) {
  return value_seed ? 1.0 - orc_u01(value_seed, (uint64_t)gidx) : v;
}

void orc_fill_cqmat(orc_mat *m, int n_row, int n_col, int c, double q,
                    unsigned seed_mult, int pr, int pc, int NR, int NC,
                    uint64_t value_seed) {
  clear_arrays(m);
  int cc = c < n_col ? c : n_col;
  if (m->fmt == ORC_COO) {
    /* MatrixCOO.cpp:332-375 */
    m->n_row = n_row; m->n_col = n_col; m->pr = pr; m->pc = pc; m->NR = NR; m->NC = NC;
    m->ln_row = NR; m->ln_col = NC; m->f_row = 0; m->f_col = 0;
    long gnv = 0;
    for (long i = 0; i < lmin((long)n_col - cc + 1, n_row); i++) gnv += cc;
    for (long i = 0; i < lmin(n_row, n_col) - lmin((long)n_col - cc + 1, n_row); i++) gnv += cc - i - 1;
    if (gnv == 0) return;
    long s = (long)orc_pflv((size_t)gnv, pr * NC + pc, NR * NC);
    long n = (long)orc_lnv((size_t)gnv, pr * NC + pc, NR * NC);
    m->nnz = n;
    for (long i = s; i < s + n; i++) {
      uint64_t r, cl; double v;
      orc_cqmat_value((size_t)i, (size_t)n_row, (size_t)n_col, (size_t)c, q, seed_mult, &r, &cl, &v);
      int ri = (int)r, ci = (int)cl;
      if (ri >= n_row || ri < 0 || ci >= n_col || ci < 0) continue;
      push_val(m, entry_value(value_seed, i, v)); push_i0(m, ri); push_i1(m, ci);
    }
    return;
  }
  set_block(m, n_row, n_col, pr, pc, NR, NC);
  const long f_row = m->f_row, ln_row = m->ln_row;
  /* first generation index of the block and candidate count
   * (MatrixCSR.cpp:317-341, identical in SCOO/ELL/DENSE) */
  int min_ = (int)lmin((long)n_col - cc + 1, n_row);
  long incr = 0, nv = 0;
  for (long i = 0; i < min_; i++) {
    if (i < f_row) incr += cc;
    if (i >= f_row && i < f_row + ln_row) nv += cc;
    if (i >= f_row + ln_row) break;
  }
  for (long i = 0; i < lmin(n_row, n_col) - min_; i++) {
    if (i + min_ < f_row) incr += cc - i - 1;
    if (i + min_ >= f_row && i + min_ < f_row + ln_row) nv += cc - i - 1;
    if (i + min_ >= f_row + ln_row) break;
  }
  const long incr_save = incr;
  const long lim3 = lmin(lmin(n_row, n_col), f_row + ln_row);
#define GEN(idx) uint64_t gr_, gc_; double gv_; \
  orc_cqmat_value((size_t)(idx), (size_t)n_row, (size_t)n_col, (size_t)c, q, seed_mult, &gr_, &gc_, &gv_); \
  long ii = (long)gr_, jj = (long)gc_; (void)ii; (void)jj

  if (m->fmt == ORC_CSR || m->fmt == ORC_SCOO) {
    /* nnz over-count pass: MatrixCSR.cpp:343-358, MatrixSCOO.cpp:331-347 */
    m->nnz = 0;
    long i;
    for (i = f_row; i < lim3; i++)
      for (long j = 0; j < cc; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) m->nnz++;
        incr++;
      }
    incr = incr_save;
    if (nv == 0) return;
    long lincr = 0;
    if (m->fmt == ORC_CSR) push_i0(m, 0);
    for (i = f_row; i < lmin(min_, f_row + ln_row); i++) {
      for (long j = 0; j < cc; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) {
          double v = entry_value(value_seed, incr, gv_);
          if (m->fmt == ORC_CSR) { push_i1(m, (int)jj); push_val(m, v); lincr++; }
          else if (ii < n_row && jj < n_col) { push_val(m, v); push_i0(m, (int)ii); push_i1(m, (int)jj); }
        }
        incr++;
      }
      if (m->fmt == ORC_CSR) push_i0(m, (int)lincr);
    }
    for (; i < lim3; i++) {
      for (long j = 0; j < cc - i + min_ - 1; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) {
          double v = entry_value(value_seed, incr, gv_);
          if (m->fmt == ORC_CSR) { push_i1(m, (int)jj); push_val(m, v); lincr++; }
          else if (ii < n_row && jj < n_col) { push_val(m, v); push_i0(m, (int)ii); push_i1(m, (int)jj); }
        }
        incr++;
      }
      if (m->fmt == ORC_CSR) push_i0(m, (int)lincr);
    }
    return;
  }
  if (m->fmt == ORC_ELL) {
    /* MatrixELL.cpp:262-392 */
    m->nnz = 0;
    m->max_col = 0;
    long i;
    for (i = f_row; i < lmin(min_, f_row + ln_row); i++) {
      int nbc = 0;
      for (long j = 0; j < cc; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) { m->nnz++; nbc++; }
        incr++;
      }
      if (nbc > m->max_col) m->max_col = nbc;
    }
    for (; i < lim3; i++) {
      int nbc = 0;
      long j = 0;
      for (; j < cc - i + min_ - 1; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) { nbc++; m->nnz++; }
        incr++;
      }
      for (; j < cc; j++)  /* slot index compared with the column range (:338-342) */
        if (j >= m->f_col && j < (long)m->f_col + m->ln_col) nbc++;
      if (nbc > m->max_col) m->max_col = nbc;
    }
    if (nv == 0) return;
    incr = incr_save;
    for (i = f_row; i < lmin(min_, f_row + ln_row); i++) {
      long lincr = 0;
      for (long j = 0; j < cc; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) { push_i1(m, (int)jj); push_val(m, entry_value(value_seed, incr, gv_)); lincr++; }
        incr++;
      }
      for (long j = lincr; j < m->max_col; j++) { push_i1(m, 0); push_val(m, 0); }
    }
    for (; i < lmin(n_row, f_row + ln_row); i++) {
      long lincr = 0;
      for (long j = 0; j < cc - i + min_ - 1; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) { push_i1(m, (int)jj); push_val(m, entry_value(value_seed, incr, gv_)); lincr++; }
        incr++;
      }
      for (long j = lincr; j < m->max_col; j++) { push_i1(m, 0); push_val(m, 0); }
    }
    return;
  }
  if (m->fmt == ORC_DENSE) {
    /* MatrixDENSE.cpp:192-271 */
    if (nv == 0) return;
    m->nnz = 0;
    resize_val_zero(m, (size_t)(m->ln_col * m->ln_row));
    long i;
    for (i = f_row; i < lmin(min_, f_row + ln_row); i++)
      for (long j = 0; j < cc; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) {
          m->val[(ii - f_row) * m->ln_col + (jj - m->f_col)] += entry_value(value_seed, incr, gv_);
          m->nnz++;
        }
        incr++;
      }
    for (; i < lmin(n_row, f_row + ln_row); i++)
      for (long j = 0; j < cc - i + min_ - 1; j++) {
        GEN(incr);
        if (in_block(m, ii, jj)) {
          m->val[(ii - f_row) * m->ln_col + (jj - m->f_col)] += entry_value(value_seed, incr, gv_);
          m->nnz++;
        }
        incr++;
      }
  }
#undef GEN
}

/* fill_cqmat_stochastic: COO :377-383, CSR :411-421, SCOO :393-399,
 * ELL :394-404, DENSE :273-286 (global column counts) */
void orc_fill_cqmat_stochastic(orc_mat *m, int n_row, int n_col, int c, double q,
                               unsigned seed_mult, int pr, int pc, int NR, int NC) {
  orc_fill_cqmat(m, n_row, n_col, c, q, seed_mult, pr, pc, NR, NC, 0);
  if (m->fmt != ORC_COO && m->fmt != ORC_SCOO && m->n_val == 0) return;
  double *sum = (double *)malloc(((size_t)n_col ? (size_t)n_col : 1) * sizeof(double));
  orc_cqmat_sum_columns((size_t)n_row, (size_t)n_col, (size_t)c, q, seed_mult, sum);
  if (m->fmt == ORC_COO || m->fmt == ORC_SCOO) {
    for (size_t k = 0; k < m->n_val; k++) m->val[k] /= sum[m->idx1[k]];
  } else if (m->fmt == ORC_CSR) {
    for (long i = 0; i < (long)m->n_idx0 - 1; i++)
      for (long j = m->idx0[i]; j < m->idx0[i + 1]; j++) m->val[j] /= sum[m->idx1[j]];
  } else if (m->fmt == ORC_ELL) {
    long rows = (long)zmin((size_t)m->ln_row, m->n_val / (size_t)m->max_col);
    for (long i = 0; i < rows; i++)
      for (long j = 0; j < m->max_col; j++)
        if (m->val[i * m->max_col + j] != 0) m->val[i * m->max_col + j] /= sum[m->idx1[i * m->max_col + j]];
  } else if (m->fmt == ORC_DENSE) {
    for (long k = 0; k < n_col; k++) if (sum[k] == 0) sum[k] = 1;
    for (long i = 0; i < m->ln_row; i++)
      for (long j = 0; j < m->ln_col; j++) m->val[i * m->ln_col + j] /= sum[m->f_col + j];
  }
  free(sum);
}

/* normalize_columns: COO :385-393, CSR :423-435, SCOO :401-409,
 * ELL :406-422, DENSE :288-300 (local column sums) */
void orc_normalize_columns(orc_mat *m) {
  size_t ns = (size_t)(m->fmt == ORC_COO ? m->n_col : m->ln_col);
  double *sum = (double *)calloc(ns ? ns : 1, sizeof(double));
  if (m->fmt == ORC_COO) {
    for (size_t k = 0; k < m->n_val; k++) sum[m->idx1[k]] += m->val[k];
    for (size_t k = 0; k < m->n_val; k++) m->val[k] /= sum[m->idx1[k]];
  } else if (m->fmt == ORC_SCOO) {
    for (size_t k = 0; k < m->n_val; k++) sum[m->idx1[k] - m->f_col] += m->val[k];
    for (size_t k = 0; k < m->n_val; k++) m->val[k] /= sum[m->idx1[k] - m->f_col];
  } else if (m->fmt == ORC_CSR) {
    for (long i = 0; i < (long)m->n_idx0 - 1; i++)
      for (long j = m->idx0[i]; j < m->idx0[i + 1]; j++) sum[m->idx1[j] - m->f_col] += m->val[j];
    for (long i = 0; i < (long)m->n_idx0 - 1; i++)
      for (long j = m->idx0[i]; j < m->idx0[i + 1]; j++) m->val[j] /= sum[m->idx1[j] - m->f_col];
  } else if (m->fmt == ORC_ELL) {
    long rows = m->max_col ? (long)zmin((size_t)m->ln_row, m->n_val / (size_t)m->max_col) : 0;
    for (long i = 0; i < rows; i++)
      for (long j = 0; j < m->max_col; j++) {
        long idx = m->idx1[i * m->max_col + j] - m->f_col;
        if (idx < 0) idx = 0;
        sum[idx] += m->val[i * m->max_col + j];
      }
    for (long i = 0; i < rows; i++)
      for (long j = 0; j < m->max_col; j++) {
        long idx = m->idx1[i * m->max_col + j] - m->f_col;
        if (idx < 0) idx = 0;
        m->val[i * m->max_col + j] /= sum[idx];
      }
  } else if (m->fmt == ORC_DENSE) {
    for (long i = 0; i < m->ln_row; i++)
      for (long j = 0; j < m->ln_col; j++) sum[j] += m->val[i * m->ln_col + j];
    for (long i = 0; i < m->ln_row; i++)
      for (long j = 0; j < m->ln_col; j++) m->val[i * m->ln_col + j] /= sum[j];
  }
  free(sum);
}

/* ------------------------------------------------------------------------ */
/* Ax / spmv                                                                  */
/*   CSR   MatrixCSR.cpp:123-140   (assign, ordered row sums)                 */
/*   ELL   MatrixELL.cpp:95-114    (assign, negative local index -> 0)        */
/*   COO   MatrixCOO.cpp:101-116   (accumulate, global ids)                   */
/*   SCOO  MatrixSCOO.cpp:108-123  (accumulate, block-local access)           */
/*   DENSE MatrixDENSE.cpp:58-73   (accumulate)                               */
/* Row-parallel formats keep the per-row order under OpenMP; COO/SCOO are   */
/* restated with the single-thread order of their vector reduction.         */
/* ------------------------------------------------------------------------ */
void orc_Ax(const orc_mat *m, double *r, const double *v, int vect_incr) {
  if (m->fmt == ORC_CSR) {
    if (m->n_val == 0) return;
    const int *rp = m->idx0, *ci = m->idx1;
    const double *val = m->val;
    const long nrows = (long)m->n_idx0 - 1;
    const int f_col = m->f_col;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < nrows; i++) {
      double tmp = 0;
      for (long j = rp[i] - rp[0]; j < rp[i + 1] - rp[0]; j++) tmp += val[j] * v[ci[j] - f_col];
      r[i] = tmp;
    }
  } else if (m->fmt == ORC_ELL) {
    if (m->nnz == 0 || m->max_col == 0) return;
    const long K = m->max_col;
    const long rows = (long)zmin((size_t)m->ln_row, m->n_val / (size_t)K);
    const int *col = m->idx1;
    const double *val = m->val;
    const int f_col = m->f_col;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < rows; i++) {
      double tmp = 0;
      for (long j = 0; j < K; j++) {
        int idx = col[i * K + j] - f_col;
        if (idx < 0) idx = 0;
        tmp += val[i * K + j] * v[idx];
      }
      r[i] = tmp;
    }
  } else if (m->fmt == ORC_COO) {
    for (size_t k = 0; k < m->n_val; k++) r[m->idx0[k] + vect_incr] += m->val[k] * v[m->idx1[k]];
  } else if (m->fmt == ORC_SCOO) {
    for (size_t k = 0; k < m->n_val; k++)
      r[m->idx0[k] + vect_incr - m->f_row] += m->val[k] * v[m->idx1[k] - m->f_col];
  } else if (m->fmt == ORC_DENSE) {
    if (m->nnz == 0 || m->n_val == 0) return;
    const long lr = m->ln_row, lc = m->ln_col;
    const double *val = m->val;
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < lr; i++)
      for (long j = 0; j < lc; j++) r[i] += val[i * lc + j] * v[j];
  }
}

void orc_spmv(const orc_mat *m, const double *v, int vect_incr, double *out) {
  memset(out, 0, (size_t)m->n_row * sizeof(double));
  orc_Ax(m, out, v, vect_incr);
}

/* Matrix.cpp:6-11 (square matrices) */
void orc_a_axpx(const orc_mat *m, const double *v, int vect_incr, double *out) {
  size_t n = (size_t)m->n_row;
  double *r = (double *)malloc((n ? n : 1) * sizeof(double));
  orc_spmv(m, v, vect_incr, r);
  for (size_t i = 0; i < n; i++) r[i] = r[i] + v[i];
  orc_spmv(m, r, vect_incr, out);
  free(r);
}

/* Matrix.cpp:13-114.  Per iteration (reference order):
 *   b_t = b; b = A b_t; tele = (1-beta)/n * sum(b_t)        (personalised:
 *   (1-beta)/|nodes| * sum(b_t), added to listed nodes only); b = beta*b + tele;
 *   max = max(b); b /= max; err = sum|b - b_t|; stop when err < epsilon;
 *   at the end b /= sum(b).
 * The arithmetic of every element and the order of every sum that feeds the
 * iterate (sum(b_t), sum(b)) or the stopping test (err) is the reference's:
 * those sums run serially in ascending i.  Elementwise steps and the max
 * (exact whatever the order) run in parallel, and sum(b_t) of the next
 * iteration is accumulated in the same serial loop as err, so the iterate is
 * bit-identical to the reference's at a fraction of its serial passes. */
static int orc_pr_loop(const orc_mat *m, double epsilon, double beta, int max_iteration,
                       const char *member, double tele_div, double b0, double *out,
                       double *last_error) {
  const long n = m->n_col;
  double *buf[2];
  buf[0] = out;
  buf[1] = (double *)malloc((size_t)(n ? n : 1) * sizeof(double));
  double *b = buf[0];
  for (long i = 0; i < n; i++) b[i] = b0;
  double sum_bt = 0.0;                       /* sum(b_t) of the next iteration */
  for (long i = 0; i < n; i++) sum_bt += b[i];
  int converge = 0, it = 0, cur = 0;
  double error = 0;
  while (!converge && it < max_iteration) {
    double *bt = buf[cur], *nb = buf[cur ^ 1];
    orc_spmv(m, bt, 0, nb);
    const double tele = sum_bt * ((1 - beta) / tele_div);
    double max = -INFINITY;
    if (member) {
      #pragma omp parallel for schedule(static) reduction(max:max)
      for (long i = 0; i < n; i++) {
        nb[i] = beta * nb[i];
        if (member[i]) nb[i] += tele;
        if (max < nb[i]) max = nb[i];
      }
    } else {
      #pragma omp parallel for schedule(static) reduction(max:max)
      for (long i = 0; i < n; i++) {
        nb[i] = beta * nb[i] + tele;
        if (max < nb[i]) max = nb[i];
      }
    }
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) nb[i] = nb[i] / max;
    error = 0.0;
    sum_bt = 0.0;
    for (long i = 0; i < n; i++) {
      error += fabs(nb[i] - bt[i]);
      sum_bt += nb[i];
    }
    if (error < epsilon) converge = 1;
    it++;
    cur ^= 1;
  }
  b = buf[cur];
  /* final sum normalisation: the same serial sum as sum(b_t) above */
  const double sum = sum_bt;
  #pragma omp parallel for schedule(static)
  for (long i = 0; i < n; i++) out[i] = b[i] / sum;
  if (last_error) *last_error = error;
  free(buf[1]);
  return it;
}

/* Matrix.cpp:13-61 (square matrices; n_row == n_col) */
int orc_page_rank(const orc_mat *m, double epsilon, double beta, int max_iteration,
                  double *out, double *last_error) {
  return orc_pr_loop(m, epsilon, beta, max_iteration, NULL, (double)m->n_col, 1.0, out, last_error);
}

/* Matrix.cpp:63-114 */
int orc_personalized_page_rank(const orc_mat *m, double epsilon, double beta,
                               int max_iteration, const int *nodes, int n_nodes,
                               double *out, double *last_error) {
  const int n = m->n_col;
  char *member = (char *)calloc((size_t)n ? (size_t)n : 1, 1);
  for (int k = 0; k < n_nodes; k++) if (nodes[k] >= 0 && nodes[k] < n) member[nodes[k]] = 1;
  const int it = orc_pr_loop(m, epsilon, beta, max_iteration, member, (double)(size_t)n_nodes, 0.25, out,
                             last_error);
  free(member);
  return it;
}

/* ------------------------------------------------------------------------ */
/* Binary format: MatrixCSR.cpp:179-232, MatrixCOO.cpp:174-228,              */
/* MatrixSCOO.cpp:174-228, MatrixELL.cpp:142-190, MatrixDENSE.cpp:98-135.    */
/* Header: n_row n_col ln_row ln_col f_row f_col (int) nnz (long)           */
/* pr pc NR NC (int) [ELL: max_col (int)], then size_t-prefixed arrays.      */
/* ------------------------------------------------------------------------ */
static void wr_arr(FILE *f, const void *p, size_t n, size_t elt) {
  fwrite(&n, sizeof(size_t), 1, f);
  if (n) fwrite(p, elt, n, f);
}
long orc_write(const orc_mat *m, const char *path) {
  FILE *f = fopen(path, "ab");
  if (!f) return -1;
  int h1[6] = {m->n_row, m->n_col, m->ln_row, m->ln_col, m->f_row, m->f_col};
  int h2[4] = {m->pr, m->pc, m->NR, m->NC};
  fwrite(h1, sizeof(int), 6, f);
  fwrite(&m->nnz, sizeof(long), 1, f);
  fwrite(h2, sizeof(int), 4, f);
  if (m->fmt == ORC_ELL) fwrite(&m->max_col, sizeof(int), 1, f);
  wr_arr(f, m->val, m->n_val, sizeof(double));
  if (m->fmt == ORC_CSR || m->fmt == ORC_COO || m->fmt == ORC_SCOO) wr_arr(f, m->idx0, m->n_idx0, sizeof(int));
  if (m->fmt != ORC_DENSE) wr_arr(f, m->idx1, m->n_idx1, sizeof(int));
  long pos = ftell(f);
  fclose(f);
  return pos;
}
static int rd_arr(FILE *f, void **p, size_t *n, size_t *cap, size_t elt) {
  size_t k;
  if (fread(&k, sizeof(size_t), 1, f) != 1) return -1;
  *p = grow(*p, cap, k ? k : 1, elt);
  if (k && fread(*p, elt, k, f) != k) return -1;
  *n = k;
  return 0;
}
long orc_read(orc_mat *m, const char *path, long seek) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  fseek(f, seek, SEEK_SET);
  int h1[6], h2[4];
  long rc = 0;
  if (fread(h1, sizeof(int), 6, f) != 6 || fread(&m->nnz, sizeof(long), 1, f) != 1 ||
      fread(h2, sizeof(int), 4, f) != 4) rc = -1;
  m->n_row = h1[0]; m->n_col = h1[1]; m->ln_row = h1[2]; m->ln_col = h1[3];
  m->f_row = h1[4]; m->f_col = h1[5];
  m->pr = h2[0]; m->pc = h2[1]; m->NR = h2[2]; m->NC = h2[3];
  if (!rc && m->fmt == ORC_ELL && fread(&m->max_col, sizeof(int), 1, f) != 1) rc = -1;
  if (!rc) rc = rd_arr(f, (void **)&m->val, &m->n_val, &m->cap_val, sizeof(double));
  if (!rc && (m->fmt == ORC_CSR || m->fmt == ORC_COO || m->fmt == ORC_SCOO))
    rc = rd_arr(f, (void **)&m->idx0, &m->n_idx0, &m->cap_idx0, sizeof(int));
  if (!rc && m->fmt != ORC_DENSE) rc = rd_arr(f, (void **)&m->idx1, &m->n_idx1, &m->cap_idx1, sizeof(int));
  long pos = rc ? -1 : ftell(f);
  fclose(f);
  return pos;
}
