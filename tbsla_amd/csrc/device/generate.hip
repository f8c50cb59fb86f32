// GPU generation of cqmat CSR blocks, bit-identical to the host generators.
//
// gen_count      one thread per row of the block: in-block entries of the row
// gen_fill       one thread per row: columns (and overlay values) written at
//                the row's offset, in the row's generation order
// gen_colcount   one thread per row of the whole matrix: entries per column
//                (the divisor of fill_cqmat_stochastic)
// gen_stochastic values = 1 / count of the entry's column
//
// The entry stream restates values_generation.cpp:26-45 entry by entry with
// the stream index i of the reference; its random numbers come from
// dev_rand_r, a restatement of glibc's rand_r (glibc 2.35,
// stdlib/rand_r.c), which the reference calls on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>

#include "common.hpp"
#include "generate.hpp"

namespace tbsla {
namespace hip {
namespace {

constexpr int kGenThreads = 256;

// glibc rand_r: three steps of the LCG next = next * 1103515245 + 12345,
// taking 11, 10 and 10 bits of next / 65536.
__device__ inline int dev_rand_r(unsigned int *seed) {
  unsigned int next = *seed;
  int result;
  next *= 1103515245u;
  next += 12345u;
  result = static_cast<int>((next / 65536u) % 2048u);
  next *= 1103515245u;
  next += 12345u;
  result <<= 10;
  result ^= static_cast<int>((next / 65536u) % 1024u);
  next *= 1103515245u;
  next += 12345u;
  result <<= 10;
  result ^= static_cast<int>((next / 65536u) % 1024u);
  *seed = next;
  return result;
}

constexpr int kRandMax = 2147483647;  // glibc RAND_MAX

// Stream geometry and block bounds (the fields of host::CqmatBlock).
struct Geo {
  unsigned long long nr, nc, c;  // n_row, n_col, c as the reference's size_t
  double q;
  unsigned seed;
  unsigned long long value_seed;
  long f_row, ln_row, f_col, ln_col;
  long cc, min_, lim;  // CqmatStream: min(c, n_col), full rows, min(n_row, n_col)

  __device__ long candidates(long r) const {
    if (r < min_) return cc;
    if (r >= lim) return 0;
    return max(0L, cc - (r - min_) - 1);
  }
  __device__ long base(long r) const {
    if (r <= min_) return r * cc;
    const long k = min(r, lim) - min_;
    long s = 0;
    if (k > 0) {
      const long kk = min(k, cc > 0 ? cc - 1 : 0L);
      s = kk * (cc - 1) - kk * (kk - 1) / 2;
    }
    return min_ * cc + s;
  }
  // values_generation.cpp:26-45 for stream entry i
  __device__ void entry(unsigned long long i, long &row, long &col) const {
    unsigned int seedp = static_cast<unsigned int>(i);
    if (seed > 0) seedp = static_cast<unsigned int>(static_cast<unsigned long long>(seed) * i);
    const unsigned long long cq = nc < c ? nc : c;
    const unsigned long long nfull = (nc - cq + 1) < nr ? (nc - cq + 1) : nr;
    if (i < cq * nfull) {
      row = static_cast<long>(i / cq);
      if ((static_cast<double>(dev_rand_r(&seedp)) / kRandMax) < q)
        col = static_cast<long>(static_cast<unsigned long long>(dev_rand_r(&seedp)) % nc);
      else
        col = static_cast<long>(i / cq + i % cq);
      return;
    }
    unsigned long long it = i - cq * nfull;
    unsigned long long curr_row = 1;
    while (curr_row < nr - nfull && it >= cq - curr_row) {
      it -= (cq - curr_row);
      curr_row++;
    }
    curr_row += nfull - 1;
    row = static_cast<long>(curr_row);
    // integer division before the conversion, as values_generation.cpp:43
    if (static_cast<double>(dev_rand_r(&seedp) / kRandMax) < q)
      col = static_cast<long>(static_cast<unsigned long long>(dev_rand_r(&seedp)) % nc);
    else
      col = static_cast<long>(curr_row + it);
  }
  __device__ bool in_block(long ii, long jj) const {
    return ii >= f_row && ii < f_row + ln_row && jj >= f_col && jj < f_col + ln_col;
  }
};

// u01(seed, i) = (splitmix64(splitmix64(seed) ^ i) >> 11) * 2^-53, the value
// overlay of the host generators (1 - u01)
__device__ inline unsigned long long splitmix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ __launch_bounds__(kGenThreads) void gen_count(Geo g, long nrows, int *__restrict__ counts) {
  const long t = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (t >= nrows) return;
  const long r = g.f_row + t;
  const long i0 = g.base(r), nc = g.candidates(r);
  int kept = 0;
  for (long k = 0; k < nc; k++) {
    long ii, jj;
    g.entry(static_cast<unsigned long long>(i0 + k), ii, jj);
    kept += g.in_block(ii, jj) ? 1 : 0;
  }
  counts[t] = kept;
}

__global__ __launch_bounds__(kGenThreads) void gen_fill(Geo g, long nrows, const int *__restrict__ rowptr,
                                                        int *__restrict__ colidx, double *__restrict__ vals,
                                                        unsigned long long vseed_mixed) {
  const long t = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (t >= nrows) return;
  const long r = g.f_row + t;
  const long i0 = g.base(r), nc = g.candidates(r);
  long o = rowptr[t];
  for (long k = 0; k < nc; k++) {
    long ii, jj;
    g.entry(static_cast<unsigned long long>(i0 + k), ii, jj);
    if (!g.in_block(ii, jj)) continue;
    colidx[o] = static_cast<int>(jj);
    double v = 1.0;
    if (g.value_seed) {
      const unsigned long long h = splitmix64(vseed_mixed ^ static_cast<unsigned long long>(i0 + k));
      v = 1.0 - static_cast<double>(h >> 11) * 0x1.0p-53;
    }
    vals[o] = v;
    o++;
  }
}

__global__ __launch_bounds__(kGenThreads) void gen_colcount(Geo g, unsigned *__restrict__ cnt) {
  const long r = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (r >= g.lim) return;
  const long i0 = g.base(r), nc = g.candidates(r);
  for (long k = 0; k < nc; k++) {
    long ii, jj;
    g.entry(static_cast<unsigned long long>(i0 + k), ii, jj);
    atomicAdd(cnt + jj, 1u);
  }
}

__global__ __launch_bounds__(kGenThreads) void gen_stochastic(long nnz, const int *__restrict__ colidx,
                                                              const unsigned *__restrict__ cnt,
                                                              double *__restrict__ vals) {
  const long j = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (j >= nnz) return;
  vals[j] = vals[j] / static_cast<double>(cnt[colidx[j]]);
}

// COO: stream entries [first, first + n) with global row and column ids
// and the overlay value or 1.0 (MatrixCOO.cpp:332-375)
__global__ __launch_bounds__(kGenThreads) void gen_coo(Geo g, long first, long n, int *__restrict__ row,
                                                       int *__restrict__ col, double *__restrict__ vals,
                                                       unsigned long long vseed_mixed) {
  const long k = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (k >= n) return;
  const unsigned long long i = static_cast<unsigned long long>(first + k);
  long ii, jj;
  g.entry(i, ii, jj);
  row[k] = static_cast<int>(ii);
  col[k] = static_cast<int>(jj);
  double v = 1.0;
  if (g.value_seed) v = 1.0 - static_cast<double>(splitmix64(vseed_mixed ^ i) >> 11) * 0x1.0p-53;
  vals[k] = v;
}

// SCOO: the global row id of every entry of a generated CSR block
__global__ __launch_bounds__(kGenThreads) void gen_row_ids(long nrows, long f_row, const int *__restrict__ rowptr,
                                                           int *__restrict__ rows) {
  const long t = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (t >= nrows) return;
  for (int o = rowptr[t]; o < rowptr[t + 1]; o++) rows[o] = static_cast<int>(f_row + t);
}

// ELL: row t's entries in slots 0.. of its K slots, the rest (column 0,
// value 0.0), row-major (MatrixELL.cpp:262-392)
__global__ __launch_bounds__(kGenThreads) void gen_ell_spread(long nrows, long ngen, int K,
                                                              const int *__restrict__ rowptr,
                                                              const int *__restrict__ colidx,
                                                              const double *__restrict__ vals, int *__restrict__ ecol,
                                                              double *__restrict__ eval) {
  const long slot = static_cast<long>(blockIdx.x) * kGenThreads + threadIdx.x;
  if (slot >= nrows * K) return;
  const long t = slot / K;
  const int j = static_cast<int>(slot - t * K);
  // rows past the generated ones (no candidates) are padding only
  const int a = t < ngen ? rowptr[t] : 0, len = t < ngen ? rowptr[t + 1] - a : 0;
  ecol[slot] = j < len ? colidx[a + j] : 0;
  eval[slot] = j < len ? vals[a + j] : 0.0;
}

unsigned grid_of(long n) { return static_cast<unsigned>((n + kGenThreads - 1) / kGenThreads); }

Geo geo_of(const host::CqmatBlock &b) {
  Geo g{};
  g.nr = static_cast<unsigned long long>(b.n_row);
  g.nc = static_cast<unsigned long long>(b.n_col);
  g.c = static_cast<unsigned long long>(b.c);
  g.q = b.q;
  g.seed = b.seed;
  g.value_seed = b.value_seed;
  g.f_row = b.f_row;
  g.ln_row = b.ln_row;
  g.f_col = b.f_col;
  g.ln_col = b.ln_col;
  g.cc = b.s.cc;
  g.min_ = b.s.min_;
  g.lim = std::min(b.s.n_row, b.s.n_col);
  return g;
}

// values / count of their column over the whole matrix (fill_cqmat_stochastic)
void divide_by_column_counts(const Geo &g, long total, const int *colidx, double *vals) {
  DevBuf cc;
  cc.alloc(sizeof(unsigned) * static_cast<std::size_t>(g.nc));
  TBSLA_HIP_CHECK(hipMemset(cc.get(), 0, sizeof(unsigned) * static_cast<std::size_t>(g.nc)));
  hipLaunchKernelGGL(gen_colcount, dim3(grid_of(g.lim)), dim3(kGenThreads), 0, 0, g, cc.as<unsigned>());
  TBSLA_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(gen_stochastic, dim3(grid_of(total)), dim3(kGenThreads), 0, 0, total, colidx,
                     cc.as<unsigned>(), vals);
  TBSLA_HIP_CHECK(hipGetLastError());
}

// The in-block entries of rows [b.f_row, b.f_row + nrows), left on the
// device (rp, ci, va) with the row pointers on the host.
struct DeviceRows {
  std::vector<int> rowptr;
  long total = 0;
  DevBuf rp, ci, va;
};

void generate_rows_device(const host::CqmatBlock &b, long nrows, bool stochastic, DeviceRows &d) {
  d.rowptr.assign(static_cast<std::size_t>(std::max(0L, nrows)) + 1, 0);
  d.total = 0;
  if (nrows <= 0) return;
  if (static_cast<long>(b.n_col) >= (1L << 31) || b.f_row + nrows > b.n_row)
    throw std::invalid_argument("tbsla: cqmat generation outside the matrix");
  const Geo g = geo_of(b);
  // the rows generated all lie below min(n_row, n_col)
  if (b.f_row + nrows > g.lim) throw std::invalid_argument("tbsla: cqmat rows past min(n_row, n_col)");
  DevBuf counts;
  counts.alloc(sizeof(int) * static_cast<std::size_t>(nrows));
  hipLaunchKernelGGL(gen_count, dim3(grid_of(nrows)), dim3(kGenThreads), 0, 0, g, nrows, counts.as<int>());
  TBSLA_HIP_CHECK(hipGetLastError());
  std::vector<int> cnt_h(static_cast<std::size_t>(nrows));
  TBSLA_HIP_CHECK(hipMemcpy(cnt_h.data(), counts.get(), cnt_h.size() * sizeof(int), hipMemcpyDeviceToHost));
  long total = 0;
  for (long t = 0; t < nrows; t++) {
    total += cnt_h[static_cast<std::size_t>(t)];
    if (total >= (1L << 31)) throw std::invalid_argument("tbsla: cqmat block exceeds 2^31 entries");
    d.rowptr[static_cast<std::size_t>(t) + 1] = static_cast<int>(total);
  }
  d.total = total;
  d.rp.upload(d.rowptr.data(), d.rowptr.size() * sizeof(int));
  if (total == 0) return;
  d.ci.alloc(sizeof(int) * static_cast<std::size_t>(total));
  d.va.alloc(sizeof(double) * static_cast<std::size_t>(total));
  const unsigned long long vmix = b.value_seed ? gen::splitmix64(b.value_seed) : 0ull;
  hipLaunchKernelGGL(gen_fill, dim3(grid_of(nrows)), dim3(kGenThreads), 0, 0, g, nrows, d.rp.as<int>(),
                     d.ci.as<int>(), d.va.as<double>(), vmix);
  TBSLA_HIP_CHECK(hipGetLastError());
  if (stochastic) divide_by_column_counts(g, total, d.ci.as<int>(), d.va.as<double>());
}

template <class T>
void copy_back(std::vector<T> &h, const DevBuf &d, long n) {
  h.resize(static_cast<std::size_t>(n));
  if (n > 0) TBSLA_HIP_CHECK(hipMemcpy(h.data(), d.get(), h.size() * sizeof(T), hipMemcpyDeviceToHost));
}

}  // namespace

void generate_cqmat_scoo(const host::CqmatBlock &b, long nrows, bool stochastic, std::vector<int> &rowptr,
                         std::vector<int> &row, std::vector<int> &col, std::vector<double> &values) {
  DeviceRows d;
  generate_rows_device(b, nrows, stochastic, d);
  rowptr = d.rowptr;
  row.clear();
  col.clear();
  values.clear();
  if (d.total == 0) return;
  DevBuf rows;
  rows.alloc(sizeof(int) * static_cast<std::size_t>(d.total));
  hipLaunchKernelGGL(gen_row_ids, dim3(grid_of(nrows)), dim3(kGenThreads), 0, 0, nrows, static_cast<long>(b.f_row),
                     d.rp.as<int>(), rows.as<int>());
  TBSLA_HIP_CHECK(hipGetLastError());
  copy_back(row, rows, d.total);
  copy_back(col, d.ci, d.total);
  copy_back(values, d.va, d.total);
}

void generate_cqmat_ell(const host::CqmatBlock &b, long nrows, long fill_rows, bool stochastic,
                        const std::function<int(const std::vector<int> &)> &width, int &K, std::vector<int> &columns,
                        std::vector<double> &values, bool fill) {
  DeviceRows d;
  generate_rows_device(b, nrows, stochastic, d);
  K = width(d.rowptr);
  columns.clear();
  values.clear();
  if (!fill || fill_rows <= 0 || K <= 0) return;
  const long slots = fill_rows * static_cast<long>(K);
  DevBuf ec, ev;
  ec.alloc(sizeof(int) * static_cast<std::size_t>(slots));
  ev.alloc(sizeof(double) * static_cast<std::size_t>(slots));
  // (a block with no entries has no value buffer: every slot is padding)
  hipLaunchKernelGGL(gen_ell_spread, dim3(grid_of(slots)), dim3(kGenThreads), 0, 0, fill_rows, nrows, K,
                     d.rp.as<int>(), d.ci.as<int>(), d.va.as<double>(), ec.as<int>(), ev.as<double>());
  TBSLA_HIP_CHECK(hipGetLastError());
  copy_back(columns, ec, slots);
  copy_back(values, ev, slots);
}

void generate_cqmat_coo(const host::CqmatBlock &b, long first, long n, bool stochastic, std::vector<int> &row,
                        std::vector<int> &col, std::vector<double> &values) {
  row.clear();
  col.clear();
  values.clear();
  if (n <= 0) return;
  if (static_cast<long>(b.n_col) >= (1L << 31) || n >= (1L << 31))
    throw std::invalid_argument("tbsla: cqmat COO slice too large");
  const Geo g = geo_of(b);
  DevBuf r, c, v;
  r.alloc(sizeof(int) * static_cast<std::size_t>(n));
  c.alloc(sizeof(int) * static_cast<std::size_t>(n));
  v.alloc(sizeof(double) * static_cast<std::size_t>(n));
  const unsigned long long vmix = b.value_seed ? gen::splitmix64(b.value_seed) : 0ull;
  hipLaunchKernelGGL(gen_coo, dim3(grid_of(n)), dim3(kGenThreads), 0, 0, g, first, n, r.as<int>(), c.as<int>(),
                     v.as<double>(), vmix);
  TBSLA_HIP_CHECK(hipGetLastError());
  if (stochastic) divide_by_column_counts(g, n, c.as<int>(), v.as<double>());
  copy_back(row, r, n);
  copy_back(col, c, n);
  copy_back(values, v, n);
}

void generate_cqmat_csr(const host::CqmatBlock &b, long nrows, bool stochastic, GeneratedCsr &out) {
  out.rowptr.assign(static_cast<std::size_t>(nrows) + 1, 0);
  out.colidx.clear();
  out.values.clear();
  if (nrows <= 0) return;
  if (static_cast<long>(b.n_col) >= (1L << 31) || b.f_row + nrows > b.n_row)
    throw std::invalid_argument("tbsla: cqmat generation outside the matrix");
  Geo g{};
  g.nr = static_cast<unsigned long long>(b.n_row);
  g.nc = static_cast<unsigned long long>(b.n_col);
  g.c = static_cast<unsigned long long>(b.c);
  g.q = b.q;
  g.seed = b.seed;
  g.value_seed = b.value_seed;
  g.f_row = b.f_row;
  g.ln_row = b.ln_row;
  g.f_col = b.f_col;
  g.ln_col = b.ln_col;
  g.cc = b.s.cc;
  g.min_ = b.s.min_;
  g.lim = std::min(b.s.n_row, b.s.n_col);
  // the rows generated all lie below min(n_row, n_col)
  if (b.f_row + nrows > g.lim) throw std::invalid_argument("tbsla: cqmat rows past min(n_row, n_col)");
  DevBuf counts;
  counts.alloc(sizeof(int) * static_cast<std::size_t>(nrows));
  hipLaunchKernelGGL(gen_count, dim3(grid_of(nrows)), dim3(kGenThreads), 0, 0, g, nrows, counts.as<int>());
  TBSLA_HIP_CHECK(hipGetLastError());
  std::vector<int> cnt_h(static_cast<std::size_t>(nrows));
  TBSLA_HIP_CHECK(hipMemcpy(cnt_h.data(), counts.get(), cnt_h.size() * sizeof(int), hipMemcpyDeviceToHost));
  long total = 0;
  for (long t = 0; t < nrows; t++) {
    total += cnt_h[static_cast<std::size_t>(t)];
    if (total >= (1L << 31)) throw std::invalid_argument("tbsla: cqmat block exceeds 2^31 entries");
    out.rowptr[static_cast<std::size_t>(t) + 1] = static_cast<int>(total);
  }
  if (total == 0) return;
  DevBuf rp, ci, va;
  rp.upload(out.rowptr.data(), out.rowptr.size() * sizeof(int));
  ci.alloc(sizeof(int) * static_cast<std::size_t>(total));
  va.alloc(sizeof(double) * static_cast<std::size_t>(total));
  const unsigned long long vmix = b.value_seed ? gen::splitmix64(b.value_seed) : 0ull;
  hipLaunchKernelGGL(gen_fill, dim3(grid_of(nrows)), dim3(kGenThreads), 0, 0, g, nrows, rp.as<int>(), ci.as<int>(),
                     va.as<double>(), vmix);
  TBSLA_HIP_CHECK(hipGetLastError());
  if (stochastic) {
    DevBuf cc;
    cc.alloc(sizeof(unsigned) * static_cast<std::size_t>(b.n_col));
    TBSLA_HIP_CHECK(hipMemset(cc.get(), 0, sizeof(unsigned) * static_cast<std::size_t>(b.n_col)));
    hipLaunchKernelGGL(gen_colcount, dim3(grid_of(g.lim)), dim3(kGenThreads), 0, 0, g, cc.as<unsigned>());
    TBSLA_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(gen_stochastic, dim3(grid_of(total)), dim3(kGenThreads), 0, 0, total, ci.as<int>(),
                       cc.as<unsigned>(), va.as<double>());
    TBSLA_HIP_CHECK(hipGetLastError());
  }
  out.colidx.resize(static_cast<std::size_t>(total));
  out.values.resize(static_cast<std::size_t>(total));
  TBSLA_HIP_CHECK(hipMemcpy(out.colidx.data(), ci.get(), out.colidx.size() * sizeof(int), hipMemcpyDeviceToHost));
  TBSLA_HIP_CHECK(hipMemcpy(out.values.data(), va.get(), out.values.size() * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace hip
}  // namespace tbsla
