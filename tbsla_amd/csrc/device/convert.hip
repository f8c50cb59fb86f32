// GPU COO -> CSR / ELL conversion.
//
// The reference sorts a permutation by (row, column) with std::sort
// (MatrixCSR.cpp:29-74, MatrixELL.cpp:12-54; csr.cpp:3-8); the host
// conversions here use a stable order, ties in input order.  On the GPU the
// same order comes from a stable LSD radix sort (hipCUB) of the 64-bit keys
// row << 32 | (column with its sign bit flipped), carrying the entry index;
// the entries are then gathered in that order (CSR) and, for ELL, spread to
// the padded layout on the host.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "../host/gen.hpp"
#include "common.hpp"
#include "convert.hpp"

namespace tbsla {
namespace hip {
namespace {

constexpr int kThreads = 256;

unsigned grid_of(long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

__global__ __launch_bounds__(kThreads) void coo_keys(long n, const int *__restrict__ row, const int *__restrict__ col,
                                                     unsigned long long *__restrict__ key, int *__restrict__ idx,
                                                     int *__restrict__ cnt) {
  const long k = static_cast<long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  const int r = row[k];
  key[k] = (static_cast<unsigned long long>(static_cast<unsigned>(r)) << 32) |
           static_cast<unsigned long long>(static_cast<unsigned>(col[k]) ^ 0x80000000u);
  idx[k] = static_cast<int>(k);
  atomicAdd(cnt + r, 1);
}

// CSR: entry k of the sorted order
__global__ __launch_bounds__(kThreads) void coo_gather(long n, const int *__restrict__ idx, const int *__restrict__ col,
                                                       const double *__restrict__ val, int *__restrict__ ocol,
                                                       double *__restrict__ oval) {
  const long k = static_cast<long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  const int s = idx[k];
  ocol[k] = col[s];
  oval[k] = val[s];
}

// Sorted keys, permutation and row pointers of a COO on the device.
struct Sorted {
  long n = 0;
  DevBuf col, val, key, idx, rowptr;
  std::vector<int> rowptr_h;
};

void sort_coo(int n_row, const std::vector<int> &row, const std::vector<int> &col, const std::vector<double> &val,
              Sorted &s) {
  const long n = static_cast<long>(val.size());
  if (static_cast<long>(row.size()) < n || static_cast<long>(col.size()) < n)
    throw std::invalid_argument("tbsla: COO row/col shorter than values");
  if (n >= (1L << 31)) throw std::invalid_argument("tbsla: GPU conversion limited to 2^31 entries");
  for (long k = 0; k < n; k++)
    if (row[static_cast<std::size_t>(k)] < 0 || row[static_cast<std::size_t>(k)] >= n_row)
      throw std::invalid_argument("tbsla: COO row index outside [0, n_row)");
  s.n = n;
  s.rowptr_h.assign(static_cast<std::size_t>(n_row) + 1, 0);
  if (n == 0) return;
  DevBuf row_d, cnt, key_in, idx_in;
  row_d.upload(row.data(), static_cast<std::size_t>(n) * sizeof(int));
  s.col.upload(col.data(), static_cast<std::size_t>(n) * sizeof(int));
  s.val.upload(val.data(), static_cast<std::size_t>(n) * sizeof(double));
  cnt.alloc(sizeof(int) * static_cast<std::size_t>(n_row));
  TBSLA_HIP_CHECK(hipMemset(cnt.get(), 0, sizeof(int) * static_cast<std::size_t>(n_row)));
  key_in.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(n));
  idx_in.alloc(sizeof(int) * static_cast<std::size_t>(n));
  s.key.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(n));
  s.idx.alloc(sizeof(int) * static_cast<std::size_t>(n));
  hipLaunchKernelGGL(coo_keys, dim3(grid_of(n)), dim3(kThreads), 0, 0, n, row_d.as<int>(), s.col.as<int>(),
                     key_in.as<unsigned long long>(), idx_in.as<int>(), cnt.as<int>());
  TBSLA_HIP_CHECK(hipGetLastError());
  // row bits above the 32 column bits
  int end_bit = 32;
  while (end_bit < 64 && (static_cast<unsigned long long>(n_row - 1) >> (end_bit - 32)) != 0) end_bit++;
  std::size_t tmp_bytes = 0;
  TBSLA_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key_in.as<unsigned long long>(),
                                                     s.key.as<unsigned long long>(), idx_in.as<int>(),
                                                     s.idx.as<int>(), static_cast<int>(n), 0, end_bit));
  DevBuf tmp;
  tmp.alloc(std::max<std::size_t>(tmp_bytes, 256));
  TBSLA_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.get(), tmp_bytes, key_in.as<unsigned long long>(),
                                                     s.key.as<unsigned long long>(), idx_in.as<int>(),
                                                     s.idx.as<int>(), static_cast<int>(n), 0, end_bit));
  std::vector<int> c(static_cast<std::size_t>(n_row));
  TBSLA_HIP_CHECK(hipMemcpy(c.data(), cnt.get(), c.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int r = 0; r < n_row; r++)
    s.rowptr_h[static_cast<std::size_t>(r) + 1] = s.rowptr_h[static_cast<std::size_t>(r)] + c[static_cast<std::size_t>(r)];
  s.rowptr.upload(s.rowptr_h.data(), s.rowptr_h.size() * sizeof(int));
}

}  // namespace

void coo_to_csr_gpu(int n_row, const std::vector<int> &row, const std::vector<int> &col,
                    const std::vector<double> &val, std::vector<int> &rowptr, std::vector<int> &colidx,
                    std::vector<double> &values) {
  Sorted s;
  sort_coo(n_row, row, col, val, s);
  rowptr = s.rowptr_h;
  colidx.assign(static_cast<std::size_t>(s.n), 0);
  values.assign(static_cast<std::size_t>(s.n), 0.0);
  if (s.n == 0) return;
  DevBuf oc, ov;
  oc.alloc(sizeof(int) * static_cast<std::size_t>(s.n));
  ov.alloc(sizeof(double) * static_cast<std::size_t>(s.n));
  hipLaunchKernelGGL(coo_gather, dim3(grid_of(s.n)), dim3(kThreads), 0, 0, s.n, s.idx.as<int>(), s.col.as<int>(),
                     s.val.as<double>(), oc.as<int>(), ov.as<double>());
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipMemcpy(colidx.data(), oc.get(), colidx.size() * sizeof(int), hipMemcpyDeviceToHost));
  TBSLA_HIP_CHECK(hipMemcpy(values.data(), ov.get(), values.size() * sizeof(double), hipMemcpyDeviceToHost));
}

void coo_to_ell_gpu(int n_row, const std::vector<int> &row, const std::vector<int> &col,
                    const std::vector<double> &val, int &K, std::vector<int> &columns, std::vector<double> &values) {
  // sorted on the GPU, brought back in the compact CSR form, then spread to
  // the padded n_row x K layout on the host cores (the padded arrays are K/avg
  // times larger than what crosses the bus)
  std::vector<int> rowptr, colidx;
  std::vector<double> v;
  coo_to_csr_gpu(n_row, row, col, val, rowptr, colidx, v);
  K = 0;
  for (int r = 0; r < n_row; r++)
    K = std::max(K, rowptr[static_cast<std::size_t>(r) + 1] - rowptr[static_cast<std::size_t>(r)]);
  const std::size_t slots = static_cast<std::size_t>(n_row) * static_cast<std::size_t>(K);
  columns.resize(slots);
  values.resize(slots);
  gen::parallel_for(n_row, [&](long b, long e) {
    for (long r = b; r < e; r++) {
      const std::size_t o = static_cast<std::size_t>(r) * static_cast<std::size_t>(K);
      const int a = rowptr[static_cast<std::size_t>(r)], len = rowptr[static_cast<std::size_t>(r) + 1] - a;
      for (int j = 0; j < K; j++) {
        columns[o + static_cast<std::size_t>(j)] = j < len ? colidx[static_cast<std::size_t>(a + j)] : 0;
        values[o + static_cast<std::size_t>(j)] = j < len ? v[static_cast<std::size_t>(a + j)] : 0.0;
      }
    }
  });
}

}  // namespace hip
}  // namespace tbsla
