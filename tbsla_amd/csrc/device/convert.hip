// GPU COO -> CSR / ELL conversion.
//
// The reference sorts a permutation by (row, column) with std::sort
// (MatrixCSR.cpp:29-74, MatrixELL.cpp:12-54; csr.cpp:3-8); the host
// conversions here use a stable order, ties in input order.  On the GPU the
// same order comes from a stable LSD radix sort of the keys
// row << cb | column (cb = bits of the largest column; with negative
// columns, cb = 32 and the column's sign bit is flipped), carrying the entry
// index: per 8-bit digit, one pass of per-tile digit histograms (LDS), an
// exclusive scan of the (digit, tile) counts, and a stable scatter in which
// each wavefront ranks its keys by matching digits across lanes.  The
// entries are then gathered in that order (CSR) and, for ELL, spread to the
// padded layout on the host.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../host/gen.hpp"
#include "common.hpp"
#include "convert.hpp"
#include "sort.hpp"

namespace tbsla {
namespace hip {
namespace {

constexpr int kThreads = 256;

unsigned grid_of(long n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// key = row << cb | column (flip: the column's sign bit flipped, cb = 32)
__global__ __launch_bounds__(kThreads) void coo_keys(long n, const int *__restrict__ row, const int *__restrict__ col,
                                                     int cb, unsigned flip, unsigned long long *__restrict__ key,
                                                     int *__restrict__ idx, int *__restrict__ cnt) {
  const long k = static_cast<long>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  const int r = row[k];
  key[k] = (static_cast<unsigned long long>(static_cast<unsigned>(r)) << cb) |
           static_cast<unsigned long long>(static_cast<unsigned>(col[k]) ^ flip);
  idx[k] = static_cast<int>(k);
  atomicAdd(cnt + r, 1);
}

// ---- stable LSD radix sort of (64-bit key, int) pairs, 8-bit digits ----
constexpr int kRadix = 256;
constexpr int kSortItems = 16;                       // keys per thread per tile
constexpr int kSortTile = kThreads * kSortItems;     // 4096 keys per tile
constexpr int kWaves = kThreads / 64;
static_assert(kThreads == kRadix, "one thread per digit in the histogram and scatter kernels");

// hist[d * ntiles + t] = keys of tile t whose digit (key >> shift) & 255 is d
__global__ __launch_bounds__(kThreads) void radix_hist(const unsigned long long *__restrict__ key, long n, int shift,
                                                       int *__restrict__ hist, int ntiles) {
  __shared__ int h[kRadix];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long base = static_cast<long>(blockIdx.x) * kSortTile;
  for (int j = 0; j < kSortItems; j++) {
    const long k = base + j * kThreads + threadIdx.x;
    if (k < n) atomicAdd(&h[static_cast<int>(key[k] >> shift) & (kRadix - 1)], 1);
  }
  __syncthreads();
  hist[static_cast<long>(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of a[0, m) in chunks of kSortTile: each chunk scanned in
// place, its total in sums[chunk].
__global__ __launch_bounds__(kThreads) void scan_chunks(int *__restrict__ a, long m, int *__restrict__ sums) {
  __shared__ int part[kThreads];
  const long base = static_cast<long>(blockIdx.x) * kSortTile + static_cast<long>(threadIdx.x) * kSortItems;
  int v[kSortItems], t = 0;
  for (int j = 0; j < kSortItems; j++) {
    v[j] = base + j < m ? a[base + j] : 0;
    t += v[j];
  }
  part[threadIdx.x] = t;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {  // inclusive Hillis-Steele scan of the thread totals
    const int x = static_cast<int>(threadIdx.x) >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += x;
    __syncthreads();
  }
  int run = part[threadIdx.x] - t;
  for (int j = 0; j < kSortItems; j++) {
    if (base + j < m) a[base + j] = run;
    run += v[j];
  }
  if (threadIdx.x == kThreads - 1) sums[blockIdx.x] = part[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void add_chunk_offsets(int *__restrict__ a, long m,
                                                              const int *__restrict__ offs) {
  const long base = static_cast<long>(blockIdx.x) * kSortTile;
  const int o = offs[blockIdx.x];
  for (int j = 0; j < kSortItems; j++) {
    const long k = base + j * kThreads + threadIdx.x;
    if (k < m) a[k] += o;
  }
}

}  // namespace

// In-place exclusive scan of a[0, m) on the device (chunk scans, then the
// chunk totals scanned the same way, recursively).
void exclusive_scan(int *a, long m, hipStream_t s) {
  const long chunks = (m + kSortTile - 1) / kSortTile;
  DevBuf sums;
  sums.alloc(sizeof(int) * static_cast<std::size_t>(chunks));
  hipLaunchKernelGGL(scan_chunks, dim3(static_cast<unsigned>(chunks)), dim3(kThreads), 0, s, a, m, sums.as<int>());
  if (chunks > 1) {
    exclusive_scan(sums.as<int>(), chunks, s);
    hipLaunchKernelGGL(add_chunk_offsets, dim3(static_cast<unsigned>(chunks)), dim3(kThreads), 0, s, a, m,
                       sums.as<int>());
  }
  TBSLA_HIP_CHECK(hipGetLastError());
}

namespace {

// Lanes of this wavefront whose 8-bit digit (and validity) equal this lane's.
__device__ inline unsigned long long digit_peers(unsigned d, bool valid) {
  unsigned long long m = __ballot(valid);
  if (!valid) m = ~m;
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const bool bit = (d >> b) & 1u;
    const unsigned long long bal = __ballot(bit);
    m &= bit ? bal : ~bal;
  }
  return m;
}

// Stable scatter of one tile: every key goes to offs[digit] (the scanned
// position of (digit, tile)) plus its rank among the tile's keys of that
// digit in input order.  The tile is taken in rounds of kThreads keys; in a
// round, wavefront w holds keys 64w .. 64w+63 and ranks each lane among the
// lanes with the same digit; the per-wavefront digit counts then give each
// key the keys of earlier wavefronts and rounds.
__global__ __launch_bounds__(kThreads) void radix_scatter(const unsigned long long *__restrict__ kin,
                                                          const int *__restrict__ vin, long n, int shift,
                                                          const int *__restrict__ scanned, int ntiles,
                                                          unsigned long long *__restrict__ kout,
                                                          int *__restrict__ vout) {
  __shared__ int offs[kRadix];
  __shared__ int wcount[kWaves][kRadix];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  offs[tid] = scanned[static_cast<long>(tid) * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < kWaves; w++) wcount[w][tid] = 0;
  __syncthreads();
  const unsigned long long below = (1ull << lane) - 1ull;
  const long base = static_cast<long>(blockIdx.x) * kSortTile;
  for (int j = 0; j < kSortItems; j++) {
    const long k = base + j * kThreads + tid;
    const bool valid = k < n;
    const unsigned long long key = valid ? kin[k] : 0ull;
    const int val = valid ? vin[k] : 0;
    const unsigned d = static_cast<unsigned>(key >> shift) & (kRadix - 1);
    const unsigned long long peers = digit_peers(d, valid);
    const int rank = __popcll(peers & below);
    if (valid && rank == 0) wcount[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int pos = offs[d] + rank;
      for (int w = 0; w < wave; w++) pos += wcount[w][d];
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    // the round's counts move into offs, and the per-wavefront counts clear
    int add = 0;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
      add += wcount[w][tid];
      wcount[w][tid] = 0;
    }
    offs[tid] += add;
    __syncthreads();
  }
}

}  // namespace

// Sorts (key, val)[0, n) by key bits [0, bits), stable; the result is in
// (key, val) again (the ping-pong buffers k2, v2 hold n entries each).
void radix_sort_pairs(unsigned long long *key, int *val, unsigned long long *k2, int *v2, long n, int bits,
                      hipStream_t s) {
  const int ntiles = static_cast<int>((n + kSortTile - 1) / kSortTile);
  DevBuf hist;
  hist.alloc(sizeof(int) * static_cast<std::size_t>(kRadix) * static_cast<std::size_t>(ntiles));
  unsigned long long *ka = key, *kb = k2;
  int *va = val, *vb = v2;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(radix_hist, dim3(ntiles), dim3(kThreads), 0, s, ka, n, shift, hist.as<int>(), ntiles);
    exclusive_scan(hist.as<int>(), static_cast<long>(kRadix) * ntiles, s);
    hipLaunchKernelGGL(radix_scatter, dim3(ntiles), dim3(kThreads), 0, s, ka, va, n, shift, hist.as<int>(), ntiles,
                       kb, vb);
    TBSLA_HIP_CHECK(hipGetLastError());
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != key) {
    TBSLA_HIP_CHECK(hipMemcpyAsync(key, ka, sizeof(unsigned long long) * static_cast<std::size_t>(n),
                                   hipMemcpyDeviceToDevice, s));
    TBSLA_HIP_CHECK(hipMemcpyAsync(val, va, sizeof(int) * static_cast<std::size_t>(n), hipMemcpyDeviceToDevice, s));
  }
}

namespace {

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
  DevBuf row_d, cnt, key2, idx2;
  row_d.upload(row.data(), static_cast<std::size_t>(n) * sizeof(int));
  s.col.upload(col.data(), static_cast<std::size_t>(n) * sizeof(int));
  s.val.upload(val.data(), static_cast<std::size_t>(n) * sizeof(double));
  cnt.alloc(sizeof(int) * static_cast<std::size_t>(n_row));
  TBSLA_HIP_CHECK(hipMemset(cnt.get(), 0, sizeof(int) * static_cast<std::size_t>(n_row)));
  s.key.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(n));
  s.idx.alloc(sizeof(int) * static_cast<std::size_t>(n));
  key2.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(n));
  idx2.alloc(sizeof(int) * static_cast<std::size_t>(n));
  // column bits: those of the largest column when none is negative, else
  // all 32 with the sign bit flipped (so that unsigned order is signed order)
  const auto mm = std::minmax_element(col.begin(), col.begin() + n);
  int cb = 0;
  unsigned flip = 0;
  if (*mm.first < 0) {
    cb = 32;
    flip = 0x80000000u;
  } else {
    while (cb < 31 && (static_cast<unsigned>(*mm.second) >> cb) != 0) cb++;
  }
  int rb = 0;
  while (rb < 31 && (static_cast<unsigned>(n_row - 1) >> rb) != 0) rb++;
  hipLaunchKernelGGL(coo_keys, dim3(grid_of(n)), dim3(kThreads), 0, 0, n, row_d.as<int>(), s.col.as<int>(), cb, flip,
                     s.key.as<unsigned long long>(), s.idx.as<int>(), cnt.as<int>());
  TBSLA_HIP_CHECK(hipGetLastError());
  radix_sort_pairs(s.key.as<unsigned long long>(), s.idx.as<int>(), key2.as<unsigned long long>(), idx2.as<int>(), n,
                   cb + rb, nullptr);
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
