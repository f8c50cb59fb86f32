// Device building blocks shared by spmv.hip and pagerank.hip (device code;
// include only from .hip translation units).
#pragma once

#include <hip/hip_runtime.h>

#include "mirror.hpp"

namespace tbsla {
namespace hip {
namespace dev {

struct bf16 {
  unsigned short bits;
};

template <typename T>
struct Val;
template <>
struct Val<double> {
  using A = double;
  __device__ static double get(const double *p, long i) { return __builtin_nontemporal_load(p + i); }
};
template <>
struct Val<float> {
  using A = float;
  __device__ static float get(const float *p, long i) { return __builtin_nontemporal_load(p + i); }
};
template <>
struct Val<bf16> {
  using A = float;
  __device__ static float get(const bf16 *p, long i) {
    unsigned short b = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p) + i);
    return __uint_as_float(static_cast<unsigned>(b) << 16);
  }
};

__device__ inline int nt_load(const int *p, long i) { return __builtin_nontemporal_load(p + i); }

// Sum of one value per thread over a block of kTileThreads, fixed order.
template <typename AT>
__device__ inline AT block_tree_sum(AT v, AT *part) {
  part[threadIdx.x] = v;
  __syncthreads();
  for (int s = kTileThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  AT r = part[0];
  __syncthreads();
  return r;
}

template <typename AT>
__device__ inline AT block_tree_max(AT v, AT *part) {
  part[threadIdx.x] = v;
  __syncthreads();
  for (int s = kTileThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + s]);
    __syncthreads();
  }
  AT r = part[0];
  __syncthreads();
  return r;
}

// LDS of one CSR tile workgroup.
template <typename AT>
struct CsrTileLds {
  AT prod[kCsrTileNnz];
  int rp[kCsrTileRows + 1];
};

// Row sums of one CSR tile (rows [tiles[b], tiles[b+1])).  Products are
// rounded and staged in LDS, then each row is summed left to right from 0 by
// one thread, which reproduces the sequential FP64 row sum exactly.  A tile
// holding one row longer than kCsrTileNnz is summed chunk by chunk with a
// fixed-order tree.  epi(row, sum) is called once per row, by the thread that
// owns it (thread 0 for a long row).  Templated on the x element type XT,
// whose values are converted to the accumulation type.
template <typename VT, typename XT, typename Epi>
__device__ inline void csr_tile_rows(const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                     const VT *__restrict__ vals, const XT *__restrict__ x,
                                     const int *__restrict__ tiles, int rp0, int f_col,
                                     CsrTileLds<typename Val<VT>::A> &lds, Epi &epi) {
  using AT = typename Val<VT>::A;
  const int tid = threadIdx.x;
  const int r0 = tiles[blockIdx.x], r1 = tiles[blockIdx.x + 1];
  const int nr = r1 - r0;
  for (int i = tid; i <= nr; i += kTileThreads) lds.rp[i] = rowptr[r0 + i] - rp0;
  __syncthreads();
  const long e0 = lds.rp[0];
  const int ne = lds.rp[nr] - lds.rp[0];
  if (ne <= kCsrTileNnz) {
    int c[kTilePer];
    AT v[kTilePer];
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < ne) {
        c[k] = nt_load(colidx, e0 + e) - f_col;
        v[k] = Val<VT>::get(vals, e0 + e);
      }
    }
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < ne) lds.prod[e] = v[k] * static_cast<AT>(x[c[k]]);
    }
    __syncthreads();
    for (int i = tid; i < nr; i += kTileThreads) {
      AT s = 0;
      const int a = lds.rp[i] - lds.rp[0], b = lds.rp[i + 1] - lds.rp[0];
      for (int e = a; e < b; e++) s += lds.prod[e];
      epi(r0 + i, s);
    }
    return;
  }
  AT total = 0;
  for (int base = 0; base < ne; base += kCsrTileNnz) {
    const int cnt = min(kCsrTileNnz, ne - base);
    AT acc = 0;
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < cnt) {
        const int cc = nt_load(colidx, e0 + base + e) - f_col;
        acc += Val<VT>::get(vals, e0 + base + e) * static_cast<AT>(x[cc]);
      }
    }
    total += block_tree_sum<AT>(acc, lds.prod);
  }
  if (tid == 0) epi(r0, total);
}

// Value-free variant for count-stochastic matrices: every value of column j
// equals colscale[j] and z[j] = colscale[j] * x[j] was formed beforehand, so
// the row sum is a plain sum of gathered z (bit-identical to values * x).
template <typename Epi>
__device__ inline void csr_pattern_tile_rows(const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                             const double *__restrict__ z, const int *__restrict__ tiles, int rp0,
                                             int f_col, CsrTileLds<double> &lds, Epi &epi) {
  const int tid = threadIdx.x;
  const int r0 = tiles[blockIdx.x], r1 = tiles[blockIdx.x + 1];
  const int nr = r1 - r0;
  for (int i = tid; i <= nr; i += kTileThreads) lds.rp[i] = rowptr[r0 + i] - rp0;
  __syncthreads();
  const long e0 = lds.rp[0];
  const int ne = lds.rp[nr] - lds.rp[0];
  if (ne <= kCsrTileNnz) {
    int c[kTilePer];
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < ne) c[k] = nt_load(colidx, e0 + e) - f_col;
    }
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < ne) lds.prod[e] = z[c[k]];
    }
    __syncthreads();
    for (int i = tid; i < nr; i += kTileThreads) {
      double s = 0;
      const int a = lds.rp[i] - lds.rp[0], b = lds.rp[i + 1] - lds.rp[0];
      for (int e = a; e < b; e++) s += lds.prod[e];
      epi(r0 + i, s);
    }
    return;
  }
  double total = 0;
  for (int base = 0; base < ne; base += kCsrTileNnz) {
    const int cnt = min(kCsrTileNnz, ne - base);
    double acc = 0;
#pragma unroll
    for (int k = 0; k < kTilePer; k++) {
      const int e = k * kTileThreads + tid;
      if (e < cnt) acc += z[nt_load(colidx, e0 + base + e) - f_col];
    }
    total += block_tree_sum<double>(acc, lds.prod);
  }
  if (tid == 0) epi(r0, total);
}

}  // namespace dev
}  // namespace hip
}  // namespace tbsla
