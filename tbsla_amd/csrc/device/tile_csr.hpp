// Device building blocks shared by spmv.hip and pagerank.hip (device code;
// include only from .hip translation units).
//
// CSR products use fixed entry tiles: tile t holds entries [t*T, (t+1)*T)
// (T = kCsrTileNnz, 3072, or the caller's TN) and owns the rows whose first
// entry lies in it
// (tile_row[t] .. tile_row[t+1]).  A tile's entries are read with 16-byte
// loads at an aligned offset that needs no descriptor load, the products
// v * x[c] are rounded and staged in LDS, and each owned row is summed by one
// thread from 0, left to right, finishing a row that runs past the tile from
// global memory in the same order.  That reproduces the sequential FP64 row
// sum of MatrixCSR.cpp:133-139 exactly.  Rows longer than T / 2 entries are
// left to csr_long_row (one workgroup per row, fixed-order tree).
#pragma once

#include <hip/hip_runtime.h>

#include "mirror.hpp"

namespace tbsla {
namespace hip {
namespace dev {

struct bf16 {
  unsigned short bits;
};

typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef unsigned short v4us __attribute__((ext_vector_type(4)));

template <typename T>
struct Val;
template <>
struct Val<double> {
  using A = double;
  __device__ static double get(const double *p, long i) { return p[i]; }
  // four consecutive values starting at a multiple of 4
  __device__ static void get4(const double *p, long i, double (&v)[4]) {
    const v2d a = *reinterpret_cast<const v2d *>(p + i);
    const v2d b = *reinterpret_cast<const v2d *>(p + i + 2);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
  }
};
template <>
struct Val<float> {
  using A = float;
  __device__ static float get(const float *p, long i) { return p[i]; }
  __device__ static void get4(const float *p, long i, float (&v)[4]) {
    const v4f a = *reinterpret_cast<const v4f *>(p + i);
    v[0] = a.x;
    v[1] = a.y;
    v[2] = a.z;
    v[3] = a.w;
  }
};
template <>
struct Val<bf16> {
  using A = float;
  __device__ static float cvt(unsigned short b) { return __uint_as_float(static_cast<unsigned>(b) << 16); }
  __device__ static float get(const bf16 *p, long i) { return cvt(reinterpret_cast<const unsigned short *>(p)[i]); }
  __device__ static void get4(const bf16 *p, long i, float (&v)[4]) {
    const v4us a = *reinterpret_cast<const v4us *>(reinterpret_cast<const unsigned short *>(p) + i);
    v[0] = cvt(a.x);
    v[1] = cvt(a.y);
    v[2] = cvt(a.z);
    v[3] = cvt(a.w);
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

// Device view of a CSR mirror (entry j of the matrix is colidx[j], vals[j],
// with j = rowptr[i] - rp0 + k for the k-th entry of row i).  colidx is the
// mirror's padded buffer; vals may be a caller's array of exactly nnz
// values, so no value is read at index nnz or beyond.
template <typename VT>
struct CsrView {
  const int *rowptr;
  const int *colidx;
  const VT *vals;  // nullptr for the value-free view
  const int *tile_row;
  const int *long_rows;
  long nnz;
  int rp0, f_col;
};

template <typename VT>
inline CsrView<VT> csr_view(const DeviceMatrix &d) {
  return CsrView<VT>{d.rowptr.as<int>(), d.colidx.as<int>(), d.vals.as<VT>(), d.tiles.as<int>(),
                     d.long_rows.as<int>(), d.entries, d.rp0, d.f_col};
}

// A row's sum reads its products from LDS kSumRun at a time, all reads issued
// before the adds; the slack slots after the tile keep those reads in bounds
// (their contents are never added).
constexpr int kSumRun = 16;

// TN: entries per tile (kCsrTileNnz unless a caller builds its own tile
// table with another size); rows longer than TN / 2 go to csr_long_row.
template <typename AT, int TN = kCsrTileNnz>
struct CsrTileLds {
  AT prod[TN + kSumRun];
};

// Value source: stored values (VT) or value-free (every product is z[c]).
// load4 reads four consecutive values (issued before the gathers that need
// the column indices), apply4 gathers x and forms the rounded products.
template <typename VT, typename XT>
struct Prod {
  using AT = typename Val<VT>::A;
  __device__ static void load4(const CsrView<VT> &m, long e, AT (&v)[4]) { Val<VT>::get4(m.vals, e, v); }
  __device__ static void apply4(const AT (&v)[4], const XT *__restrict__ x, const int (&c)[4], AT (&p)[4]) {
#pragma unroll
    for (int q = 0; q < 4; q++) p[q] = v[q] * static_cast<AT>(x[c[q]]);
  }
  __device__ static AT one(const CsrView<VT> &m, const XT *__restrict__ x, long e) {
    const AT v = Val<VT>::get(m.vals, e);
    const AT xv = static_cast<AT>(x[m.colidx[e] - m.f_col]);
    return v * xv;
  }
};
struct Pattern {};
template <>
struct Val<Pattern> {
  using A = double;
};
template <typename XT>
struct Prod<Pattern, XT> {
  using AT = double;
  __device__ static void load4(const CsrView<Pattern> &, long, double (&)[4]) {}
  __device__ static void apply4(const double (&)[4], const XT *__restrict__ z, const int (&c)[4], double (&p)[4]) {
#pragma unroll
    for (int q = 0; q < 4; q++) p[q] = static_cast<double>(z[c[q]]);
  }
  __device__ static double one(const CsrView<Pattern> &m, const XT *__restrict__ z, long e) {
    return static_cast<double>(z[m.colidx[e] - m.f_col]);
  }
};

// A row sum starts from epi.init<AT>(row) when the epilogue has one (an
// accumulating product, y[row] += ...), else from 0.
template <typename AT, typename E>
__device__ inline auto epi_init(const E &e, int r, int) -> decltype(e.template init<AT>(r)) {
  return e.template init<AT>(r);
}
template <typename AT, typename E>
__device__ inline AT epi_init(const E &, int, long) {
  return AT(0);
}

// The entries [e0, e0 + ne) of one tile held by the calling thread: groups
// of four at 4 * (k * kTileThreads + threadIdx.x).  load() issues the index and
// value loads; products() gathers x and writes the rounded products to
// prod[0 .. ne).  FULL (ne == TN, every tile but the matrix's last)
// reads with no per-entry guard; otherwise entries at or past ne are not read
// (a caller's value array holds exactly nnz values).
template <bool FULL, typename VT, typename XT, int TN>
struct TileEntries {
  using AT = typename Val<VT>::A;
  static constexpr int G = TN / 4 / kTileThreads;  // 4-entry groups per thread
  int c[G][4];
  AT v[G][4];

  __device__ void load(const CsrView<VT> &m, long e0, int ne) {
#pragma unroll
    for (int k = 0; k < G; k++) {
      const int g = k * kTileThreads + threadIdx.x;
      if (FULL || 4 * g < ne) {
        const v4i ci = *reinterpret_cast<const v4i *>(m.colidx + e0 + 4 * g);
        c[k][0] = ci.x - m.f_col;
        c[k][1] = ci.y - m.f_col;
        c[k][2] = ci.z - m.f_col;
        c[k][3] = ci.w - m.f_col;
        if (!FULL)  // entries past the matrix end read the zeroed padding: column 0
          for (int q = 0; q < 4; q++)
            if (4 * g + q >= ne) c[k][q] = 0;
      }
      if (FULL || 4 * g + 4 <= ne) Prod<VT, XT>::load4(m, e0 + 4 * g, v[k]);
    }
  }

  // gathers and rounded products into p
  __device__ void products(const CsrView<VT> &m, const XT *__restrict__ x, long e0, int ne, AT (&p)[G][4]) const {
#pragma clang fp contract(off)
#pragma unroll
    for (int k = 0; k < G; k++) {
      const int g = k * kTileThreads + threadIdx.x;
      if (FULL || 4 * g + 4 <= ne)
        Prod<VT, XT>::apply4(v[k], x, c[k], p[k]);
      else if (4 * g < ne)  // the matrix's last entries: no value read past nnz
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (4 * g + q < ne) p[k][q] = Prod<VT, XT>::one(m, x, e0 + 4 * g + q);
    }
  }

  static __device__ void store(const AT (&p)[G][4], AT *prod, int ne) {
#pragma unroll
    for (int k = 0; k < G; k++) {
      const int g = k * kTileThreads + threadIdx.x;
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (FULL || 4 * g + q < ne) prod[4 * g + q] = p[k][q];
    }
  }
};

// The bounds of the thread's first owned row, then the index/value loads,
// gathers and products into prod.  (Issuing the row table first measured
// faster than deferring it behind the tile loads.)
template <bool FULL, typename VT, typename XT, int TN>
__device__ inline void csr_tile_stage(const CsrView<VT> &m, const XT *__restrict__ x, typename Val<VT>::A *prod,
                                      long e0, int ne, int rlo, int rhi, int &a, int &b) {
  // loaded with no branch so they stay in flight under the tile loads
  // (indices clamped to rhi <= rows: in bounds, and exact for every thread
  // that owns a row)
  a = m.rowptr[min(rlo + static_cast<int>(threadIdx.x), rhi)] - m.rp0;
  b = m.rowptr[min(rlo + static_cast<int>(threadIdx.x) + 1, rhi)] - m.rp0;
  TileEntries<FULL, VT, XT, TN> t;
  t.load(m, e0, ne);
  typename Val<VT>::A p[TileEntries<FULL, VT, XT, TN>::G][4];
  t.products(m, x, e0, ne, p);
  TileEntries<FULL, VT, XT, TN>::store(p, prod, ne);
}

// Rows of entry tile `tile`; epi(row, sum) once per owned row by the thread
// that sums it, rows longer than LT left to csr_long_row.  The caller
// synchronises the block before reusing `lds`.
template <typename VT, typename XT, typename Epi, int TN = kCsrTileNnz, int LT = TN / 2>
__device__ inline void csr_tile_rows_at(const CsrView<VT> &m, const XT *__restrict__ x,
                                        CsrTileLds<typename Val<VT>::A, TN> &lds, Epi &epi, int tile) {
#pragma clang fp contract(off)
  using AT = typename Val<VT>::A;
  const int tid = threadIdx.x;
  const long e0 = static_cast<long>(tile) * TN;
  const int ne = static_cast<int>(min(static_cast<long>(TN), m.nnz - e0));
  const int rlo = m.tile_row[tile], rhi = m.tile_row[tile + 1];
  int a, b;
  if (ne == TN)
    csr_tile_stage<true, VT, XT, TN>(m, x, lds.prod, e0, ne, rlo, rhi, a, b);
  else
    csr_tile_stage<false, VT, XT, TN>(m, x, lds.prod, e0, ne, rlo, rhi, a, b);
  __syncthreads();
  const long e1 = e0 + ne;
  for (int r = rlo + tid; r < rhi; r += kTileThreads) {
    if (r != rlo + tid) {
      a = m.rowptr[r] - m.rp0;
      b = m.rowptr[r + 1] - m.rp0;
    }
    if (b - a > LT) continue;  // summed by csr_long_row
    AT s = epi_init<AT>(epi, r, 0);
    // a row owned by the tile starts in it: 0 <= a - e0 < TN, so the reads
    // below stay inside prod[TN + kSumRun]
    const int lend = static_cast<int>(min(static_cast<long>(b), e1) - e0);
    for (int e = a - static_cast<int>(e0); e < lend; e += kSumRun) {
      AT v[kSumRun];
#pragma unroll
      for (int u = 0; u < kSumRun; u++) v[u] = lds.prod[e + u];
#pragma unroll
      for (int u = 0; u < kSumRun; u++) s = e + u < lend ? s + v[u] : s;
    }
    // the rest of a row that runs past the tile, same order, 8 loads in flight
    for (long e = e1; e < b; e += 8) {
      AT q8[8];
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (e + u < b) q8[u] = Prod<VT, XT>::one(m, x, e + u);
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (e + u < b) s += q8[u];
    }
    epi(r, s);
  }
}

// Rows of entry tile blockIdx.x.
template <typename VT, typename XT, typename Epi>
__device__ inline void csr_tile_rows(const CsrView<VT> &m, const XT *__restrict__ x,
                                     CsrTileLds<typename Val<VT>::A> &lds, Epi &epi) {
  csr_tile_rows_at<VT, XT, Epi>(m, x, lds, epi, blockIdx.x);
}

// One long row (m.long_rows[idx]) per workgroup: chunk products (LC
// entries per chunk), a fixed-order tree per chunk, chunk totals added in
// order.  epi by thread 0.
template <typename VT, typename XT, typename Epi, int TN = kCsrTileNnz, int LC = TN>
__device__ inline void csr_long_row_at(const CsrView<VT> &m, const XT *__restrict__ x,
                                       CsrTileLds<typename Val<VT>::A, TN> &lds, Epi &epi, int idx) {
  using AT = typename Val<VT>::A;
  const int r = m.long_rows[idx];
  const long a = m.rowptr[r] - m.rp0, b = m.rowptr[r + 1] - m.rp0;
  AT total = 0;
  for (long base = a; base < b; base += LC) {
    const long cnt = min(static_cast<long>(LC), b - base);
    AT acc = 0;
    for (long e = threadIdx.x; e < cnt; e += kTileThreads) acc += Prod<VT, XT>::one(m, x, base + e);
    total += block_tree_sum<AT>(acc, lds.prod);
  }
  if (threadIdx.x == 0) epi(r, epi_init<AT>(epi, r, 0) + total);
}

template <typename VT, typename XT, typename Epi>
__device__ inline void csr_long_row(const CsrView<VT> &m, const XT *__restrict__ x,
                                    CsrTileLds<typename Val<VT>::A> &lds, Epi &epi) {
  csr_long_row_at<VT, XT, Epi>(m, x, lds, epi, blockIdx.x);
}

}  // namespace dev
}  // namespace hip
}  // namespace tbsla
