// SpMV kernels for gfx950 (MI355X, CDNA4, 64-wide wavefronts).
//
// csr_tile_kernel   one workgroup per fixed 3072-entry tile.  Entries are
//                   read with 16-byte loads (12 entries per thread),
//                   multiplied by the gathered x, and staged in LDS; each row
//                   that starts in the tile is then summed by one thread in
//                   entry order.  With FP64 this reproduces the reference loop
//                   MatrixCSR.cpp:133-139 bit for bit (product rounded, then a
//                   left-to-right sum starting from 0).
// csr_long_kernel   rows longer than 1536 entries, one workgroup each.
// ell_kernel        one row per thread over the sliced column-major slots (one
//                   contiguous block per workgroup), eight slots per step;
//                   same order as MatrixELL.cpp:105-113 (contraction off, so
//                   bit-identical in FP64).
// coo_tile_kernel   COO/SCOO: the CSR entry-tile kernel on the row segments of
//                   the row-sorted view, each segment's sum starting from
//                   y[row], so each row is accumulated in entry order
//                   (MatrixCOO.cpp:113-115 / MatrixSCOO.cpp:120-122 with one
//                   thread).
// dense_kernel      four rows per wavefront sharing each x load, butterfly reduction
//                   (MatrixDENSE.cpp:67-72, different summation order).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "tile_csr.hpp"

namespace tbsla {
namespace hip {
namespace {

using dev::bf16;
using dev::block_tree_sum;
using dev::nt_load;
using dev::Val;

template <typename YT>
struct StoreRow {
  YT *y;
  template <typename AT>
  __device__ void operator()(int row, AT s) const {
    y[row] = static_cast<YT>(s);
  }
};

template <typename VT, typename XT, typename YT>
__global__ __launch_bounds__(kTileThreads) void csr_tile_kernel(dev::CsrView<VT> m, const XT *__restrict__ x,
                                                                YT *__restrict__ y) {
  __shared__ dev::CsrTileLds<typename Val<VT>::A> lds;
  StoreRow<YT> epi{y};
  dev::csr_tile_rows<VT, XT>(m, x, lds, epi);
}

template <typename VT, typename XT, typename YT>
__global__ __launch_bounds__(kTileThreads) void csr_long_kernel(dev::CsrView<VT> m, const XT *__restrict__ x,
                                                                YT *__restrict__ y) {
  __shared__ dev::CsrTileLds<typename Val<VT>::A> lds;
  StoreRow<YT> epi{y};
  dev::csr_long_row<VT, XT>(m, x, lds, epi);
}

// ELL: one row per thread, one workgroup per slice of kEllSlice rows, whose
// K slots lie in one contiguous block (slot j of the slice's row t at
// j * kEllSlice + t), so a workgroup streams one block of HBM and every slot
// is one coalesced load per wavefront.  Slots go in groups of kEllGroup; the
// next group's loads are issued before this group's gathers (two register
// sets, alternating), so each group waits on one round trip.  Each row still
// sums its slots in order 0..K-1.
constexpr int kEllGroup = 8;

__device__ inline double ell_value(const double *p, long i) { return __builtin_nontemporal_load(p + i); }
__device__ inline float ell_value(const float *p, long i) { return __builtin_nontemporal_load(p + i); }
__device__ inline float ell_value(const bf16 *p, long i) {
  return Val<bf16>::cvt(__builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(p) + i));
}

template <typename VT>
struct EllGroup {
  using AT = typename Val<VT>::A;
  int c[kEllGroup];
  AT v[kEllGroup];

  // slots o, o + kEllSlice, ... (N of them)
  template <int N>
  __device__ void load(const int *__restrict__ cols, const VT *__restrict__ vals, long o) {
#pragma unroll
    for (int u = 0; u < N; u++) {
      c[u] = nt_load(cols, o + u * kEllSlice);
      v[u] = ell_value(vals, o + u * kEllSlice);
    }
  }
  template <int N, typename XT>
  __device__ void accumulate(const XT *__restrict__ x, AT &acc) const {
#pragma clang fp contract(off)
    AT xv[N];
#pragma unroll
    for (int u = 0; u < N; u++) xv[u] = static_cast<AT>(x[c[u]]);
#pragma unroll
    for (int u = 0; u < N; u++) acc += v[u] * xv[u];
  }
};

template <typename VT, typename XT, typename YT>
__global__ __launch_bounds__(kEllSlice) void ell_kernel(const int *__restrict__ cols, const VT *__restrict__ vals,
                                                        const XT *__restrict__ x, YT *__restrict__ y, int rows,
                                                        int K) {
  using AT = typename Val<VT>::A;
  const long i = static_cast<long>(blockIdx.x) * kEllSlice + threadIdx.x;
  if (i >= rows) return;
  const long base = static_cast<long>(blockIdx.x) * K * kEllSlice + threadIdx.x;
  const int KG = K - K % kEllGroup;  // slots in whole groups
  AT acc = 0;
  EllGroup<VT> A, B;
  int j = 0;
  if (KG > 0) A.template load<kEllGroup>(cols, vals, base);
  while (j < KG) {
    if (j + kEllGroup < KG) B.template load<kEllGroup>(cols, vals, base + static_cast<long>(j + kEllGroup) * kEllSlice);
    A.template accumulate<kEllGroup>(x, acc);
    j += kEllGroup;
    if (j >= KG) break;
    if (j + kEllGroup < KG) A.template load<kEllGroup>(cols, vals, base + static_cast<long>(j + kEllGroup) * kEllSlice);
    B.template accumulate<kEllGroup>(x, acc);
    j += kEllGroup;
  }
  for (; j < K; j++) {  // the last K % kEllGroup slots
    A.template load<1>(cols, vals, base + static_cast<long>(j) * kEllSlice);
    A.template accumulate<1>(x, acc);
  }
  y[i] = static_cast<YT>(acc);
}

// COO/SCOO segment epilogue: the segment's sum started from y[row], so the
// row is accumulated y[row] += v * x in entry order (MatrixCOO.cpp:113-115).
template <typename YT>
struct AccRow {
  YT *y;
  const int *segrow;
  long y_off;
  template <typename AT>
  __device__ AT init(int s) const {
    return static_cast<AT>(y[segrow[s] + y_off]);
  }
  template <typename AT>
  __device__ void operator()(int s, AT v) const {
    y[segrow[s] + y_off] = static_cast<YT>(v);
  }
};

template <typename VT, typename XT, typename YT, bool LONG>
__global__ __launch_bounds__(kTileThreads) void coo_tile_kernel(dev::CsrView<VT> m, const XT *__restrict__ x,
                                                                YT *__restrict__ y, const int *__restrict__ segrow,
                                                                long y_off) {
  __shared__ dev::CsrTileLds<typename Val<VT>::A> lds;
  AccRow<YT> epi{y, segrow, y_off};
  if (LONG)
    dev::csr_long_row<VT, XT>(m, x, lds, epi);
  else
    dev::csr_tile_rows<VT, XT>(m, x, lds, epi);
}

// kDenseRows consecutive rows per wavefront, so each x load serves all of
// them; per row and lane the columns are summed in the same order as with one
// row per wavefront (j = 2*lane + 128*k), then a butterfly.
constexpr int kDenseRows = 4;

template <typename XT, typename YT>
__global__ __launch_bounds__(256) void dense_kernel(const double *__restrict__ A, const XT *__restrict__ x,
                                                    YT *__restrict__ y, int rows, int cols, long ld) {
  constexpr int R = kDenseRows;
  const int w = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const long r0 = static_cast<long>(w) * R;
  if (r0 >= rows) return;
  const int nr = static_cast<int>(min(static_cast<long>(R), rows - r0));  // rows of this wavefront
  double acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = 0;
  if ((ld & 1) == 0) {
    // 16-byte non-temporal loads, two columns per lane, two steps of R rows in flight
    int j = 2 * lane;
    for (; j + 128 + 1 < cols; j += 2 * 128) {
      dev::v2d av[2][R];
      double xv[2][2];
#pragma unroll
      for (int u = 0; u < 2; u++) {
        xv[u][0] = static_cast<double>(x[j + u * 128]);
        xv[u][1] = static_cast<double>(x[j + u * 128 + 1]);
#pragma unroll
        for (int r = 0; r < R; r++)
          av[u][r] = r < nr ? __builtin_nontemporal_load(
                                  reinterpret_cast<const dev::v2d *>(A + (r0 + r) * ld + j + u * 128))
                            : dev::v2d{0, 0};
      }
#pragma unroll
      for (int u = 0; u < 2; u++)
#pragma unroll
        for (int r = 0; r < R; r++) acc[r] += av[u][r].x * xv[u][0] + av[u][r].y * xv[u][1];
    }
    for (; j < cols; j += 128)
#pragma unroll
      for (int r = 0; r < R; r++)
        if (r < nr) {
          const double *a = A + (r0 + r) * ld;
          acc[r] += __builtin_nontemporal_load(a + j) * static_cast<double>(x[j]);
          if (j + 1 < cols) acc[r] += __builtin_nontemporal_load(a + j + 1) * static_cast<double>(x[j + 1]);
        }
  } else {
#pragma unroll
    for (int r = 0; r < R; r++)
      if (r < nr) {
        const double *a = A + (r0 + r) * ld;
        for (int j = lane; j < cols; j += 64) acc[r] += __builtin_nontemporal_load(a + j) * static_cast<double>(x[j]);
      }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[r] += __shfl_xor(acc[r], off, 64);
  }
  if (lane == 0)
#pragma unroll
    for (int r = 0; r < R; r++)
      if (r < nr) y[r0 + r] = static_cast<YT>(static_cast<double>(y[r0 + r]) + acc[r]);
}

template <typename T>
__global__ void fill_kernel(T *y, long n, T v) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) y[i] = v;
}

inline unsigned blocks_for(long n, int per_block) { return static_cast<unsigned>((n + per_block - 1) / per_block); }

template <typename VT, typename XT, typename YT>
void run_typed(const DeviceMatrix &d, const XT *x, YT *y, int vect_incr, hipStream_t s) {
  switch (d.format) {
    case cpp::FORMAT_CSR: {
      const dev::CsrView<VT> m = dev::csr_view<VT>(d);
      if (d.ntiles > 0)
        hipLaunchKernelGGL((csr_tile_kernel<VT, XT, YT>), dim3(d.ntiles), dim3(kTileThreads), 0, s, m, x, y);
      if (d.nlong > 0)
        hipLaunchKernelGGL((csr_long_kernel<VT, XT, YT>), dim3(d.nlong), dim3(kTileThreads), 0, s, m, x, y);
      break;
    }
    case cpp::FORMAT_ELL:
      if (d.rows > 0)
        hipLaunchKernelGGL((ell_kernel<VT, XT, YT>), dim3(blocks_for(d.rows, kEllSlice)), dim3(kEllSlice), 0, s,
                           d.ecols.as<int>(), d.evals.as<VT>(), x, y, d.rows, d.K);
      break;
    case cpp::FORMAT_COO:
    case cpp::FORMAT_SCOO: {
      // columns: global (COO) or minus f_col (SCOO, the view's f_col);
      // rows: + vect_incr (- f_row for SCOO)
      const dev::CsrView<VT> m = dev::csr_view<VT>(d);
      const long y_off = vect_incr - (d.format == cpp::FORMAT_SCOO ? d.f_row : 0);
      if (d.ntiles > 0)
        hipLaunchKernelGGL((coo_tile_kernel<VT, XT, YT, false>), dim3(d.ntiles), dim3(kTileThreads), 0, s, m, x, y,
                           d.segrow.as<int>(), y_off);
      if (d.nlong > 0)
        hipLaunchKernelGGL((coo_tile_kernel<VT, XT, YT, true>), dim3(d.nlong), dim3(kTileThreads), 0, s, m, x, y,
                           d.segrow.as<int>(), y_off);
      break;
    }
    case cpp::FORMAT_DENSE:
      if (d.rows > 0)
        hipLaunchKernelGGL((dense_kernel<XT, YT>), dim3(blocks_for((static_cast<long>(d.rows) + kDenseRows - 1) / kDenseRows * 64, 256)),
                           dim3(256), 0, s, d.dvals.as<double>(), x, y, d.rows, d.dcols, static_cast<long>(d.dcols));
      break;
  }
  TBSLA_HIP_CHECK(hipGetLastError());
}

}  // namespace

void launch_spmv(const DeviceMatrix &d, const void *x, Dtype xdt, void *y, Dtype ydt, int vect_incr,
                 hipStream_t s) {
  if (d.noop) return;
  if (d.vdt == Dtype::F64 && xdt == Dtype::F64 && ydt == Dtype::F64) {
    run_typed<double, double, double>(d, static_cast<const double *>(x), static_cast<double *>(y), vect_incr, s);
  } else if (d.vdt == Dtype::F32 && xdt == Dtype::F32 && ydt == Dtype::F32 && d.format != cpp::FORMAT_DENSE) {
    run_typed<float, float, float>(d, static_cast<const float *>(x), static_cast<float *>(y), vect_incr, s);
  } else if (d.vdt == Dtype::BF16 && xdt == Dtype::F32 && ydt == Dtype::F32 && d.format != cpp::FORMAT_DENSE) {
    run_typed<bf16, float, float>(d, static_cast<const float *>(x), static_cast<float *>(y), vect_incr, s);
  } else {
    throw std::invalid_argument("tbsla: unsupported value/vector type combination for this format");
  }
}

void launch_fill(void *y, Dtype ydt, long n, double value, hipStream_t s) {
  if (n <= 0) return;
  const unsigned g = static_cast<unsigned>(std::min<long>((n + 255) / 256, 65536));
  if (ydt == Dtype::F64)
    hipLaunchKernelGGL(fill_kernel<double>, dim3(g), dim3(256), 0, s, static_cast<double *>(y), n, value);
  else if (ydt == Dtype::F32)
    hipLaunchKernelGGL(fill_kernel<float>, dim3(g), dim3(256), 0, s, static_cast<float *>(y), n,
                       static_cast<float>(value));
  else
    throw std::invalid_argument("tbsla: fill supports F64/F32 vectors");
  TBSLA_HIP_CHECK(hipGetLastError());
}

const char *spmv_kernel_name(cpp::Format f) {
  switch (f) {
    case cpp::FORMAT_CSR: return "csr_tile_kernel";
    case cpp::FORMAT_ELL: return "ell_kernel";
    case cpp::FORMAT_COO:
    case cpp::FORMAT_SCOO: return "coo_tile_kernel";
    case cpp::FORMAT_DENSE: return "dense_kernel";
  }
  return "";
}

}  // namespace hip
}  // namespace tbsla
