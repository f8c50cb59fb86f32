// Adjoint PageRank kernels and driver (the mathematics is in adjoint.hpp).
//
// Launch sequence:
//   set-up     adj_gather (values in transposed order, product precision),
//              adj_spmv_sum (sum of A p), adj_sums (1^T p, 1^T v, p^T w,
//              v^T w), adj_init (s, a, c)
//   iteration  adj_prep (u = (lambda - a)/s), adj_update (tile product on
//              A^T with the fused lambda' epilogue and four partial sums),
//              adj_finalize (merge the partials, convergence)
//   gradients  adj_grad_teleport, adj_grad_values
// Every iteration kernel returns at once once the state says done, so the
// host reads the state only every `check_every` iterations.
#include <hip/hip_runtime.h>

#include <tbsla/cpp/MatrixCSR.hpp>

#include <algorithm>
#include <stdexcept>

#include "adjoint.hpp"
#include "mirror.hpp"
#include "tile_csr.hpp"

namespace tbsla {
namespace hip {
namespace {

struct AdjState {
  double beta, s, sv, sp, eps, inv_n;  // s = 1^T M p, sv = 1^T v, sp = 1^T p
  double a, c;                         // p^T lambda, v^T lambda of the current lambda
  double diff, norm;                   // |lambda_k - lambda_{k-1}|_1, |lambda_k|_1
  int it, max_it, done, pad;
};

constexpr int kSumBlocks = 1024;

// Fixed-order sums of Q values per thread over a block of kTileThreads.
template <int Q>
__device__ inline void block_sums(double (&v)[Q], double (*red)[kTileThreads]) {
#pragma unroll
  for (int q = 0; q < Q; q++) red[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int s = kTileThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
#pragma unroll
      for (int q = 0; q < Q; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < Q; q++) v[q] = red[q][0];
  __syncthreads();
}

template <typename VT>
__global__ __launch_bounds__(256) void adj_gather(const int *__restrict__ perm, const double *__restrict__ src,
                                                  VT *__restrict__ dst, long n) {
  for (long k = blockIdx.x * 256L + threadIdx.x; k < n; k += static_cast<long>(gridDim.x) * 256)
    dst[k] = static_cast<VT>(src[perm[k]]);
}

// sum_i (A p)_i over the CSR tiles / long rows, one partial per block.
struct SumRow {
  double acc;
  __device__ void operator()(int, double s) { acc += s; }
};
template <bool LONG>
__global__ __launch_bounds__(kTileThreads) void adj_spmv_sum(dev::CsrView<double> m, const double *__restrict__ x,
                                                             double *__restrict__ part, int part_off) {
  __shared__ dev::CsrTileLds<double> lds;
  __shared__ double red[1][kTileThreads];
  SumRow epi{0.0};
  if (LONG)
    dev::csr_long_row<double, double>(m, x, lds, epi);
  else
    dev::csr_tile_rows<double, double>(m, x, lds, epi);
  double v[1] = {epi.acc};
  block_sums<1>(v, red);
  if (threadIdx.x == 0) part[part_off + blockIdx.x] = v[0];
}

// Partials: 1^T p, 1^T v, p^T w, v^T w (v = nullptr: uniform inv_n).
__global__ __launch_bounds__(kTileThreads) void adj_sums(const double *__restrict__ p, const double *__restrict__ v,
                                                         const double *__restrict__ w, double inv_n, long n,
                                                         double *__restrict__ part) {
  __shared__ double red[4][kTileThreads];
  double q[4] = {0, 0, 0, 0};
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kTileThreads) {
    const double vi = v ? v[i] : inv_n;
    q[0] += p[i];
    q[1] += vi;
    q[2] += p[i] * w[i];
    q[3] += vi * w[i];
  }
  block_sums<4>(q, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) part[4 * blockIdx.x + k] = q[k];
}

// One block: sum the set-up partials, fix s, a, c.
__global__ __launch_bounds__(kTileThreads) void adj_init(const double *__restrict__ part_ap, int n_ap,
                                                         const double *__restrict__ part4, int n4,
                                                         AdjState *st) {
  __shared__ double red[5][kTileThreads];
  double q[5] = {0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < n_ap; i += kTileThreads) q[0] += part_ap[i];
  for (int i = threadIdx.x; i < n4; i += kTileThreads)
    for (int k = 0; k < 4; k++) q[1 + k] += part4[4 * i + k];
  block_sums<5>(q, red);
  if (threadIdx.x == 0) {
    const double beta = st->beta;
    st->sp = q[1];
    st->sv = q[2];
    st->s = beta * q[0] + (1.0 - beta) * q[2] * q[1];
    st->a = q[3];
    st->c = q[4];
    st->done = st->max_it <= 0 ? 1 : 0;
  }
}

template <typename UT>
__global__ __launch_bounds__(256) void adj_prep(const double *__restrict__ lam, UT *__restrict__ u, long n,
                                                const AdjState *st) {
  if (st->done) return;
  const double a = st->a, inv_s = 1.0 / st->s;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    u[i] = static_cast<UT>((lam[i] - a) * inv_s);
}

// lambda'_j = beta * (A^T u)_j + (1-beta) * (v^T u) + w_j, with partial sums.
template <typename AT>
struct AdjRow {
  const double *w, *lam, *p, *v;
  double *lam_new;
  double beta, tv, inv_n;  // tv = (1-beta) * v^T u
  double q[4];
  __device__ void operator()(int j, AT y) {
    const double ln = beta * static_cast<double>(y) + tv + w[j];
    lam_new[j] = ln;
    q[0] += p[j] * ln;
    q[1] += (v ? v[j] : inv_n) * ln;
    q[2] += fabs(ln - lam[j]);
    q[3] += fabs(ln);
  }
};

template <typename VT, typename UT, bool LONG>
__global__ __launch_bounds__(kTileThreads) void adj_update(dev::CsrView<VT> mT, const UT *__restrict__ u,
                                                           const double *__restrict__ w,
                                                           const double *__restrict__ lam,
                                                           double *__restrict__ lam_new, const double *__restrict__ p,
                                                           const double *__restrict__ v, AdjState *st,
                                                           double *__restrict__ part, int part_off) {
  if (st->done) return;
  using AT = typename dev::Val<VT>::A;
  __shared__ dev::CsrTileLds<AT> lds;
  __shared__ double red[4][kTileThreads];
  const double beta = st->beta;
  const double vu = (st->c - st->a * st->sv) / st->s;
  AdjRow<AT> epi{w, lam, p, v, lam_new, beta, (1.0 - beta) * vu, st->inv_n, {0, 0, 0, 0}};
  if (LONG)
    dev::csr_long_row<VT, UT>(mT, u, lds, epi);
  else
    dev::csr_tile_rows<VT, UT>(mT, u, lds, epi);
  block_sums<4>(epi.q, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; k++) part[4 * (part_off + blockIdx.x) + k] = epi.q[k];
}

__global__ __launch_bounds__(kTileThreads) void adj_finalize(const double *__restrict__ part, int nparts,
                                                             AdjState *st) {
  if (st->done) return;
  __shared__ double red[4][kTileThreads];
  double q[4] = {0, 0, 0, 0};
  // four partials of U blocks per thread per step, the loads issued before
  // the additions (in the order of one block per step)
  constexpr int U = 4;
  for (int i0 = threadIdx.x; i0 < nparts; i0 += U * kTileThreads) {
    double t[U][4];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = min(i0 + u * kTileThreads, nparts - 1);
#pragma unroll
      for (int k = 0; k < 4; k++) t[u][k] = part[4 * i + k];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + u * kTileThreads < nparts)
#pragma unroll
        for (int k = 0; k < 4; k++) q[k] += t[u][k];
  }
  block_sums<4>(q, red);
  if (threadIdx.x == 0) {
    st->a = q[0];
    st->c = q[1];
    st->diff = q[2];
    st->norm = q[3];
    st->it += 1;
    st->done = (q[2] <= st->eps * q[3] || st->it >= st->max_it) ? 1 : 0;
  }
}

__global__ __launch_bounds__(256) void adj_grad_teleport(const double *__restrict__ lam, double *__restrict__ gt,
                                                         long n, const AdjState *st) {
  const double a = st->a, f = (1.0 - st->beta) * st->sp / st->s;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    gt[i] = f * (lam[i] - a);
}

// grad[e] = beta * (lambda_i - a)/s * p[col(e)] for the entries of 256 rows
// per workgroup: the rows' pointers are staged in LDS, each thread takes
// consecutive entries (coalesced writes) and finds its row by binary search.
__global__ __launch_bounds__(256) void adj_grad_values(const int *__restrict__ rowptr, const int *__restrict__ colidx,
                                                       int rows, int rp0, int f_col, const double *__restrict__ lam,
                                                       const double *__restrict__ p, double *__restrict__ gv,
                                                       const AdjState *st) {
  __shared__ int rpl[257];
  __shared__ double gl[256];
  const int r0 = blockIdx.x * 256;
  const int nr = min(256, rows - r0);
  const double a = st->a, f = st->beta / st->s;
  for (int i = threadIdx.x; i <= nr; i += 256) rpl[i] = rowptr[r0 + i] - rp0;
  if (static_cast<int>(threadIdx.x) < nr) gl[threadIdx.x] = f * (lam[r0 + threadIdx.x] - a);
  __syncthreads();
  const int e0 = rpl[0], e1 = rpl[nr];
  for (int e = e0 + threadIdx.x; e < e1; e += 256) {
    int lo = 0, hi = nr;  // the row i with rpl[i] <= e < rpl[i+1]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (rpl[mid] <= e)
        lo = mid;
      else
        hi = mid;
    }
    gv[e] = gl[lo] * p[colidx[e] - f_col];
  }
}

unsigned grid_for(long n) {
  return static_cast<unsigned>(std::max<long>(1, std::min<long>(kSumBlocks, (n + 255) / 256)));
}

template <typename VT, typename UT>
void run_adjoint(const DeviceMatrix &d64, const DeviceMatrix &dT, const AdjointParams &p, const double *pr,
                 const double *w, double *grad_values, double *grad_teleport, hipStream_t s, AdjointReport &rep) {
  const long n = dT.rows;
  const long nnz = dT.entries;
  const unsigned gN = grid_for(n);
  const double *src = p.values ? p.values : d64.vals.as<double>();
  DevBuf valsT, L0, L1, U, st, part, part_ap, part4;
  valsT.alloc(sizeof(VT) * static_cast<std::size_t>(std::max(1L, nnz)));
  L0.alloc(sizeof(double) * n);
  L1.alloc(sizeof(double) * n);
  U.alloc(sizeof(UT) * n);
  st.alloc(sizeof(AdjState));
  const int nparts = dT.ntiles + dT.nlong;
  part.alloc(sizeof(double) * 4 * static_cast<std::size_t>(std::max(1, nparts)));
  const int n_ap = d64.ntiles + d64.nlong;
  part_ap.alloc(sizeof(double) * static_cast<std::size_t>(std::max(1, n_ap)));
  part4.alloc(sizeof(double) * 4 * gN);

  AdjState h{};
  h.beta = p.beta;
  h.eps = p.epsilon;
  h.inv_n = 1.0 / static_cast<double>(n);
  h.max_it = p.max_iteration;
  TBSLA_HIP_CHECK(hipMemcpyAsync(st.get(), &h, sizeof(h), hipMemcpyHostToDevice, s));
  AdjState *dst = st.as<AdjState>();

  hipEvent_t e0, e1;
  TBSLA_HIP_CHECK(hipEventCreate(&e0));
  TBSLA_HIP_CHECK(hipEventCreate(&e1));
  TBSLA_HIP_CHECK(hipEventRecord(e0, s));
  if (nnz > 0)
    hipLaunchKernelGGL((adj_gather<VT>), dim3(grid_for(nnz)), dim3(256), 0, s, dT.perm.as<int>(), src,
                       valsT.as<VT>(), nnz);
  // s = beta * 1^T(A p) + (1-beta) (1^T v)(1^T p)
  dev::CsrView<double> a = dev::csr_view<double>(d64);
  a.vals = src;
  if (d64.ntiles > 0)
    hipLaunchKernelGGL(adj_spmv_sum<false>, dim3(d64.ntiles), dim3(kTileThreads), 0, s, a, pr, part_ap.as<double>(),
                       0);
  if (d64.nlong > 0)
    hipLaunchKernelGGL(adj_spmv_sum<true>, dim3(d64.nlong), dim3(kTileThreads), 0, s, a, pr, part_ap.as<double>(),
                       d64.ntiles);
  hipLaunchKernelGGL(adj_sums, dim3(gN), dim3(kTileThreads), 0, s, pr, p.teleport, w, h.inv_n, n,
                     part4.as<double>());
  hipLaunchKernelGGL(adj_init, dim3(1), dim3(kTileThreads), 0, s, part_ap.as<double>(), n_ap, part4.as<double>(),
                     static_cast<int>(gN), dst);
  TBSLA_HIP_CHECK(hipMemcpyAsync(L0.get(), w, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  TBSLA_HIP_CHECK(hipGetLastError());
  // the state copy above reads host memory
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));

  double *L[2] = {L0.as<double>(), L1.as<double>()};
  dev::CsrView<VT> mt = dev::csr_view<VT>(dT);
  mt.vals = valsT.as<VT>();
  AdjState hs{};
  const int check = std::max(1, p.check_every);
  int launched = 0;
  while (launched < p.max_iteration) {
    const int batch = std::min(check, p.max_iteration - launched);
    for (int k = 0; k < batch; k++) {
      const int it = launched + k;
      double *lin = L[it & 1], *lout = L[(it + 1) & 1];
      hipLaunchKernelGGL((adj_prep<UT>), dim3(gN), dim3(256), 0, s, lin, U.as<UT>(), n, dst);
      if (dT.ntiles > 0)
        hipLaunchKernelGGL((adj_update<VT, UT, false>), dim3(dT.ntiles), dim3(kTileThreads), 0, s, mt, U.as<UT>(), w,
                           lin, lout, pr, p.teleport, dst, part.as<double>(), 0);
      if (dT.nlong > 0)
        hipLaunchKernelGGL((adj_update<VT, UT, true>), dim3(dT.nlong), dim3(kTileThreads), 0, s, mt, U.as<UT>(), w,
                           lin, lout, pr, p.teleport, dst, part.as<double>(), dT.ntiles);
      hipLaunchKernelGGL(adj_finalize, dim3(1), dim3(kTileThreads), 0, s, part.as<double>(), nparts, dst);
    }
    TBSLA_HIP_CHECK(hipGetLastError());
    launched += batch;
    TBSLA_HIP_CHECK(hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    if (hs.done) break;
  }
  if (p.max_iteration <= 0) {
    TBSLA_HIP_CHECK(hipMemcpyAsync(&hs, dst, sizeof(hs), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  }
  const double *lam = L[hs.it & 1];
  if (grad_teleport)
    hipLaunchKernelGGL(adj_grad_teleport, dim3(gN), dim3(256), 0, s, lam, grad_teleport, n, dst);
  if (grad_values && d64.rows > 0)
    hipLaunchKernelGGL(adj_grad_values, dim3((d64.rows + 255) / 256), dim3(256), 0, s, d64.rowptr.as<int>(),
                       d64.colidx.as<int>(), d64.rows, d64.rp0, d64.f_col, lam, pr, grad_values, dst);
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipEventRecord(e1, s));
  TBSLA_HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  TBSLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rep.iterations = hs.it;
  rep.last_change = hs.norm > 0 ? hs.diff / hs.norm : 0.0;
  rep.s = hs.s;
  rep.seconds = ms * 1e-3;
}

}  // namespace

AdjointReport page_rank_adjoint(const cpp::Matrix &m, const AdjointParams &p, const double *pr, const double *w,
                                double *grad_values, double *grad_teleport, hipStream_t s) {
  if (m.format() != cpp::FORMAT_CSR) throw std::invalid_argument("tbsla: the PageRank adjoint needs a CSR matrix");
  if (m.get_n_row() != m.get_n_col() || m.get_NR() != 1 || m.get_NC() != 1)
    throw std::invalid_argument("tbsla: page_rank needs a square matrix held in one block");
  if (!pr || !w) throw std::invalid_argument("tbsla: NULL PageRank vector or weight");
  AdjointReport rep;
  const long n = m.get_n_col();
  if (n == 0) return rep;
  const DeviceMatrix &d64 = mirror(m, Dtype::F64);
  if (d64.noop || d64.rows != n) throw std::invalid_argument("tbsla: CSR rowptr does not cover every row");
  if (d64.x_need > n) throw std::invalid_argument("tbsla: column index beyond n_col");
  const DeviceMatrix &dT = transpose_mirror(m);
  if (p.precision == Dtype::F32)
    run_adjoint<float, float>(d64, dT, p, pr, w, grad_values, grad_teleport, s, rep);
  else if (p.precision == Dtype::F64)
    run_adjoint<double, double>(d64, dT, p, pr, w, grad_values, grad_teleport, s, rep);
  else
    throw std::invalid_argument("tbsla: the PageRank adjoint runs in F32 or F64");
  return rep;
}

}  // namespace hip
}  // namespace tbsla
