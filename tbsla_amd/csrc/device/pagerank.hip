// PageRank iteration kernels and their host driver (see pagerank.hpp).
//
// One iteration of the CSR engine is four launches (five with long rows):
//   pr_update_csr     entry-tile SpMV of the current iterate fused with
//                     b = beta*(A b_t) + tele and the block max of b, merged
//                     into 64 atomic-max slots.
//   pr_finalize_max   one wavefront: max over the slots into the state.
//   pr_normalize      b /= max, the L1 residual against b_t and the
//                     compensated sum of b, as per-block partials.
//   pr_finalize_norm  one block: merges the partials and updates the state:
//                     iteration count, convergence, the next teleport term,
//                     and (mixed precision) whether to move the anchor.
// Partial results cross workgroups only at kernel boundaries, so no kernel
// needs a device-wide fence.
// Every kernel returns at once when the state says the run has converged, so
// the host launches `check_every` iterations between two reads of the state.
#include <hip/hip_runtime.h>

#include <tbsla/cpp/MatrixCSR.hpp>

#include <chrono>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "mirror.hpp"
#include <type_traits>

#include "pagerank.hpp"
#include "tile_csr.hpp"

namespace tbsla {
namespace hip {
namespace {

struct PrState {
  double beta, eps, bound_target, tele_scale;
  double tele;          // teleport term used by the next update
  double maxv;          // max of the last update
  double err;           // reference residual sum|b_k - b_{k-1}| of the last iteration
  double err_prev;      // ... and of the one before
  double sum, sum_prev; // sum of the last / previous normalised iterate (compensated)
  double bound, bound_prev, bound_anchor;
  double anchor_tau, floor_eps;
  // sum-normalised propagation-blocked runs: 1 / sum of the last iterate
  // (the next iterate's scale), and for mixed runs the sum of the anchor,
  // the teleport term and the scale of the iteration that moved it
  double inv_sum, sum_hat, t_anchor, c_anchor;
  // reference-rule runs with propagation-blocked products keep the update's
  // unnormalised b (raw = 1): the iterate is b * xscale, xscale = 1 / max of
  // the update that wrote it (1 for the start and in every other run)
  double xscale;
  int it, max_it, done, rule;
  int need_anchor, anchors, anchor_it, mixed, raw;
};

constexpr int kNormBlocks = 1024;

// Neumaier-compensated running sum and its exact-error merge (TwoSum).
struct CSum {
  double s, c;
};
__device__ inline void csum_add(CSum &a, double v) {
  const double t = a.s + v;
  if (fabs(a.s) >= fabs(v))
    a.c += (a.s - t) + v;
  else
    a.c += (v - t) + a.s;
  a.s = t;
}
__device__ inline CSum csum_merge(CSum a, CSum b) {
  const double t = a.s + b.s;
  const double bp = t - a.s;
  const double e = (a.s - (t - bp)) + (b.s - bp);
  return CSum{t, a.c + b.c + e};
}

// Fixed-order butterfly reductions over one wavefront (every lane gets the
// result) and over the 16 wavefronts of a 1024-thread workgroup (thread 0
// gets the result; red holds 48 doubles).
__device__ inline CSum wave_csum(CSum v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = csum_merge(v, CSum{__shfl_xor(v.s, off, 64), __shfl_xor(v.c, off, 64)});
  return v;
}
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline void block1024_err_csum(double err, CSum cs, double *red, double &err_out, CSum &cs_out) {
  err = wave_sum(err);
  cs = wave_csum(cs);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w] = err;
    red[16 + w] = cs.s;
    red[32 + w] = cs.c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    err_out = 0;
    cs_out = CSum{0, 0};
    for (int i = 0; i < 16; i++) {
      err_out += red[i];
      cs_out = csum_merge(cs_out, CSum{red[16 + i], red[32 + i]});
    }
  }
}

__device__ inline CSum block_csum(CSum v, double *ps, double *pc) {
  ps[threadIdx.x] = v.s;
  pc[threadIdx.x] = v.c;
  __syncthreads();
  for (int s = kTileThreads / 2; s > 0; s >>= 1) {
    if (static_cast<int>(threadIdx.x) < s) {
      CSum m = csum_merge(CSum{ps[threadIdx.x], pc[threadIdx.x]}, CSum{ps[threadIdx.x + s], pc[threadIdx.x + s]});
      ps[threadIdx.x] = m.s;
      pc[threadIdx.x] = m.c;
    }
    __syncthreads();
  }
  CSum r{ps[0], pc[0]};
  __syncthreads();
  return r;
}

// Teleport term of row i, by mode:
//   TM = 0  uniform: + tele on every row (tele = (1-beta)/n * sum b_t)
//   TM = 1  node set: + tele on listed rows (tele = (1-beta)/|nodes| * sum b_t)
//   TM = 2  distribution v: + tele * v[i] (tele = (1-beta) * sum b_t)
struct Teleport {
  const unsigned char *member;
  const double *v;
};
template <int TM>
__device__ inline double add_teleport(double bi, double tele, const Teleport &tp, long i) {
#pragma clang fp contract(off)
  if (TM == 0) return bi + tele;
  if (TM == 1) return tp.member[i] ? bi + tele : bi;
  return bi + tele * tp.v[i];
}

// Per-row update epilogue: b = beta*(yhat + s) + teleport (MIXED) or beta*s + teleport.
template <typename AT, bool MIXED, int TM>
struct UpdateRow {
  double *b;
  const double *yhat;
  Teleport tp;
  double beta, tele, lmax;
  __device__ void operator()(int row, AT s) {
#pragma clang fp contract(off)
    const double v = MIXED ? yhat[row] + static_cast<double>(s) : static_cast<double>(s);
    const double bi = add_teleport<TM>(beta * v, tele, tp, row);
    b[row] = bi;
    lmax = fmax(lmax, bi);
  }
};

// Block maxima are merged with integer atomicMax on 64 slots (a block uses
// slot blockIdx % 64), each slot holding an order-preserving encoding of a
// double (0 = empty), so the finalize reads 64 numbers however many blocks ran.
constexpr int kMaxSlots = 64;

__device__ inline unsigned long long ord_encode(double v) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double ord_decode(unsigned long long u) {
  if (u == 0) return -INFINITY;
  return __longlong_as_double(static_cast<long long>((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}
__device__ inline void slot_max(unsigned long long *slots, double v, double *red) {
  const double bm = dev::block_tree_max<double>(v, red);
  if (threadIdx.x == 0) atomicMax(slots + (blockIdx.x % kMaxSlots), ord_encode(bm));
}

// Fused update over the CSR tiles (LONG = false) or the long rows (LONG =
// true): b = beta * (A x) + tele, block max into the slots.  VT = dev::Pattern
// is the value-free product over z = colscale .* b_t.
template <typename VT, typename XT, bool MIXED, int TM, bool LONG>
__global__ __launch_bounds__(kTileThreads) void pr_update_csr(dev::CsrView<VT> m, const XT *__restrict__ x,
                                                              const double *__restrict__ yhat, Teleport tp,
                                                              double *__restrict__ b, PrState *st,
                                                              unsigned long long *slots) {
  if (st->done) return;
  using AT = typename dev::Val<VT>::A;
  __shared__ dev::CsrTileLds<AT> lds;
  __shared__ double red[kTileThreads];
  UpdateRow<AT, MIXED, TM> epi{b, yhat, tp, st->beta, st->tele, -INFINITY};
  if (LONG)
    dev::csr_long_row<VT, XT>(m, x, lds, epi);
  else
    dev::csr_tile_rows<VT, XT>(m, x, lds, epi);
  slot_max(slots, epi.lmax, red);
}

// Update for any format after a plain product y = A b_t (generic engine).
template <int TM>
__global__ __launch_bounds__(kTileThreads) void pr_update_vec(double *__restrict__ b, long n, Teleport tp,
                                                              PrState *st, unsigned long long *slots) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double red[kTileThreads];
  const double beta = st->beta, tele = st->tele;
  double lmax = -INFINITY;
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kTileThreads) {
    const double bi = add_teleport<TM>(beta * b[i], tele, tp, i);
    b[i] = bi;
    lmax = fmax(lmax, bi);
  }
  slot_max(slots, lmax, red);
}

// Decision logic of the last block of pr_normalize (one thread).
__device__ inline void finalize_step(PrState *st, double err, double S) {
  // (mixed runs) this iteration moved the anchor: its step is the FP64
  // anchor product's, so its bound is free of the low-precision error
  const bool moved = st->mixed && st->need_anchor;
  st->it += 1;
  st->err_prev = st->err;
  st->err = err;
  st->sum_prev = st->sum;
  st->sum = S;
  const double bound = err / S + fabs(st->sum_prev / S - 1.0);
  st->bound_prev = st->bound;
  st->bound = bound;
  bool converged = st->rule == 0 ? (err < st->eps) : (bound <= st->bound_target);
  int anchor = 0;
  if (st->mixed) {
    // A low-precision phase converges to a point off the FP64 fixed point by
    // up to about floor_eps * (bound at its anchor) / (1 - rho)^2, rho the
    // phase's mean contraction per iteration.  When that may exceed the
    // target, convergence is only declared after an iteration that moved the
    // anchor (an FP64 step) meets the bound.
    if (converged && !moved) {
      const int k = st->it - st->anchor_it > 0 ? st->it - st->anchor_it : 1;
      double rho = 0.999;
      if (st->bound_anchor > 0.0 && bound > 0.0) rho = fmin(0.999, pow(bound / st->bound_anchor, 1.0 / k));
      const double floor = st->floor_eps * st->bound_anchor * 10.0 / ((1.0 - rho) * (1.0 - rho));
      if (floor > st->bound_target) {
        converged = false;
        anchor = 1;
      }
    }
    if (!converged) {
      if (bound <= st->anchor_tau * st->bound_anchor) anchor = 1;
      // progress has stopped (the phase reached its low-precision floor)
      if (st->it - st->anchor_it >= 3 && bound > 0.99 * st->bound_prev) anchor = 1;
    }
    if (anchor) {
      st->bound_anchor = bound;
      st->anchor_it = st->it;
      st->anchors += 1;
    }
  }
  st->need_anchor = anchor;
  st->done = (converged || st->it >= st->max_it) ? 1 : 0;
  st->tele = S * st->tele_scale;
}

// MODE 0: FP64, b normalised in place, previous iterate `prev`.
// MODE 4: MODE 0 with b and prev left unnormalised (raw runs): b * (1/max)
//         and prev * xscale are the iterates, nothing is written.
// MODE 1: mixed, previous iterate xhat + dold, writes dnew = b - xhat (FP32).
// MODE 2: FP64 + z = colscale .* b for the next value-free product.
// MODE 3: MODE 1 + zd = colscale .* (b - xhat) in FP32 (colscale32) for the
//         next value-free FP32 product.
template <int MODE>
__global__ __launch_bounds__(kTileThreads) void pr_normalize(double *__restrict__ b, const double *__restrict__ prev,
                                                             const double *__restrict__ xhat,
                                                             const float *__restrict__ dold, float *__restrict__ dnew,
                                                             double *__restrict__ z,
                                                             const double *__restrict__ colscale,
                                                             float *__restrict__ zd,
                                                             const float *__restrict__ colscale32, long n, PrState *st,
                                                             double *__restrict__ perr, double *__restrict__ ps,
                                                             double *__restrict__ pc) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  const double m = st->maxv;
  const double inv_m = 1.0 / m, pscale = st->xscale;  // (MODE 4)
  const bool moved = st->need_anchor != 0;
  double err = 0;
  CSum cs{0, 0};
  if (MODE == 4) {
    // four elements of the thread's sequence per step, loads first (the
    // same order of additions as one element per step)
    const long stride = static_cast<long>(gridDim.x) * kTileThreads;
    for (long i0 = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i0 < n; i0 += 4 * stride) {
      double bv[4], pv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const long i = min(i0 + u * stride, n - 1);
        bv[u] = b[i];
        pv[u] = prev[i];
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (i0 + u * stride < n) {
          const double bi = bv[u] * inv_m;
          err += fabs(bi - pv[u] * pscale);
          csum_add(cs, bi);
        }
    }
  }
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; MODE != 4 && i < n;
       i += static_cast<long>(gridDim.x) * kTileThreads) {
    const double bi = b[i] / m;
    const bool mixed = MODE == 1 || MODE == 3;
    // an iteration that moved the anchor starts from xhat itself (its delta
    // was folded into xhat and is not read)
    const double pi = mixed ? (moved ? xhat[i] : xhat[i] + static_cast<double>(dold[i])) : prev[i];
    err += fabs(bi - pi);
    csum_add(cs, bi);
    if (mixed) {
      const double dd = bi - xhat[i];
      dnew[i] = static_cast<float>(dd);
      if (MODE == 3) zd[i] = static_cast<float>(static_cast<double>(colscale32[i]) * dd);
    } else {
      b[i] = bi;
      if (MODE == 2) z[i] = colscale[i] * bi;
    }
  }
  const double berr = dev::block_tree_sum<double>(err, red_s);
  const CSum bs = block_csum(cs, red_s, red_c);
  if (threadIdx.x == 0) {
    perr[blockIdx.x] = berr;
    ps[blockIdx.x] = bs.s;
    pc[blockIdx.x] = bs.c;
  }
}

// One wavefront: the maximum over the slots, which are cleared for the next
// update.  Slots are read and cleared with device-scope atomics.
__global__ __launch_bounds__(64) void pr_finalize_max(unsigned long long *slots, PrState *st) {
  if (st->done) return;
  unsigned long long u = __hip_atomic_load(slots + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(slots + threadIdx.x, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(u, off, 64);
    u = o > u ? o : u;
  }
  if (threadIdx.x == 0) st->maxv = ord_decode(u);
}

// Sum of the per-block partials (err, compensated sum) over nb blocks by one
// kTileThreads block, in the fixed order thread t: t, t + kTileThreads, ...
// The loads of 8 partials per thread are issued together (a loop that loads
// one partial per step waits for each load in turn).
__device__ inline void sum_partials(const double *__restrict__ perr, const double *__restrict__ ps,
                                    const double *__restrict__ pc, int nb, double &e, CSum &t) {
  constexpr int U = 8;
  for (int i0 = threadIdx.x; i0 < nb; i0 += U * kTileThreads) {
    double pe[U], pv[U], pw[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = min(i0 + u * kTileThreads, nb - 1);
      pe[u] = perr ? perr[i] : 0.0;
      pv[u] = ps[i];
      pw[u] = pc[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + u * kTileThreads < nb) {
        e += pe[u];
        t = csum_merge(t, CSum{pv[u], pw[u]});
      }
  }
}

// One block: residual and compensated sum of the normalise pass, then the
// step decision.
__global__ __launch_bounds__(kTileThreads) void pr_finalize_norm(const double *__restrict__ perr,
                                                                 const double *__restrict__ ps,
                                                                 const double *__restrict__ pc, int nb, PrState *st) {
  if (st->done) return;
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  double e = 0;
  CSum t{0, 0};
  sum_partials(perr, ps, pc, nb, e, t);
  e = dev::block_tree_sum<double>(e, red_s);
  t = block_csum(t, red_s, red_c);
  if (threadIdx.x == 0) {
    finalize_step(st, e, t.s + t.c);
    if (st->raw) st->xscale = 1.0 / st->maxv;  // the scale of the b just written
  }
}

// Mixed precision with tile products: move the anchor to the current iterate
// when requested, xhat += delta (and z = colscale .* xhat for the value-free
// anchor product).  The tile products still add the delta product in that
// iteration, so its input `clear` (the delta, or the scaled delta of
// PATTERN_F32) is zeroed; the normalise pass of that iteration starts from
// xhat.  (With propagation-blocked products the move is part of the scatter,
// and that iteration has no delta product.)
__global__ __launch_bounds__(256) void pr_anchor_move(double *__restrict__ xhat, const float *delta, float *clear,
                                                      double *__restrict__ z, const double *__restrict__ colscale,
                                                      long n, const PrState *st) {
  if (st->done || !st->need_anchor) return;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) {
    const double v = xhat[i] + static_cast<double>(delta[i]);
    xhat[i] = v;
    if (clear) clear[i] = 0.0f;
    if (z) z[i] = colscale[i] * v;
  }
}

template <typename YT>
struct StoreRowD {
  YT *y;
  template <typename AT>
  __device__ void operator()(int row, AT s) {
    y[row] = static_cast<YT>(s);
  }
};

// Mixed precision: the FP64 anchor product yhat = A64 xhat (VT = double) or
// its value-free form over z = colscale .* xhat (VT = dev::Pattern).
template <typename VT, bool LONG>
__global__ __launch_bounds__(kTileThreads) void pr_anchor_spmv(dev::CsrView<VT> m, const double *__restrict__ x,
                                                               double *__restrict__ yhat, const PrState *st) {
  if (st->done || !st->need_anchor) return;
  __shared__ dev::CsrTileLds<double> lds;
  StoreRowD<double> epi{yhat};
  if (LONG)
    dev::csr_long_row<VT, double>(m, x, lds, epi);
  else
    dev::csr_tile_rows<VT, double>(m, x, lds, epi);
}

// v = a (+ d) ; out = v / sum(v) with a compensated sum (two passes).
// (the element is (a + d) * xscale: xscale is 1 except in raw runs)
__global__ __launch_bounds__(kTileThreads) void pr_sum_kernel(const double *__restrict__ a,
                                                              const float *__restrict__ d, long n,
                                                              const PrState *st, double *__restrict__ ps,
                                                              double *__restrict__ pc) {
#pragma clang fp contract(off)
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  const double sc = st->xscale;
  CSum cs{0, 0};
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kTileThreads)
    csum_add(cs, (d ? a[i] + static_cast<double>(d[i]) : a[i]) * sc);
  const CSum bs = block_csum(cs, red_s, red_c);
  if (threadIdx.x == 0) {
    ps[blockIdx.x] = bs.s;
    pc[blockIdx.x] = bs.c;
  }
}

__global__ __launch_bounds__(kTileThreads) void pr_sum_finalize(const double *__restrict__ ps,
                                                                const double *__restrict__ pc, int nb,
                                                                double *__restrict__ total) {
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  CSum t{0, 0};
  for (int i = threadIdx.x; i < nb; i += kTileThreads) t = csum_merge(t, CSum{ps[i], pc[i]});
  t = block_csum(t, red_s, red_c);
  if (threadIdx.x == 0) *total = t.s + t.c;
}

__global__ __launch_bounds__(256) void pr_scale_kernel(const double *__restrict__ a, const float *__restrict__ d,
                                                       long n, const PrState *st, const double *__restrict__ total,
                                                       double *__restrict__ out) {
#pragma clang fp contract(off)
  const double S = *total, sc = st->xscale;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    out[i] = ((d ? a[i] + static_cast<double>(d[i]) : a[i]) * sc) / S;
}

// cs[colidx[k] - f_col] = vals[k] over the stored entries: for a
// count-stochastic matrix every entry of a column holds the same value, so
// the order of the writes does not matter.
__global__ __launch_bounds__(256) void colscale_kernel(const int *__restrict__ colidx, const double *__restrict__ vals,
                                                       long nnz, int f_col, long n, double *__restrict__ cs) {
  for (long k = blockIdx.x * 256L + threadIdx.x; k < nnz; k += static_cast<long>(gridDim.x) * 256) {
    const long c = static_cast<long>(colidx[k]) - f_col;
    if (c >= 0 && c < n) cs[c] = vals[k];
  }
}
__global__ __launch_bounds__(256) void to_f32(const double *__restrict__ a, float *__restrict__ b, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    b[i] = static_cast<float>(a[i]);
}

__global__ __launch_bounds__(256) void fill_f64(double *y, long n, double v) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) y[i] = v;
}
__global__ __launch_bounds__(256) void fill_f32(float *y, long n, float v) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) y[i] = v;
}
__global__ __launch_bounds__(256) void scale_into(const double *__restrict__ s, const double *__restrict__ x,
                                                  double *__restrict__ z, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) z[i] = s[i] * x[i];
}

__global__ __launch_bounds__(256) void scale_into_f32(const double *__restrict__ s, double v, float *__restrict__ z,
                                                      long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    z[i] = static_cast<float>(s[i] * v);
}

unsigned grid_for(long n, long cap = kNormBlocks) {
  return static_cast<unsigned>(std::max<long>(1, std::min<long>(cap, (n + 255) / 256)));
}

// ---------------------------------------------------------------------------
// Propagation-blocked product (layout in mirror.hpp, PbMatrix).
//   pb_scatter  one workgroup per column superblock (8192 columns): the
//               superblock's slice of x is staged in LDS, then every far
//               product v * x[c] is written to its slot in the bin-grouped
//               product array; entries come in (bin, column) order, so
//               consecutive threads write runs of consecutive slots.
//   pb_gather   one 1024-thread workgroup per bin of 8192 rows: the bin's
//               far products and local products (x from the bin's own slice)
//               are added into LDS, then the row epilogue runs on the sums.
// The LDS additions are atomic, so the summation order within a row is not
// fixed (the result can differ in the last bits from run to run).
template <typename VT>
struct PbVal {
  template <typename AT>
  __device__ static AT prod(const VT *v, long k, AT x) {
    return static_cast<AT>(dev::Val<VT>::get(v, k)) * x;
  }
  // four values from a multiple of four
  template <typename AT>
  __device__ static void get4(const VT *v, long k, AT (&out)[4]) {
    typename dev::Val<VT>::A t[4];
    dev::Val<VT>::get4(v, k, t);
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = static_cast<AT>(t[q]);
  }
  template <typename AT>
  __device__ static AT mul(AT v, AT x) {
    return v * x;
  }
};
template <>
struct PbVal<dev::Pattern> {
  template <typename AT>
  __device__ static AT prod(const dev::Pattern *, long, AT x) {
    return x;
  }
  template <typename AT>
  __device__ static void get4(const dev::Pattern *, long, AT (&out)[4]) {
#pragma unroll
    for (int q = 0; q < 4; q++) out[q] = AT(1);
  }
  template <typename AT>
  __device__ static AT mul(AT, AT x) {
    return x;
  }
};

constexpr int kPbScatterThreads = 1024;

struct PbScatterArgs {
  const int *sp;              // far entries of superblock s: [sp[s], sp[s+1])
  const unsigned short *fkey; // column within the superblock | kPbRunStart
  const int *rp;              // runs of superblock s: [rp[s], rp[s+1])
  const int *rdelta;          // slot - entry of each run
  const int *cbase;           // run of the first entry of each 64-entry chunk
  const int *cp;              // chunks of superblock s: [cp[s], cp[s+1])
  int nsb, n;
  long nfar;                  // product slots; kPbChunk spare slots follow them
};

// The anchor move of a mixed run, folded into the scatter of the iteration
// that makes it: over the superblock's columns c, xhat[c] += delta[c], and
// the product's input is colscale[c] * xhat[c] (value-free anchor product;
// the gather forms the same product for its local entries) or xhat[c] itself
// (colscale == nullptr), and xhat32[c] = fl32(xhat[c]).  Each superblock
// leaves the compensated sum of its new xhat in (psum, pcomp)[superblock].
struct AnchorMove {
  double *xhat;
  float *xhat32;  // fl32(xhat), for the scale term of the mixed iterations (or nullptr)
  const float *delta;
  // value-free anchor products: 1/count of each column (the product's input
  // is colscale .* xhat, formed where it is read), else nullptr
  const double *colscale;
  double *psum, *pcomp;
  int first;  // set on the device: the run's first move, whose anchor is still 0
};

// The far products of one superblock (picked from blockIdx): its slice of x
// staged in xs (LDS), then every far product v * x[c] written to its slot.
// With MOVE the slice is the anchor move's result instead of x.  COLS: the
// superblock width of the layout `a` (kPbCols or kPbColsWide).
template <typename VT, typename XT, typename AT, bool MOVE, int COLS>
__device__ inline void pb_scatter_sb(const PbScatterArgs &a, const VT *__restrict__ val, const XT *__restrict__ x,
                                     AT *__restrict__ prod, XT *xs, const AnchorMove &am, double *red,
                                     XT xsc = XT(1)) {
#pragma clang fp contract(off)
  // superblocks of one XCD (blockIdx % 8) are contiguous, so the runs that
  // share a cache line at a run boundary are written through the same L2
  // (the grid may be larger than the layout needs: the mixed kernel serves
  // two layouts)
  const int per = (a.nsb + 7) / 8;
  if (static_cast<int>(blockIdx.x) >= 8 * per) return;
  const int sb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (sb >= a.nsb) return;
  const int k0 = a.sp[sb], k1 = a.sp[sb + 1];
  const int *__restrict__ delta = a.rdelta + a.rp[sb];  // slot - entry of each run of this superblock
  const int *__restrict__ cbase = a.cbase + a.cp[sb];   // run of the first entry of each chunk
  const long c0 = static_cast<long>(sb) * COLS;
  const int nc = static_cast<int>(min(static_cast<long>(COLS), static_cast<long>(a.n) - c0));
  {
    constexpr int S = COLS / kPbScatterThreads;  // x values per thread, all loads issued first
    XT t[S];
    CSum cs{0, 0};
#pragma unroll
    for (int u = 0; u < S; u++) {
      const int i = threadIdx.x + u * kPbScatterThreads;
      if (MOVE) {
        if (i < nc) {
          const long c = c0 + i;
          // (0 + d is d: the first move need not read the zero anchor)
          const double v = am.first ? static_cast<double>(am.delta[c]) : am.xhat[c] + static_cast<double>(am.delta[c]);
          am.xhat[c] = v;
          if (am.xhat32) am.xhat32[c] = static_cast<float>(v);
          csum_add(cs, v);
          t[u] = static_cast<XT>(am.colscale ? am.colscale[c] * v : v);
        } else {
          t[u] = XT(0);
        }
      } else {
        t[u] = i < nc ? static_cast<XT>(x[c0 + i] * xsc) : XT(0);
      }
    }
#pragma unroll
    for (int u = 0; u < S; u++) xs[threadIdx.x + u * kPbScatterThreads] = t[u];
    if (MOVE) {
      double e_unused;
      CSum tot;
      block1024_err_csum(0.0, cs, red, e_unused, tot);
      if (threadIdx.x == 0) {
        am.psum[sb] = tot.s;
        am.pcomp[sb] = tot.c;
      }
    }
  }
  __syncthreads();
  // Each wavefront walks chunks wv + 16u + 16*UC*it (u < UC) of the
  // superblock, one entry per lane, as a software pipeline: the run of each
  // chunk's first entry (cbase) is loaded two iterations ahead; the keys,
  // the values and the slot offsets of the chunk's runs (rdelta[cb .. cb +
  // 63], one coalesced load, each lane's offset picked with a lane shuffle)
  // one iteration ahead, so no load of an iteration waits on another.  An
  // entry's run is the chunk's first run plus the run starts on lanes
  // 1..lane (a ballot).
  static_assert(kPbScatterThreads == 16 * kPbChunk && kPbChunk == 64, "16 wavefronts of 64-entry chunks");
  // chunks per wavefront per stage (two stages in flight fit 64 VGPRs)
  constexpr int UC = sizeof(AT) == 8 ? 2 : 4;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nent = k1 - k0, nck = (nent + kPbChunk - 1) / kPbChunk;
  const int nr = a.rp[sb + 1] - a.rp[sb];  // runs of this superblock
  const int nit = (nck + 16 * UC - 1) / (16 * UC);
  struct Stage {
    unsigned key[UC];
    AT v[UC];
    int dl[UC];
  };
  if (nent == 0) return;
  // Every load is unconditional (indices clamped into the superblock) and
  // every store too (lanes past the end write spare slots), so the loop is
  // straight-line code and each wait covers only the loads it needs; the
  // loop is unrolled twice with the two stages in fixed registers (no
  // register copies, which would wait for the loads in flight).  On the
  // 2^24-node graph this took the FP32 delta scatter from 193 to 177 us.
  auto load_cb = [&](int it, int (&cb)[UC]) {
#pragma unroll
    for (int u = 0; u < UC; u++) cb[u] = cbase[min(wv + 16 * u + 16 * UC * it, nck - 1)];
  };
  // (lanes past the end keep what they load: their run starts only count
  // for lanes past the end, and their products go to the spare slots)
  auto load_stage = [&](int it, const int (&cb)[UC], Stage &g) {
#pragma unroll
    for (int u = 0; u < UC; u++) {
      const int ec = k0 + min((wv + 16 * u + 16 * UC * it) * kPbChunk + lane, nent - 1);
      g.key[u] = a.fkey[ec];
      g.v[u] = PbVal<VT>::template prod<AT>(val, ec, AT(1));
      g.dl[u] = delta[min(cb[u] + lane, nr - 1)];
    }
  };
  auto store = [&](int it, const Stage &g) {
#pragma unroll
    for (int u = 0; u < UC; u++) {
      const int e = (wv + 16 * u + 16 * UC * it) * kPbChunk + lane;
      const unsigned long long starts = __ballot((g.key[u] & kPbRunStart) != 0u);
      // run starts on lanes 1..lane: those below this lane (mbcnt) plus its own
      const unsigned long long s1 = starts & ~1ull;
      const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
          static_cast<unsigned>(s1 >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(s1), 0u)));
      const int d = __shfl(g.dl[u], below + static_cast<int>((s1 >> lane) & 1ull), 64);
      // lanes past the end write a spare slot, so the store is unconditional
      const long o = e < nent ? static_cast<long>(k0 + e + d) : a.nfar + lane;
      prod[o] = g.v[u] * static_cast<AT>(xs[g.key[u] & static_cast<unsigned>(COLS - 1)]);
    }
  };
  // iteration it: the loads of stage it + 1 (its chunk runs were loaded in
  // iteration it - 1) and of the chunk runs of it + 2, then the stores of
  // stage it (loaded in iteration it - 1); chunks past the end load the
  // superblock's last entry and store to the spare slots
  int cb0[UC], cb1[UC];
  Stage s0, s1;
  load_cb(0, cb0);
  load_cb(1, cb1);
  __builtin_amdgcn_sched_barrier(0);  // (the loop's order: chunk runs, then a stage)
  load_stage(0, cb0, s0);
  // (both halves of an iteration use what they load: the compiler would
  // sink a load past a branch into the block that uses it)
  int it = 0;
  for (; it + 1 < nit; it += 2) {
    load_cb(it + 2, cb0);  // (first: the next half waits for it, not for the stage loads)
    load_stage(it + 1, cb1, s1);
    __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the stores
    store(it, s0);
    load_cb(it + 3, cb1);
    load_stage(it + 2, cb0, s0);
    __builtin_amdgcn_sched_barrier(0);
    store(it + 1, s1);
  }
  if (it < nit) store(it, s0);
}

// FP64 runs: the far products of x.
template <typename VT, typename XT, typename AT, int COLS>
__global__ __launch_bounds__(kPbScatterThreads) void pb_scatter(PbScatterArgs a, const VT *__restrict__ val,
                                                                const XT *__restrict__ x, AT *__restrict__ prod,
                                                                const PrState *st) {
  if (st->done) return;
  __shared__ XT xs[COLS];
  // (raw reference-rule runs: x is the unnormalised b, the iterate x * xscale)
  pb_scatter_sb<VT, XT, AT, false, COLS>(a, val, x, prod, xs, AnchorMove{}, nullptr,
                                         static_cast<XT>(st->raw ? st->xscale : 1.0));
}

template <typename VT>
constexpr bool pb_value_free() {
  return std::is_same<VT, dev::Pattern>::value;
}
// The value-free FP64 anchor product takes the wide scatter order (a 128 KB
// x slice, runs of about 32 products), as the value-free FP64 runs do; with
// FP64 values the narrow order was faster.
template <typename VA>
constexpr bool pb_anchor_wide() {
  return pb_value_free<VA>();
}

// Mixed runs, one launch per iteration: in an iteration that moves the
// anchor, the anchor move and the FP64 anchor product's far products (values
// VA, FP64, layout a of superblock width kPbColsWide for value-free anchors,
// else kPbCols); otherwise the delta product's (values VD, delta XD, sums AD)
// over layout ad of superblock width DCOLS (kPbColsWide: the 4-byte x slice
// fills the LDS of a narrow FP64 slice).
template <typename VA, typename VD, typename XD, typename AD, int DCOLS>
__global__ __launch_bounds__(kPbScatterThreads) void pb_scatter_mx(PbScatterArgs a, PbScatterArgs ad,
                                                                   const VA *__restrict__ vala,
                                                                   const VD *__restrict__ vald,
                                                                   const XD *__restrict__ xd, void *prod,
                                                                   AnchorMove am, const PrState *st) {
  constexpr int ACOLS = pb_anchor_wide<VA>() ? kPbColsWide : kPbCols;
  static_assert(sizeof(XD) * DCOLS <= sizeof(double) * ACOLS, "the delta slice fits the LDS");
  if (st->done) return;
  __shared__ double xs[ACOLS];
  __shared__ double red[48];
  if (st->need_anchor) {
    AnchorMove m = am;
    m.first = st->anchors == 1;  // (the count includes the move this iteration makes)
    pb_scatter_sb<VA, double, double, true, ACOLS>(a, vala, nullptr, static_cast<double *>(prod), xs, m, red);
  } else {
    pb_scatter_sb<VD, XD, AD, false, DCOLS>(ad, vald, xd, static_cast<AD *>(prod), reinterpret_cast<XD *>(xs), am,
                                            nullptr);
  }
}

struct PbView {
  const int *bp, *lp;
  const unsigned short *frow, *lkey, *lbase;
  int n;
};

// Two products of type T in one vector load.
template <typename T>
struct PbPair;
template <>
struct PbPair<float> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct PbPair<double> {
  typedef double type __attribute__((ext_vector_type(2)));
};

// The row sums of one bin (blockIdx) into the FP64 LDS accumulator ybin
// (kPbRows entries): the bin's far products and its local products, x read
// from the bin's own slice (XCS: x[j] scaled by xcs[j], the value-free anchor
// product's input colscale .* xhat; else by xsc).  Ends with a barrier.
template <typename VT, typename XT, typename AT, bool XCS = false>
__device__ inline void pb_accumulate(const PbView &v, const VT *__restrict__ lval, const XT *__restrict__ x,
                                     const AT *__restrict__ prod, double *ybin, XT xsc = XT(1),
                                     const double *__restrict__ xcs = nullptr) {
#pragma clang fp contract(off)
  const int bin = blockIdx.x, tid = threadIdx.x;
  const int base = bin << kPbShift;
  for (int i = tid; i < kPbRows; i += 1024) ybin[i] = 0.0;
  __syncthreads();
  const int f1 = v.bp[bin + 1], l1 = v.lp[bin + 1];
  // far products in pairs: one 4-byte load of two rows and one 8- or
  // 16-byte load of two products, four pairs per thread per round, 1024
  // pairs apart (an odd first or last entry of the bin's run goes alone).
  // Rounds with all four pairs in range run unmasked as a software
  // pipeline: the next round's loads are issued before this round's LDS
  // additions (the compiler does not move global loads across LDS atomics);
  // one masked round ends the run
  int f0 = v.bp[bin], fe = f1;
  if ((f0 & 1) && f0 < fe) {
    if (tid == 0) atomicAdd(&ybin[v.frow[f0]], static_cast<double>(prod[f0]));
    f0++;
  }
  if (((fe - f0) & 1) && fe > f0) {
    if (tid == 1023) atomicAdd(&ybin[v.frow[fe - 1]], static_cast<double>(prod[fe - 1]));
    fe--;
  }
  {
    using AT2 = typename PbPair<AT>::type;
    const unsigned *__restrict__ frow2 = reinterpret_cast<const unsigned *>(v.frow);
    const AT2 *__restrict__ prod2 = reinterpret_cast<const AT2 *>(prod);
    const int p1 = fe >> 1;
    constexpr int U = 4;
    int kp = (f0 >> 1) + tid;
    auto add = [&](unsigned rr, AT2 pp) {
      atomicAdd(&ybin[rr & 0xFFFFu], static_cast<double>(pp.x));
      atomicAdd(&ybin[rr >> 16], static_cast<double>(pp.y));
    };
    if (kp + (U - 1) * 1024 < p1) {
      unsigned rA[U];
      AT2 pA[U];
#pragma unroll
      for (int u = 0; u < U; u++) rA[u] = frow2[kp + u * 1024], pA[u] = prod2[kp + u * 1024];
      for (;;) {
        const int kn = kp + U * 1024;
        const bool more = kn + (U - 1) * 1024 < p1;
        unsigned rB[U];
        AT2 pB[U];
        if (more)
#pragma unroll
          for (int u = 0; u < U; u++) rB[u] = frow2[kn + u * 1024], pB[u] = prod2[kn + u * 1024];
#pragma unroll
        for (int u = 0; u < U; u++) add(rA[u], pA[u]);
        kp = kn;
        if (!more) break;
#pragma unroll
        for (int u = 0; u < U; u++) rA[u] = rB[u], pA[u] = pB[u];
      }
    }
    if (kp < p1) {
      unsigned r[U];
      AT2 pv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int kk = kp + u * 1024;
        r[u] = kk < p1 ? frow2[kk] : 0u;
        pv[u] = kk < p1 ? prod2[kk] : AT2{AT(0), AT(0)};
      }
#pragma unroll
      for (int u = 0; u < U; u++)
        if (kp + u * 1024 < p1) add(r[u], pv[u]);
    }
  }
  // local entries: one group of four per thread (a 2-byte base row, one
  // 8-byte key load, aligned value loads), rows decoded from the keys; runs
  // of one row summed and added once
  constexpr int LU = 2;  // groups per thread in flight (4 measured slower)
  constexpr int kPad = -1;
  for (int k = v.lp[bin] + kPbLocalGroup * tid; k < l1; k += LU * kPbLocalGroup * 1024) {
    int r[LU][4], c[LU][4];
    AT lv[LU][4], xv[LU][4];
#pragma unroll
    for (int g = 0; g < LU; g++) {
      const int kg = k + g * kPbLocalGroup * 1024;
      if (g == 0 || kg < l1) {
        const int rb = v.lbase[kg / kPbLocalGroup];
        const dev::v4us kk = *reinterpret_cast<const dev::v4us *>(v.lkey + kg);
        PbVal<VT>::template get4<AT>(lval, kg, lv[g]);
        const unsigned key[4] = {kk.x, kk.y, kk.z, kk.w};
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int dr = static_cast<int>(key[u] >> kPbShift);
          r[g][u] = dr == kPbLocalPad ? kPad : rb + dr;
          c[g][u] = static_cast<int>(key[u] & static_cast<unsigned>(kPbRows - 1));
        }
      } else {  // past the bin's run: a group of padding
#pragma unroll
        for (int u = 0; u < 4; u++) r[g][u] = kPad, c[g][u] = 0, lv[g][u] = AT(0);
      }
    }
#pragma unroll
    for (int g = 0; g < LU; g++)
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int j = base + c[g][u];
        xv[g][u] = static_cast<AT>(XCS ? xcs[j] * x[j] : x[j] * xsc);
      }
#pragma unroll
    for (int g = 0; g < LU; g++) {
      double acc = 0.0;
      int row = r[g][0];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (r[g][u] == kPad) break;  // padding closes a group
        if (r[g][u] != row) {
          atomicAdd(&ybin[row], acc);
          acc = 0.0;
          row = r[g][u];
        }
        acc += static_cast<double>(PbVal<VT>::template mul<AT>(lv[g][u], xv[g][u]));
      }
      if (row != kPad) atomicAdd(&ybin[row], acc);
    }
  }
  __syncthreads();
}

// FP64 runs under the reference rule: b = beta*(A x) + t into out, block
// max into the slots (normalised by a separate pass).
template <typename VT, typename XT, typename AT, int TM>
__global__ __launch_bounds__(1024) void pb_gather(PbView v, const VT *__restrict__ lval, const XT *__restrict__ x,
                                                  const AT *__restrict__ prod, Teleport tp, double *__restrict__ out,
                                                  PrState *st, unsigned long long *slots) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double ybin[kPbRows];  // FP64 sums (products may be FP32)
  __shared__ double red[16];
  pb_accumulate<VT, XT, AT>(v, lval, x, prod, ybin, static_cast<XT>(st->raw ? st->xscale : 1.0));
  const int bin = blockIdx.x, tid = threadIdx.x;
  const int base = bin << kPbShift;
  const int rows = min(kPbRows, v.n - base);
  double lmax = -INFINITY;
  const double beta = st->beta, tele = st->tele;
  for (int i = tid; i < rows; i += 1024) {
    const int r = base + i;
    const double bi = add_teleport<TM>(beta * ybin[i], tele, tp, r);
    out[r] = bi;
    lmax = fmax(lmax, bi);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) lmax = fmax(lmax, __shfl_xor(lmax, off, 64));
  if ((tid & 63) == 0) red[tid >> 6] = lmax;
  __syncthreads();
  if (tid == 0) {
    double m = red[0];
    for (int w = 1; w < 16; w++) m = fmax(m, red[w]);
    atomicMax(slots + (bin % kMaxSlots), ord_encode(m));
  }
}

// FP64 runs under the sum-normalised rule, update and normalisation in one
// pass: x' = (beta*(A x) + t) * inv_sum, where inv_sum = 1 / sum(x) (the
// sum-normalised iterate does not depend on the scale, so the scale of the
// last iterate serves), with sum|x' - x| and the compensated sum of x' as
// per-bin partials.  PAT: also z' = colscale .* x' for the next value-free
// product (x is then the product's input z, xprev the iterate).
template <typename VT, typename XT, typename AT, int TM, bool PAT>
__global__ __launch_bounds__(1024) void pb_gather_sn(PbView v, const VT *__restrict__ lval, const XT *__restrict__ x,
                                                     const AT *__restrict__ prod, Teleport tp,
                                                     const double *__restrict__ xprev, double *__restrict__ xout,
                                                     double *__restrict__ zout, const double *__restrict__ colscale,
                                                     const PrState *st, double *__restrict__ perr,
                                                     double *__restrict__ ps, double *__restrict__ pc) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double ybin[kPbRows];
  __shared__ double red[48];
  pb_accumulate<VT, XT, AT>(v, lval, x, prod, ybin);
  const int bin = blockIdx.x, tid = threadIdx.x;
  const int base = bin << kPbShift;
  const int rows = min(kPbRows, v.n - base);
  const double beta = st->beta, tele = st->tele, inv = st->inv_sum;
  double err = 0.0;
  CSum cs{0, 0};
  for (int i = tid; i < rows; i += 1024) {
    const int r = base + i;
    const double xn = add_teleport<TM>(beta * ybin[i], tele, tp, r) * inv;
    xout[r] = xn;
    if (PAT) zout[r] = colscale[r] * xn;
    err += fabs(xn - xprev[r]);
    csum_add(cs, xn);
  }
  double e;
  CSum t;
  block1024_err_csum(err, cs, red, e, t);
  if (tid == 0) {
    perr[bin] = e;
    ps[bin] = t.s;
    pc[bin] = t.c;
  }
}

// Mixed runs (the partner of pb_scatter_mx).  The iterate is xhat + delta
// (FP64 anchor, FP32 delta) and is scaled by c = 1 / sum each iteration, as
// the FP64 runs are: x' = c*(beta*A x + t*e), e the teleport pattern.  The
// iteration that moves the anchor (xhat += delta, then delta = 0) sums the
// FP64 anchor product (values VA, input xhat; value-free anchors: acs .*
// xhat, acs = 1/count of each column, as the scatter forms it) and stores the
// anchor's scaled one-step change w = c_a*(beta*A xhat + t_a*e) - xhat in
// FP32 (c_a, t_a: that iteration's scale and teleport term); the next delta
// is d' = w.  The other iterations sum the delta product (values VD, input
// xd): with r = c / c_a,
//   d' = r*w + (r - 1)*xhat + c*(beta*(A_lo d) + (t - t_a)*e),
// which is x' - xhat.  w and r - 1 stay of the order of the iterate's
// change (the scale moves with the iterate's sum, which leaks at dangling
// columns), so the FP32 delta carries no part of the iterate itself.
// Writes d' (PAT32: also zd' = fl32(colscale32 * d')), with sum|d' - d| and
// the compensated sum of d' as per-bin partials.
template <typename VA, typename VD, typename XD, typename AD, int TM, bool PAT32>
__global__ __launch_bounds__(1024) void pb_gather_mx(PbView v, const VA *__restrict__ lvala,
                                                     const double *__restrict__ acs, const VD *__restrict__ lvald,
                                                     const XD *__restrict__ xd, const void *prod, Teleport tp,
                                                     const double *__restrict__ xhat,
                                                     const float *__restrict__ xhat32, float *__restrict__ w,
                                                     const float *__restrict__ dold, float *__restrict__ dnew,
                                                     float *__restrict__ zdnew, const float *__restrict__ colscale32,
                                                     const PrState *st, double *__restrict__ perr,
                                                     double *__restrict__ ps, double *__restrict__ pc) {
#pragma clang fp contract(off)
  if (st->done) return;
  __shared__ double ybin[kPbRows];
  __shared__ double red[48];
  const bool anchor = st->need_anchor != 0;
  if (anchor)
    pb_accumulate<VA, double, double, pb_value_free<VA>()>(v, lvala, xhat, static_cast<const double *>(prod), ybin,
                                                          1.0, acs);
  else
    pb_accumulate<VD, XD, AD>(v, lvald, xd, static_cast<const AD *>(prod), ybin);
  const int bin = blockIdx.x, tid = threadIdx.x;
  const int base = bin << kPbShift;
  const int rows = min(kPbRows, v.n - base);
  const double beta = st->beta, c = st->inv_sum;
  // the anchor iteration's teleport term and scale (kept by the finalize
  // pass of that iteration for the phase)
  const double ta = anchor ? st->tele : st->t_anchor;
  const double r = anchor ? 1.0 : c / st->c_anchor, r1 = r - 1.0;
  const double dt = st->tele - ta;
  double err = 0.0;
  CSum cs{0, 0};
  for (int i = tid; i < rows; i += 1024) {
    const int row = base + i;
    float d32;
    if (anchor) {
      d32 = static_cast<float>(c * add_teleport<TM>(beta * ybin[i], ta, tp, row) - xhat[row]);
      w[row] = d32;
      err += fabs(static_cast<double>(d32));
    } else {
      // (r - 1) is of the order of the change of the leak since the anchor,
      // so the FP32 anchor copy carries this term to well below FP32 delta
      // precision
      d32 = static_cast<float>(r * static_cast<double>(w[row]) + r1 * static_cast<double>(xhat32[row]) +
                               c * add_teleport<TM>(beta * ybin[i], dt, tp, row));
      err += fabs(static_cast<double>(d32) - static_cast<double>(dold[row]));
    }
    dnew[row] = d32;
    if (PAT32) zdnew[row] = static_cast<float>(static_cast<double>(colscale32[row]) * static_cast<double>(d32));
    csum_add(cs, static_cast<double>(d32));
  }
  double e;
  CSum t;
  block1024_err_csum(err, cs, red, e, t);
  if (tid == 0) {
    perr[bin] = e;
    ps[bin] = t.s;
    pc[bin] = t.c;
  }
}

PbView pb_view(const PbMatrix &pb) {
  return PbView{pb.bp.as<int>(), pb.lp.as<int>(), pb.frow.as<unsigned short>(), pb.lkey.as<unsigned short>(),
                pb.lbase.as<unsigned short>(), pb.n};
}

PbScatterArgs pb_args(const PbMatrix &pb) {
  return PbScatterArgs{pb.sp.as<int>(),    pb.fkey.as<unsigned short>(), pb.rp.as<int>(), pb.rdelta.as<int>(),
                       pb.cbase.as<int>(), pb.cp.as<int>(),              pb.nsb,          pb.n,
                       pb.n_far};
}

// The scatter order of the delta products with values VD (the far values of
// dtype F32 are stored narrow, BF16 wide; the pb_mirror build follows this).
template <typename VD>
constexpr bool pb_delta_wide() {
  return !std::is_same<VD, float>::value;
}

PbScatterArgs pb_args_wide(const PbMatrix &pb) {
  return PbScatterArgs{pb.sp_w.as<int>(),    pb.fkey_w.as<unsigned short>(), pb.rp_w.as<int>(), pb.rdelta_w.as<int>(),
                       pb.cbase_w.as<int>(), pb.cp_w.as<int>(),              pb.nsb_w,          pb.n,
                       pb.n_far};
}

dim3 pb_scatter_grid(const PbMatrix &pb) { return dim3((pb.nsb + 7) / 8 * 8); }
dim3 pb_scatter_grid_wide(const PbMatrix &pb) { return dim3((pb.nsb_w + 7) / 8 * 8); }

// Dynamic LDS added to the scatter launches whose static LDS (a 64 KB x
// slice) would let two blocks share a CU: one scatter block per CU writes
// its products faster (2^24-node graph: FP64 scatter 345 -> 315 us, and the
// gather after it 374 -> 362 us; two blocks per CU for the mixed scatters:
// 199 -> 207 us).  The value-free FP64 scatter takes the wide order, whose
// 128 KB slice already holds a CU alone (234 -> 210 us; with values, FP64
// and FP32, the wide order measured slower).
constexpr unsigned kPbScatterLdsPad = 24 * 1024;

// The far products of one FP64-sum iteration (values VT, x of type XT):
// value-free ones over the wide order, the others over the narrow one.
template <typename VT, typename XT, typename AT>
void launch_pb_scatter(const PbMatrix &pb, const VT *fval, const XT *x, AT *prod, const PrState *st, hipStream_t s) {
  if (pb.n_far == 0) return;
  if constexpr (std::is_same<VT, dev::Pattern>::value && sizeof(XT) == 8)
    hipLaunchKernelGGL((pb_scatter<VT, XT, AT, kPbColsWide>), pb_scatter_grid_wide(pb), dim3(kPbScatterThreads), 0, s,
                       pb_args_wide(pb), fval, x, prod, st);
  else
    hipLaunchKernelGGL((pb_scatter<VT, XT, AT, kPbCols>), pb_scatter_grid(pb), dim3(kPbScatterThreads),
                       kPbScatterLdsPad, s, pb_args(pb), fval, x, prod, st);
}

// One propagation-blocked product with the reference-rule update (FP64
// runs): values of type VT from `pb` (dev::Pattern: value-free), x of type
// XT, sums of type AT.
template <typename VT, typename XT, typename AT, int TM>
void launch_pb(const PbMatrix &pb, int vslot, const XT *x, const Teleport &tp, double *out, PrState *st,
               unsigned long long *slots, hipStream_t s) {
  const VT *fval = vslot >= 0 ? pb.fval[vslot].as<VT>() : nullptr;
  const VT *lval = vslot >= 0 ? pb.lval[vslot].as<VT>() : nullptr;
  AT *prod = pb.prod.as<AT>();
  launch_pb_scatter<VT, XT, AT>(pb, fval, x, prod, st, s);
  hipLaunchKernelGGL((pb_gather<VT, XT, AT, TM>), dim3(pb.nbins), dim3(1024), 0, s, pb_view(pb), lval, x, prod, tp,
                     out, st, slots);
}

template <typename VT, typename XT, typename AT>
void launch_pb_update(int tmode, const PbMatrix &pb, int vslot, const XT *x, const Teleport &tp, double *out,
                      PrState *st, unsigned long long *slots, hipStream_t s) {
  if (tmode == 1)
    launch_pb<VT, XT, AT, 1>(pb, vslot, x, tp, out, st, slots, s);
  else if (tmode == 2)
    launch_pb<VT, XT, AT, 2>(pb, vslot, x, tp, out, st, slots, s);
  else
    launch_pb<VT, XT, AT, 0>(pb, vslot, x, tp, out, st, slots, s);
}

// The per-bin partials of a sum-normalised iteration.
struct SnPartials {
  double *perr, *ps, *pc;
};

// One sum-normalised FP64 iteration (PAT: value-free over z): scatter and
// gather; the finalize pass follows.
template <typename VT, int TM, bool PAT>
void launch_pb_sn_tm(const PbMatrix &pb, int vslot, const double *xin, const double *z, const Teleport &tp,
                     double *xout, double *zout, const double *colscale, PrState *st, const SnPartials &sp,
                     hipStream_t s) {
  const VT *fval = vslot >= 0 ? pb.fval[vslot].as<VT>() : nullptr;
  const VT *lval = vslot >= 0 ? pb.lval[vslot].as<VT>() : nullptr;
  const double *x = PAT ? z : xin;
  double *prod = pb.prod.as<double>();
  launch_pb_scatter<VT, double, double>(pb, fval, x, prod, st, s);
  hipLaunchKernelGGL((pb_gather_sn<VT, double, double, TM, PAT>), dim3(pb.nbins), dim3(1024), 0, s, pb_view(pb), lval,
                     x, prod, tp, xin, xout, zout, colscale, st, sp.perr, sp.ps, sp.pc);
}

template <typename VT, bool PAT>
void launch_pb_sn(int tmode, const PbMatrix &pb, int vslot, const double *xin, const double *z, const Teleport &tp,
                  double *xout, double *zout, const double *colscale, PrState *st, const SnPartials &sp,
                  hipStream_t s) {
  if (tmode == 1)
    launch_pb_sn_tm<VT, 1, PAT>(pb, vslot, xin, z, tp, xout, zout, colscale, st, sp, s);
  else if (tmode == 2)
    launch_pb_sn_tm<VT, 2, PAT>(pb, vslot, xin, z, tp, xout, zout, colscale, st, sp, s);
  else
    launch_pb_sn_tm<VT, 0, PAT>(pb, vslot, xin, z, tp, xout, zout, colscale, st, sp, s);
}

// Buffers of a mixed sum-normalised iteration.
struct MxBufs {
  AnchorMove am;           // anchor xhat (and 1/count, value-free) with the scatter's sum partials
  const void *xd;          // the delta product's input: delta (F32/BF16) or zd (PATTERN_F32)
  float *w;                // FP32 one-step change of the anchor
  const float *dold;       // delta of the current iterate
  float *dnew, *zdnew;     // delta (and zd) of the next
  const float *colscale32; // PATTERN_F32
};

// One mixed iteration: the anchor move and the FP64 anchor product (values
// VA at slot va, input z or xhat) when the state asks for it, else the delta
// product (values VD at slot vd); the finalize pass follows.
template <typename VA, typename VD, typename XD, int TM, bool PAT32>
void launch_pb_mx_tm(const PbMatrix &pb, int va, int vd, const MxBufs &b, const Teleport &tp, PrState *st,
                     const SnPartials &sp, hipStream_t s) {
  const VA *fa = va >= 0 ? pb.fval[va].as<VA>() : nullptr;
  const VA *la = va >= 0 ? pb.lval[va].as<VA>() : nullptr;
  const VD *fd = vd >= 0 ? pb.fval[vd].as<VD>() : nullptr;
  const VD *ld = vd >= 0 ? pb.lval[vd].as<VD>() : nullptr;
  const XD *xd = static_cast<const XD *>(b.xd);
  // the delta products of BF16 and value-free runs take the wide scatter
  // order, FP32-valued ones the narrow one (measured: value-free FP32 delta
  // scatter 154 -> 140 us wide, BF16 196 -> 192, FP32 191 -> 196)
  constexpr bool wide = pb_delta_wide<VD>();
  // (a wide anchor's 128 KB slice leaves room for one block per CU without
  // the LDS pad)
  constexpr bool awide = pb_anchor_wide<VA>();
  hipLaunchKernelGGL((pb_scatter_mx<VA, VD, XD, float, wide ? kPbColsWide : kPbCols>), pb_scatter_grid(pb),
                     dim3(kPbScatterThreads), awide ? 0 : kPbScatterLdsPad, s, awide ? pb_args_wide(pb) : pb_args(pb),
                     wide ? pb_args_wide(pb) : pb_args(pb), fa, fd, xd, pb.prod.get(), b.am, st);
  hipLaunchKernelGGL((pb_gather_mx<VA, VD, XD, float, TM, PAT32>), dim3(pb.nbins), dim3(1024), 0, s, pb_view(pb), la,
                     b.am.colscale, ld, xd, pb.prod.get(), tp, b.am.xhat, b.am.xhat32, b.w, b.dold, b.dnew, b.zdnew, b.colscale32, st,
                     sp.perr, sp.ps, sp.pc);
}

template <typename VA, typename VD, typename XD, bool PAT32>
void launch_pb_mx(int tmode, const PbMatrix &pb, int va, int vd, const MxBufs &b, const Teleport &tp, PrState *st,
                  const SnPartials &sp, hipStream_t s) {
  if (tmode == 1)
    launch_pb_mx_tm<VA, VD, XD, 1, PAT32>(pb, va, vd, b, tp, st, sp, s);
  else if (tmode == 2)
    launch_pb_mx_tm<VA, VD, XD, 2, PAT32>(pb, va, vd, b, tp, st, sp, s);
  else
    launch_pb_mx_tm<VA, VD, XD, 0, PAT32>(pb, va, vd, b, tp, st, sp, s);
}

template <typename VT, typename XT, bool MIXED, int TM>
void launch_update(const DeviceMatrix &d, const VT *vals, const XT *x, const double *yhat, const Teleport &tp,
                   double *b, PrState *st, unsigned long long *slots, hipStream_t s) {
  dev::CsrView<VT> m = dev::csr_view<VT>(d);
  if (vals) m.vals = vals;
  if (d.ntiles > 0)
    hipLaunchKernelGGL((pr_update_csr<VT, XT, MIXED, TM, false>), dim3(d.ntiles), dim3(kTileThreads), 0, s, m, x,
                       yhat, tp, b, st, slots);
  if (d.nlong > 0)
    hipLaunchKernelGGL((pr_update_csr<VT, XT, MIXED, TM, true>), dim3(d.nlong), dim3(kTileThreads), 0, s, m, x, yhat,
                       tp, b, st, slots);
}

// Values override `vals` (nullptr: the mirror's own values).
template <typename VT, typename XT, bool MIXED>
void launch_update(int tmode, const DeviceMatrix &d, const VT *vals, const XT *x, const double *yhat,
                   const Teleport &tp, double *b, PrState *st, unsigned long long *slots, hipStream_t s) {
  if (tmode == 1)
    launch_update<VT, XT, MIXED, 1>(d, vals, x, yhat, tp, b, st, slots, s);
  else if (tmode == 2)
    launch_update<VT, XT, MIXED, 2>(d, vals, x, yhat, tp, b, st, slots, s);
  else
    launch_update<VT, XT, MIXED, 0>(d, vals, x, yhat, tp, b, st, slots, s);
}

template <typename VT>
void launch_anchor(const DeviceMatrix &d, const double *x, double *yhat, const PrState *st, hipStream_t s) {
  const dev::CsrView<VT> m = dev::csr_view<VT>(d);
  if (d.ntiles > 0)
    hipLaunchKernelGGL((pr_anchor_spmv<VT, false>), dim3(d.ntiles), dim3(kTileThreads), 0, s, m, x, yhat, st);
  if (d.nlong > 0)
    hipLaunchKernelGGL((pr_anchor_spmv<VT, true>), dim3(d.nlong), dim3(kTileThreads), 0, s, m, x, yhat, st);
}

__global__ void zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// One device buffer of a run: owned (allocated for the run) or a slice of
// the fixed scratch buffer of the allocation-free entry point.  Unlike
// DevBuf, a slice is not cleared: every buffer below is written before it
// is read.
struct WsBuf {
  DevBuf own;
  void *p = nullptr;
  template <class T>
  T *as() const {
    return static_cast<T *>(p);
  }
  void *get() const { return p; }
};

// Where a run's buffers come from: hipMalloc (arena == nullptr) or
// consecutive 256-byte-aligned slices of `arena` (cap bytes), which is
// never allocated or freed here.
struct Workspace {
  unsigned char *arena = nullptr;
  std::size_t cap = 0, used = 0;
  void alloc(WsBuf &b, std::size_t bytes) {
    if (!arena) {
      b.own.alloc(bytes);
      b.p = b.own.get();
      return;
    }
    const std::size_t need = (std::max<std::size_t>(bytes, 1) + 255) & ~std::size_t(255);
    if (need > cap - used)
      throw std::invalid_argument("tbsla: the allocation-free PageRank needs more than its " +
                                  std::to_string(cap >> 20) + " MiB scratch buffer");
    b.p = arena + used;
    used += need;
  }
  void upload(WsBuf &b, const void *src, std::size_t bytes, hipStream_t s) {
    alloc(b, bytes);
    if (bytes) TBSLA_HIP_CHECK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
  }
};

struct Scratch {
  WsBuf st, slots, perr, ps, pc, total, member, nodes, hsum, hcomp;
  WsBuf X0, X1, Z, Z1, yhat, xhat, xhat32, D0, D1, ZD0, ZD1, W;
};

// zd = fl32(colscale32 * d) of a constant FP32 delta d
__global__ __launch_bounds__(256) void scale_into_f32_32(const float *__restrict__ s, float d, float *__restrict__ z,
                                                         long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    z[i] = static_cast<float>(static_cast<double>(s[i]) * static_cast<double>(d));
}

// One block: the step decision of a sum-normalised propagation-blocked
// iteration from the gather's per-bin partials (nb) and, after a mixed run's
// anchor move, the scatter's per-superblock sums of the new anchor (nsb).
// S is the sum of the new iterate: sum(x') (FP64 runs), sum(xhat) + sum(d')
// (mixed runs).  The iteration that moved the anchor starts from the
// rescaled anchor, whose sum replaces the previous one in the bound.
__global__ __launch_bounds__(kTileThreads) void pr_finalize_sn(const double *__restrict__ perr,
                                                               const double *__restrict__ ps,
                                                               const double *__restrict__ pc, int nb,
                                                               const double *__restrict__ hs,
                                                               const double *__restrict__ hc, int nsb, PrState *st) {
  if (st->done) return;
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  double e = 0;
  CSum t{0, 0};
  sum_partials(perr, ps, pc, nb, e, t);
  e = dev::block_tree_sum<double>(e, red_s);
  t = block_csum(t, red_s, red_c);
  const bool moved = st->mixed && st->need_anchor;
  CSum h{0, 0};
  if (moved) {
    double unused = 0;
    sum_partials(nullptr, hs, hc, nsb, unused, h);
    h = block_csum(h, red_s, red_c);
  }
  if (threadIdx.x == 0) {
    double S = t.s + t.c;
    if (st->mixed) {
      if (moved) {
        st->sum_hat = h.s + h.c;
        st->sum = st->sum_hat;  // the iterate the step started from: the moved anchor
        st->t_anchor = st->tele;
        st->c_anchor = st->inv_sum;
      }
      S += st->sum_hat;
    }
    finalize_step(st, e, S);
    st->inv_sum = 1.0 / S;
  }
}

// member[i] = 1 for the listed nodes in [0, n) (member cleared first).
__global__ __launch_bounds__(256) void set_members(unsigned char *member, long n, const int *nodes, int n_nodes) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n_nodes; i += static_cast<long>(gridDim.x) * 256) {
    const int v = nodes[i];
    if (v >= 0 && v < n) member[v] = 1;
  }
}
__global__ __launch_bounds__(256) void fill_u8(unsigned char *y, long n, unsigned char v) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) y[i] = v;
}

}  // namespace

// Iterations to launch before the next state read: the iterations the last
// two residuals predict are left (geometric decay of the quantity the rule
// tests), at least 1 and at most `check`.  Kernels launched past convergence
// return at once, so this only trims those launches.
int next_batch(const PrState &h, int check) {
  const bool sumnorm = h.rule != 0;
  const double cur = sumnorm ? h.bound : h.err, prev = sumnorm ? h.bound_prev : h.err_prev;
  const double target = sumnorm ? h.bound_target : h.eps;
  if (!(cur > 0.0) || !(prev > cur) || !(target > 0.0) || cur <= target) return check;
  const double left = std::log(target / cur) / std::log(cur / prev);
  if (!(left > 0.0)) return check;
  return static_cast<int>(std::min<double>(check, std::max(1.0, std::ceil(left))));
}

namespace {

// Iterations the state h predicts are left (as next_batch, unclamped): 1
// once the bound is met (a certification step of a mixed run may follow),
// `check` when the last two residuals give no decay to extrapolate.
int iterations_left(const PrState &h, int check) {
  const bool sumnorm = h.rule != 0;
  const double cur = sumnorm ? h.bound : h.err, prev = sumnorm ? h.bound_prev : h.err_prev;
  const double target = sumnorm ? h.bound_target : h.eps;
  if (cur <= target) return 1;
  if (!(cur > 0.0) || !(prev > cur) || !(target > 0.0)) return check;
  const double left = std::log(target / cur) / std::log(cur / prev);
  if (!(left > 0.0) || !(left < 1e6)) return check;
  return static_cast<int>(std::ceil(left));
}

}  // namespace

namespace {

// The engine.  arena != nullptr: every buffer of the run is a slice of the
// arena (cap bytes) and nothing is allocated (the caller has checked that
// the mirror exists and that no lazily built layout is needed).
PageRankReport run_page_rank(const cpp::Matrix &m, const PageRankParams &p, double *out_dev, hipStream_t s,
                             unsigned char *arena, std::size_t cap) {
  const long n = m.get_n_col();
  // (COO keeps NR/NC in ln_row/ln_col, MatrixCOO.cpp:344-345, so the block
  // layout is read from NR/NC)
  if (m.get_n_row() != m.get_n_col() || m.get_NR() != 1 || m.get_NC() != 1)
    throw std::invalid_argument("tbsla: page_rank needs a square matrix held in one block");
  PageRankReport rep;
  if (n == 0) return rep;
  const bool pat32 = p.precision == PrPrecision::PATTERN_F32;
  const bool mixed = p.precision == PrPrecision::F32 || p.precision == PrPrecision::BF16 || pat32;
  const bool csr = m.format() == cpp::FORMAT_CSR;
  if ((mixed || p.precision == PrPrecision::PATTERN) && !csr)
    throw std::invalid_argument("tbsla: mixed-precision and value-free PageRank need a CSR matrix");
  if ((p.precision == PrPrecision::PATTERN || pat32) && !m.count_stochastic())
    throw std::invalid_argument("tbsla: value-free PageRank needs values from fill_cqmat_stochastic");
  if (mixed && p.rule != PrRule::SUMNORM)
    throw std::invalid_argument("tbsla: mixed-precision PageRank uses the SUMNORM stopping rule");
  if (p.values && (!csr || p.precision != PrPrecision::F64))
    throw std::invalid_argument("tbsla: explicit edge values need a CSR matrix and FP64 precision");

  const DeviceMatrix &d64 = mirror(m, Dtype::F64);
  const bool fused = csr && !d64.noop && d64.rows == n;
  // propagation-blocked products: explicit choice, or by default on graphs
  // of 2^16 nodes and more, whatever the stopping rule (their row sums are
  // rounded in a run-dependent order; gather = 1 keeps the fixed-order tile
  // products, whose sums are the reference's)
  const bool use_pb = fused && !p.values && (p.gather == 2 || (p.gather == 0 && n >= (1L << 16)));
  const PbMatrix *pbm = nullptr;
  if (use_pb) {
    const bool mixed_pat = mixed && m.count_stochastic();
    pbm = &pb_mirror(m, (p.precision == PrPrecision::PATTERN || mixed_pat) ? Dtype::PATTERN : Dtype::F64);
    if (p.precision == PrPrecision::F32) pbm = &pb_mirror(m, Dtype::F32);
    if (p.precision == PrPrecision::BF16) pbm = &pb_mirror(m, Dtype::BF16);
  }
  // the tile products of a mixed run read the values in the low precision
  // (the blocked ones read them from the blocked layout)
  const DeviceMatrix *dlo = nullptr;
  if (!use_pb && p.precision == PrPrecision::F32) dlo = &mirror(m, Dtype::F32);
  if (!use_pb && p.precision == PrPrecision::BF16) dlo = &mirror(m, Dtype::BF16);
  if (csr && !d64.noop && d64.rows != n)
    throw std::invalid_argument("tbsla: CSR rowptr does not cover every row");
  if (d64.x_need > n) throw std::invalid_argument("tbsla: column index beyond n_col");

  Scratch w;
  Workspace ws{arena, cap, 0};
  const unsigned gN = grid_for(n);
  // the mixed normalise pass (six streams) measured faster with twice the
  // blocks (four and eight times: slower); the FP64 one (three streams) slower
  const unsigned gM = mixed ? grid_for(n, 2L * kNormBlocks) : gN;
  // sum-normalised rule with propagation-blocked products: the update, the
  // normalisation and the residual in the gather pass, one finalize pass
  const bool sn = pbm && p.rule == PrRule::SUMNORM;
  const unsigned nparts = std::max(std::max(gN, gM), sn ? static_cast<unsigned>(pbm->nbins) : 0u);
  ws.alloc(w.st, sizeof(PrState));
  ws.alloc(w.slots, sizeof(unsigned long long) * kMaxSlots);
  ws.alloc(w.perr, sizeof(double) * nparts);
  ws.alloc(w.ps, sizeof(double) * nparts);
  ws.alloc(w.pc, sizeof(double) * nparts);
  ws.alloc(w.total, sizeof(double));
  ws.alloc(w.X0, sizeof(double) * n);
  ws.alloc(w.X1, sizeof(double) * n);
  // teleport mode: 0 uniform, 1 node set (Matrix.cpp:63-114), 2 distribution vector
  const int tmode = p.teleport ? 2 : (p.personalized ? 1 : 0);
  Teleport tp{nullptr, p.teleport};
  const double b0 = p.personalized && !p.teleport ? 0.25 : 1.0;
  std::vector<unsigned char> mem;
  if (tmode == 1 && arena) {
    // no host staging vector: the node list goes to the scratch buffer and
    // a kernel sets the flags
    ws.alloc(w.member, static_cast<std::size_t>(n));
    hipLaunchKernelGGL(fill_u8, dim3(grid_for(n)), dim3(256), 0, s, w.member.as<unsigned char>(), n,
                       static_cast<unsigned char>(0));
    if (p.n_nodes > 0) {
      ws.upload(w.nodes, p.nodes, sizeof(int) * static_cast<std::size_t>(p.n_nodes), s);
      hipLaunchKernelGGL(set_members, dim3(grid_for(p.n_nodes)), dim3(256), 0, s, w.member.as<unsigned char>(), n,
                         w.nodes.as<int>(), p.n_nodes);
    }
    tp.member = w.member.as<unsigned char>();
  } else if (tmode == 1) {
    mem.assign(static_cast<std::size_t>(n), 0);
    for (int k = 0; k < p.n_nodes; k++)
      if (p.nodes[k] >= 0 && p.nodes[k] < n) mem[static_cast<std::size_t>(p.nodes[k])] = 1;
    ws.upload(w.member, mem.data(), mem.size(), s);
    tp.member = w.member.as<unsigned char>();
  }
  const bool pattern = p.precision == PrPrecision::PATTERN;
  // 1/count per column (the value every entry of column j holds after
  // fill_cqmat_stochastic), computed from the FP64 mirror on the GPU once
  // per matrix and kept in its device cache
  const double *colscale64 = nullptr;
  const float *colscale32 = nullptr;
  if (pattern || pat32 || (mixed && m.count_stochastic())) {
    if (arena) throw std::invalid_argument("tbsla: the allocation-free entry point runs FP64 only");
    DeviceCache &cache = m.device_cache();
    std::lock_guard<std::mutex> g(cache.mu);
    if (!cache.colscale64.get()) {
      DevBuf c64, c32;
      c64.alloc(sizeof(double) * static_cast<std::size_t>(n));
      c32.alloc(sizeof(float) * static_cast<std::size_t>(n));
      hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, c64.as<double>(), n, 0.0);
      if (d64.entries > 0)
        hipLaunchKernelGGL(colscale_kernel, dim3(grid_for(d64.entries)), dim3(256), 0, s, d64.colidx.as<int>(),
                           d64.vals.as<double>(), d64.entries, d64.f_col, n, c64.as<double>());
      hipLaunchKernelGGL(to_f32, dim3(gN), dim3(256), 0, s, c64.as<double>(), c32.as<float>(), n);
      TBSLA_HIP_CHECK(hipGetLastError());
      TBSLA_HIP_CHECK(hipStreamSynchronize(s));
      cache.colscale64 = std::move(c64);
      cache.colscale32 = std::move(c32);
    }
    colscale64 = cache.colscale64.as<double>();
    colscale32 = cache.colscale32.as<float>();
  }
  PrState h{};
  h.beta = p.beta;
  h.eps = p.epsilon;
  h.bound_target = p.epsilon * (1.0 - p.beta) / p.beta;
  h.tele_scale = tmode == 2   ? (1.0 - p.beta)
                 : tmode == 1 ? (1.0 - p.beta) / static_cast<double>(static_cast<std::size_t>(p.n_nodes))
                              : (1.0 - p.beta) / static_cast<double>(n);
  // sum-normalised propagation-blocked runs start from 1/n (any uniform
  // start gives the same sum-normalised iterates; this one keeps the
  // iterates near sum 1 from the start)
  const double x0 = sn ? 1.0 / static_cast<double>(n) : b0;
  const double x0s = mixed ? static_cast<double>(static_cast<float>(x0)) : x0;  // as stored
  h.sum = x0s * static_cast<double>(n);
  h.tele = h.sum * h.tele_scale;
  h.inv_sum = 1.0 / h.sum;
  h.sum_hat = 0.0;
  h.t_anchor = 0.0;
  h.c_anchor = 1.0;
  h.bound = 2.0;
  h.bound_prev = 2.0;
  h.bound_anchor = 2.0;
  h.max_it = p.max_iteration;
  h.rule = static_cast<int>(p.rule);
  h.mixed = mixed ? 1 : 0;
  // FP64 reference-rule runs on propagation-blocked products keep each
  // update's b unnormalised and carry its scale (1/max) to the readers of
  // the iterate, so the normalise pass writes nothing (a pass over the
  // iterate less per iteration)
  h.raw = (fused && pbm && !sn && !mixed && !pattern) ? 1 : 0;
  h.xscale = 1.0;
  // mixed runs start from the anchor xhat = 0 (so yhat = A xhat = 0 with no
  // product) and delta = b0: the first phase carries the whole iterate in the
  // low precision, and the first anchor move comes when it has converged to
  // that precision's floor
  h.need_anchor = 0;
  h.anchors = 0;
  h.floor_eps = p.precision == PrPrecision::BF16 ? 4e-3 : 6e-8;  // FP32 and PATTERN_F32: 6e-8
  // (a sweep on fill_cqmat_stochastic(2^24,2^24,8,0.5,7): BF16 5e-2 -> 7
  // anchor moves, 2e-2 -> 6, 1e-2 -> 5 and the fastest, 5e-3 -> 5 with one
  // more iteration; FP32 3e-6 and 1e-6 -> 1 move, 1e-5 and above -> 2)
  h.anchor_tau = p.anchor_tau > 0 ? p.anchor_tau : (p.precision == PrPrecision::BF16 ? 1e-2 : 3e-6);
  h.done = p.max_iteration <= 0 ? 1 : 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(w.st.get(), &h, sizeof(h), hipMemcpyHostToDevice, s));
  PrState *st = w.st.as<PrState>();
  double *X[2] = {w.X0.as<double>(), w.X1.as<double>()};
  // (mixed runs on propagation-blocked products keep the iterate as the
  // anchor plus the delta and never read X)
  const bool pb_mixed = mixed && sn;
  if (!pb_mixed) {
    hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, X[0], n, x0);
    hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, X[1], n, 0.0);
  }
  unsigned long long *slots = w.slots.as<unsigned long long>();
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(kMaxSlots), 0, s, slots, kMaxSlots);
  float *D[2] = {nullptr, nullptr};
  if (mixed) {
    ws.alloc(w.xhat, sizeof(double) * n);
    ws.alloc(w.D0, sizeof(float) * n);
    ws.alloc(w.D1, sizeof(float) * n);
    D[0] = w.D0.as<float>();
    D[1] = w.D1.as<float>();
    hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, w.xhat.as<double>(), n, 0.0);
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, D[0], n, static_cast<float>(x0));
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, D[1], n, 0.0f);
    if (sn) {
      // the anchor's FP32 one-step change (0 before the first anchor, whose
      // anchor is 0) and the scatter's per-superblock sums of a moved anchor
      ws.alloc(w.W, sizeof(float) * n);
      hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, w.W.as<float>(), n, 0.0f);
      ws.alloc(w.xhat32, sizeof(float) * n);
      hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, w.xhat32.as<float>(), n, 0.0f);
      ws.alloc(w.hsum, sizeof(double) * std::max(1, pbm->nsb));
      ws.alloc(w.hcomp, sizeof(double) * std::max(1, pbm->nsb));
    } else {
      ws.alloc(w.yhat, sizeof(double) * n);
      hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, w.yhat.as<double>(), n, 0.0);
    }
  }
  float *ZD[2] = {nullptr, nullptr};
  if (pat32) {
    ws.alloc(w.ZD0, sizeof(float) * n);
    ws.alloc(w.ZD1, sizeof(float) * n);
    ZD[0] = w.ZD0.as<float>();
    ZD[1] = w.ZD1.as<float>();
    if (sn)
      hipLaunchKernelGGL(scale_into_f32_32, dim3(gN), dim3(256), 0, s, colscale32,
                         static_cast<float>(x0), ZD[0], n);
    else
      hipLaunchKernelGGL(scale_into_f32, dim3(gN), dim3(256), 0, s, colscale64, b0, ZD[0], n);
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, ZD[1], n, 0.0f);
  }
  // (propagation-blocked mixed runs form z = colscale .* xhat where they
  // read it)
  if (pattern || (mixed && m.count_stochastic() && !pb_mixed)) {
    ws.alloc(w.Z, sizeof(double) * n);
    hipLaunchKernelGGL(scale_into, dim3(gN), dim3(256), 0, s, colscale64, X[0], w.Z.as<double>(), n);
  }
  // value-free FP64 sum-normalised runs write the next z beside the next
  // iterate, so z alternates between two buffers like the iterate
  double *ZP[2] = {w.Z.as<double>(), nullptr};
  if (sn && pattern) {
    ws.alloc(w.Z1, sizeof(double) * n);
    ZP[1] = w.Z1.as<double>();
  }
  TBSLA_HIP_CHECK(hipGetLastError());
  // the set-up copies read pageable host memory (state, node set, 1/count)
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));

  // device time of the loop: HIP events, or (allocation-free runs, which
  // create no events) the host clock between two stream synchronisations
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const auto h0 = std::chrono::steady_clock::now();
  if (!arena) {
    TBSLA_HIP_CHECK(hipEventCreate(&e0));
    TBSLA_HIP_CHECK(hipEventCreate(&e1));
    TBSLA_HIP_CHECK(hipEventRecord(e0, s));
  }
  const int check = fused ? std::max(1, p.check_every) : 1;
  int next = check;  // iterations in the next launch batch
  PrState hs{};
  int launched = 0;
  const bool mixed_pattern = mixed && m.count_stochastic();
  auto fin_max = [&]() { hipLaunchKernelGGL(pr_finalize_max, dim3(1), dim3(kMaxSlots), 0, s, slots, st); };
  auto fin_norm = [&](unsigned nb) {
    hipLaunchKernelGGL(pr_finalize_norm, dim3(1), dim3(kTileThreads), 0, s, w.perr.as<double>(), w.ps.as<double>(),
                       w.pc.as<double>(), static_cast<int>(nb), st);
  };
  double *Z = w.Z.as<double>();
  const double *colscale = colscale64;
  auto launch_batch = [&](int batch) {
    for (int k = 0; k < batch; k++) {
      const int it = launched + k;
      double *xin = X[it & 1], *xout = X[(it + 1) & 1];
      if (sn) {
        const SnPartials sp{w.perr.as<double>(), w.ps.as<double>(), w.pc.as<double>()};
        if (mixed) {
          MxBufs b;
          b.am = AnchorMove{w.xhat.as<double>(), w.xhat32.as<float>(), D[it & 1], mixed_pattern ? colscale : nullptr,
                            w.hsum.as<double>(), w.hcomp.as<double>(), 0};
          b.xd = pat32 ? static_cast<const void *>(ZD[it & 1]) : static_cast<const void *>(D[it & 1]);
          b.w = w.W.as<float>();
          b.dold = D[it & 1];
          b.dnew = D[(it + 1) & 1];
          b.zdnew = pat32 ? ZD[(it + 1) & 1] : nullptr;
          b.colscale32 = colscale32;
          if (pat32)
            launch_pb_mx<dev::Pattern, dev::Pattern, float, true>(tmode, *pbm, -1, -1, b, tp, st, sp, s);
          else if (mixed_pattern && p.precision == PrPrecision::F32)
            launch_pb_mx<dev::Pattern, float, float, false>(tmode, *pbm, -1, 1, b, tp, st, sp, s);
          else if (mixed_pattern)
            launch_pb_mx<dev::Pattern, dev::bf16, float, false>(tmode, *pbm, -1, 2, b, tp, st, sp, s);
          else if (p.precision == PrPrecision::F32)
            launch_pb_mx<double, float, float, false>(tmode, *pbm, 0, 1, b, tp, st, sp, s);
          else
            launch_pb_mx<double, dev::bf16, float, false>(tmode, *pbm, 0, 2, b, tp, st, sp, s);
        } else if (pattern) {
          launch_pb_sn<dev::Pattern, true>(tmode, *pbm, -1, xin, ZP[it & 1], tp, xout, ZP[(it + 1) & 1], colscale, st,
                                           sp, s);
        } else {
          launch_pb_sn<double, false>(tmode, *pbm, 0, xin, nullptr, tp, xout, nullptr, nullptr, st, sp, s);
        }
        // (the anchor move's sum partials: one per superblock of the anchor
        // product's scatter order)
        const int nsb_anchor = (pat32 || mixed_pattern) ? pbm->nsb_w : pbm->nsb;
        hipLaunchKernelGGL(pr_finalize_sn, dim3(1), dim3(kTileThreads), 0, s, sp.perr, sp.ps, sp.pc, pbm->nbins,
                           mixed ? w.hsum.as<double>() : nullptr, mixed ? w.hcomp.as<double>() : nullptr, nsb_anchor,
                           st);
      } else if (fused) {
        if (mixed) {
          float *dcur = D[it & 1], *dnext = D[(it + 1) & 1];
          float *zdcur = ZD[it & 1], *zdnext = ZD[(it + 1) & 1];
          double *xhat = w.xhat.as<double>(), *yhat = w.yhat.as<double>();
          {
            // (mixed runs take the tile products here: with propagation-
            // blocked products they run the sum-normalised branch above)
            hipLaunchKernelGGL(pr_anchor_move, dim3(gN), dim3(256), 0, s, xhat, dcur, pat32 ? zdcur : dcur,
                               mixed_pattern ? Z : nullptr, colscale, n, st);
            if (mixed_pattern)
              launch_anchor<dev::Pattern>(d64, Z, yhat, st, s);
            else
              launch_anchor<double>(d64, xhat, yhat, st, s);
            if (pat32)
              launch_update<dev::Pattern, float, true>(tmode, d64, nullptr, zdcur, yhat, tp, xout, st, slots, s);
            else if (p.precision == PrPrecision::F32)
              launch_update<float, float, true>(tmode, *dlo, nullptr, dcur, yhat, tp, xout, st, slots, s);
            else
              launch_update<dev::bf16, float, true>(tmode, *dlo, nullptr, dcur, yhat, tp, xout, st, slots, s);
          }
          fin_max();
          if (pat32)
            hipLaunchKernelGGL(pr_normalize<3>, dim3(gM), dim3(kTileThreads), 0, s, xout, nullptr, xhat, dcur, dnext,
                               nullptr, nullptr, zdnext, colscale32, n, st, w.perr.as<double>(),
                               w.ps.as<double>(), w.pc.as<double>());
          else
            hipLaunchKernelGGL(pr_normalize<1>, dim3(gM), dim3(kTileThreads), 0, s, xout, nullptr, xhat, dcur, dnext,
                               nullptr, nullptr, nullptr, nullptr, n, st, w.perr.as<double>(), w.ps.as<double>(),
                               w.pc.as<double>());
          fin_norm(gM);
        } else if (pattern) {
          if (pbm)
            launch_pb_update<dev::Pattern, double, double>(tmode, *pbm, -1, Z, tp, xout, st, slots, s);
          else
            launch_update<dev::Pattern, double, false>(tmode, d64, nullptr, Z, nullptr, tp, xout, st, slots, s);
          fin_max();
          hipLaunchKernelGGL(pr_normalize<2>, dim3(gN), dim3(kTileThreads), 0, s, xout, xin, nullptr, nullptr,
                             nullptr, Z, colscale, nullptr, nullptr, n, st, w.perr.as<double>(), w.ps.as<double>(),
                             w.pc.as<double>());
          fin_norm(gN);
        } else {
          if (pbm)
            launch_pb_update<double, double, double>(tmode, *pbm, 0, xin, tp, xout, st, slots, s);
          else
            launch_update<double, double, false>(tmode, d64, p.values, xin, nullptr, tp, xout, st, slots, s);
          fin_max();
          if (pbm)
            hipLaunchKernelGGL(pr_normalize<4>, dim3(gN), dim3(kTileThreads), 0, s, xout, xin, nullptr, nullptr,
                               nullptr, nullptr, nullptr, nullptr, nullptr, n, st, w.perr.as<double>(),
                               w.ps.as<double>(), w.pc.as<double>());
          else
            hipLaunchKernelGGL(pr_normalize<0>, dim3(gN), dim3(kTileThreads), 0, s, xout, xin, nullptr, nullptr,
                               nullptr, nullptr, nullptr, nullptr, nullptr, n, st, w.perr.as<double>(),
                               w.ps.as<double>(), w.pc.as<double>());
          fin_norm(gN);
        }
      } else {
        // any format: y = 0 (accumulating formats), y += A b_t, update, normalise
        hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, xout, n, 0.0);
        if (!d64.noop) launch_spmv(d64, xin, Dtype::F64, xout, Dtype::F64, 0, s);
        if (tmode == 1)
          hipLaunchKernelGGL(pr_update_vec<1>, dim3(gN), dim3(kTileThreads), 0, s, xout, n, tp, st, slots);
        else if (tmode == 2)
          hipLaunchKernelGGL(pr_update_vec<2>, dim3(gN), dim3(kTileThreads), 0, s, xout, n, tp, st, slots);
        else
          hipLaunchKernelGGL(pr_update_vec<0>, dim3(gN), dim3(kTileThreads), 0, s, xout, n, tp, st, slots);
        fin_max();
        hipLaunchKernelGGL(pr_normalize<0>, dim3(gN), dim3(kTileThreads), 0, s, xout, xin, nullptr, nullptr,
                           nullptr, nullptr, nullptr, nullptr, nullptr, n, st, w.perr.as<double>(),
                           w.ps.as<double>(), w.pc.as<double>());
        fin_norm(gN);
      }
    }
    TBSLA_HIP_CHECK(hipGetLastError());
    launched += batch;
  };
  // pinned state copies of the two batches in flight and their events
  // (released on every exit; the stream has drained by then or the run failed)
  struct ReadBack {
    PrState *hbuf = nullptr;
    hipEvent_t cev[2] = {nullptr, nullptr};
    ~ReadBack() {
      if (hbuf) (void)hipHostFree(hbuf);
      for (hipEvent_t e : cev)
        if (e) (void)hipEventDestroy(e);
    }
  } rb;
  PrState *&hbuf = rb.hbuf;
  hipEvent_t *cev = rb.cev;
  if (arena) {
    // allocation-free runs: the state is read after every batch
    while (launched < p.max_iteration) {
      launch_batch(std::min(next, p.max_iteration - launched));
      TBSLA_HIP_CHECK(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
      TBSLA_HIP_CHECK(hipStreamSynchronize(s));
      if (hs.done) break;
      next = next_batch(hs, check);
    }
  } else {
    // Up to two batches in flight, each followed by a copy of the state to
    // pinned memory and an event: while the host waits for one batch's
    // state, the next batch is already queued when the state read before
    // predicts more iterations than are in flight, so the GPU does not idle
    // at the reads (kernels launched past convergence return at once).
    TBSLA_HIP_CHECK(hipHostMalloc(reinterpret_cast<void **>(&hbuf), 2 * sizeof(PrState), hipHostMallocDefault));
    for (int k = 0; k < 2; k++) TBSLA_HIP_CHECK(hipEventCreateWithFlags(&cev[k], hipEventDisableTiming));
    int head = 0, npend = 0;  // slot of the oldest batch in flight, batches in flight
    auto enqueue = [&](int batch) {
      launch_batch(batch);
      const int sl = (head + npend) & 1;
      TBSLA_HIP_CHECK(hipMemcpyAsync(hbuf + sl, st, sizeof(PrState), hipMemcpyDeviceToHost, s));
      TBSLA_HIP_CHECK(hipEventRecord(cev[sl], s));
      npend++;
    };
    bool have = false;  // hs holds a state read back
    if (launched < p.max_iteration) enqueue(std::min(check, p.max_iteration - launched));
    while (npend > 0) {
      if (npend == 1 && launched < p.max_iteration) {
        int nb = check;
        if (have) nb = std::min(check, iterations_left(hs, check) - (launched - hs.it));
        nb = std::min(nb, p.max_iteration - launched);
        if (nb > 0) enqueue(nb);
      }
      TBSLA_HIP_CHECK(hipEventSynchronize(cev[head]));
      hs = hbuf[head];
      head ^= 1;
      npend--;
      have = true;
      if (hs.done) break;
      if (npend == 0 && launched < p.max_iteration) enqueue(std::min(next_batch(hs, check), p.max_iteration - launched));
    }
  }
  if (!arena) TBSLA_HIP_CHECK(hipEventRecord(e1, s));
  const double host_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - h0).count();
  const int it = hs.it;
  // final sum normalisation (Matrix.cpp:52-59), compensated
  const double *a = mixed ? w.xhat.as<double>() : X[it & 1];
  const float *dd = mixed ? D[it & 1] : nullptr;
  hipLaunchKernelGGL(pr_sum_kernel, dim3(gN), dim3(kTileThreads), 0, s, a, dd, n, st, w.ps.as<double>(),
                     w.pc.as<double>());
  hipLaunchKernelGGL(pr_sum_finalize, dim3(1), dim3(kTileThreads), 0, s, w.ps.as<double>(), w.pc.as<double>(),
                     static_cast<int>(gN), w.total.as<double>());
  hipLaunchKernelGGL(pr_scale_kernel, dim3(gN), dim3(256), 0, s, a, dd, n, st, w.total.as<double>(), out_dev);
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  if (!arena) {
    TBSLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  rep.iterations = it;
  rep.anchors = hs.anchors;
  rep.last_error = hs.err;
  rep.bound = hs.bound;
  rep.seconds = arena ? host_seconds : ms * 1e-3;
  return rep;
}

}  // namespace

PageRankReport page_rank_device(const cpp::Matrix &m, const PageRankParams &p, double *out_dev, hipStream_t s) {
  return run_page_rank(m, p, out_dev, s, nullptr, 0);
}

// ---- allocation-free PageRank ----------------------------------------------
// The scratch memory of page_rank_device_noalloc: one fixed buffer in each
// GPU's copy of this code object, held until the process ends.  Host code
// never passes `scratch` itself (in host code it names a host-side stand-in):
// it uses the device address hipGetSymbolAddress returns for the current
// device, looked up once per device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};

unsigned char *scratch_of_current_device() {
  int dev = 0;
  TBSLA_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("tbsla: device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    TBSLA_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw std::runtime_error("tbsla: no device address for the scratch buffer");
    g_scratch[dev] = static_cast<unsigned char *>(ptr);
  }
  return g_scratch[dev];
}

std::size_t slice(std::size_t bytes) { return (std::max<std::size_t>(bytes, 1) + 255) & ~std::size_t(255); }

}  // namespace

std::size_t page_rank_noalloc_bytes(long n, int n_nodes) {
  const std::size_t gN = grid_for(n);
  std::size_t b = slice(sizeof(PrState)) + slice(sizeof(unsigned long long) * kMaxSlots) +
                  3 * slice(sizeof(double) * gN) + slice(sizeof(double)) +
                  2 * slice(sizeof(double) * static_cast<std::size_t>(n));
  if (n_nodes > 0) b += slice(static_cast<std::size_t>(n)) + slice(sizeof(int) * static_cast<std::size_t>(n_nodes));
  return b;
}

PageRankReport page_rank_device_noalloc(const cpp::Matrix &m, const PageRankParams &p, double *out_dev,
                                        hipStream_t s) {
  if (p.precision != PrPrecision::F64)
    throw std::invalid_argument("tbsla: the allocation-free PageRank computes in FP64 (precision TBSLA_F64)");
  if (p.gather == 2)
    throw std::invalid_argument("tbsla: the allocation-free PageRank uses the fixed-order tile products "
                                "(gather 0 or 1): the propagation-blocked layout is built, and allocated, on first use");
  const long n = m.get_n_col();
  const int nodes = p.personalized && !p.teleport ? std::max(p.n_nodes, 0) : 0;
  const std::size_t need = page_rank_noalloc_bytes(n, nodes);
  if (need > kNoAllocScratchBytes)
    throw std::invalid_argument("tbsla: allocation-free PageRank on " + std::to_string(n) + " nodes needs " +
                                std::to_string(need) + " bytes of scratch; the buffer holds " +
                                std::to_string(kNoAllocScratchBytes));
  if (!mirror_if_built(m, Dtype::F64))
    throw std::invalid_argument("tbsla: the allocation-free PageRank needs the FP64 mirror: call "
                                "tbsla_hip_prepare(m, TBSLA_F64) first");
  PageRankParams q = p;
  q.gather = 1;
  return run_page_rank(m, q, out_dev, s, scratch_of_current_device(), kNoAllocScratchBytes);
}

PageRankReport page_rank(const cpp::Matrix &m, const PageRankParams &p, double *out) {
  const long n = m.get_n_col();
  DevBuf o;
  o.alloc(sizeof(double) * static_cast<std::size_t>(std::max(1L, n)));
  PageRankReport r = page_rank_device(m, p, o.as<double>(), nullptr);
  if (n > 0) TBSLA_HIP_CHECK(hipMemcpy(out, o.get(), sizeof(double) * static_cast<std::size_t>(n), hipMemcpyDeviceToHost));
  return r;
}

}  // namespace hip
}  // namespace tbsla
