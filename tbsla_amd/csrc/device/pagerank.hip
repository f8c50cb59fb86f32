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
#include <stdexcept>
#include <vector>

#include "kernels.hpp"
#include "mirror.hpp"
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
  int it, max_it, done, rule;
  int need_anchor, anchors, anchor_it, mixed;
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
    const double floor = st->floor_eps * st->bound_anchor * 10.0;
    if (converged && floor > st->bound_target) {
      converged = false;
      anchor = 1;
    }
    if (!converged) {
      if (bound <= st->anchor_tau * st->bound_anchor) anchor = 1;
      if (st->it - st->anchor_it >= 3 && bound > 0.9 * st->bound_prev) anchor = 1;
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
  const bool moved = st->need_anchor != 0;
  double err = 0;
  CSum cs{0, 0};
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i < n;
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

// One block: residual and compensated sum of the normalise pass, then the
// step decision.
__global__ __launch_bounds__(kTileThreads) void pr_finalize_norm(const double *__restrict__ perr,
                                                                 const double *__restrict__ ps,
                                                                 const double *__restrict__ pc, int nb, PrState *st) {
  if (st->done) return;
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  double e = 0;
  CSum t{0, 0};
  for (int i = threadIdx.x; i < nb; i += kTileThreads) {
    e += perr[i];
    t = csum_merge(t, CSum{ps[i], pc[i]});
  }
  e = dev::block_tree_sum<double>(e, red_s);
  t = block_csum(t, red_s, red_c);
  if (threadIdx.x == 0) finalize_step(st, e, t.s + t.c);
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
__global__ __launch_bounds__(kTileThreads) void pr_sum_kernel(const double *__restrict__ a,
                                                              const float *__restrict__ d, long n,
                                                              double *__restrict__ ps, double *__restrict__ pc) {
#pragma clang fp contract(off)
  __shared__ double red_s[kTileThreads], red_c[kTileThreads];
  CSum cs{0, 0};
  for (long i = blockIdx.x * static_cast<long>(kTileThreads) + threadIdx.x; i < n;
       i += static_cast<long>(gridDim.x) * kTileThreads)
    csum_add(cs, d ? a[i] + static_cast<double>(d[i]) : a[i]);
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
                                                       long n, const double *__restrict__ total,
                                                       double *__restrict__ out) {
#pragma clang fp contract(off)
  const double S = *total;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256)
    out[i] = (d ? a[i] + static_cast<double>(d[i]) : a[i]) / S;
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
};

// The anchor move, folded into the scatter of the iteration that makes it:
// over the superblock's columns c, xhat[c] += delta[c], and the product's
// input is z[c] = colscale[c] * xhat[c] (value-free anchor product; z is
// written for the gather) or xhat[c] itself (z == nullptr).
struct AnchorMove {
  double *xhat;
  const float *delta;
  double *z;
  const double *colscale;
};

// The far products of one superblock (picked from blockIdx): its slice of x
// staged in xs (LDS), then every far product v * x[c] written to its slot.
// With MOVE the slice is the anchor move's result instead of x.
template <typename VT, typename XT, typename AT, bool MOVE>
__device__ inline void pb_scatter_sb(const PbScatterArgs &a, const VT *__restrict__ val, const XT *__restrict__ x,
                                     AT *__restrict__ prod, XT *xs, const AnchorMove &am) {
#pragma clang fp contract(off)
  // superblocks of one XCD (blockIdx % 8) are contiguous, so the runs that
  // share a cache line at a run boundary are written through the same L2
  const int per = (gridDim.x + 7) / 8;
  const int sb = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (sb >= a.nsb) return;
  const int k0 = a.sp[sb], k1 = a.sp[sb + 1];
  const int *__restrict__ delta = a.rdelta + a.rp[sb];  // slot - entry of each run of this superblock
  const int *__restrict__ cbase = a.cbase + a.cp[sb];   // run of the first entry of each chunk
  const long c0 = static_cast<long>(sb) << kPbColShift;
  const int nc = static_cast<int>(min(static_cast<long>(kPbCols), static_cast<long>(a.n) - c0));
  {
    constexpr int S = kPbCols / kPbScatterThreads;  // x values per thread, all loads issued first
    XT t[S];
#pragma unroll
    for (int u = 0; u < S; u++) {
      const int i = threadIdx.x + u * kPbScatterThreads;
      if (MOVE) {
        if (i < nc) {
          const long c = c0 + i;
          const double v = am.xhat[c] + static_cast<double>(am.delta[c]);
          am.xhat[c] = v;
          double xv = v;
          if (am.z) {
            xv = am.colscale[c] * v;
            am.z[c] = xv;
          }
          t[u] = static_cast<XT>(xv);
        } else {
          t[u] = XT(0);
        }
      } else {
        t[u] = i < nc ? x[c0 + i] : XT(0);
      }
    }
#pragma unroll
    for (int u = 0; u < S; u++) xs[threadIdx.x + u * kPbScatterThreads] = t[u];
  }
  __syncthreads();
  // four entries per thread in flight: every global load first, then the
  // run offsets and the LDS reads of x, then the products and the stores.
  // The lanes of a wavefront hold consecutive entries of one 64-entry chunk
  // (kPbScatterThreads is a multiple of 64), so an entry's run is the
  // chunk's first run plus the run starts on lanes 1..lane, counted from a
  // ballot; lanes past the end are the chunk's last and hold no run start.
  static_assert(kPbScatterThreads % kPbChunk == 0 && kPbChunk == 64, "one chunk per wavefront");
  const int lane = threadIdx.x & 63;
  const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane
  for (int k = k0 + threadIdx.x; k < k1; k += 4 * kPbScatterThreads) {
    unsigned key[4];
    int sl[4], cb[4];
    AT v1[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k + u * kPbScatterThreads;
      const bool ok = kk < k1;
      key[u] = ok ? a.fkey[kk] : 0u;
      cb[u] = ok ? cbase[(kk - k0) >> 6] : 0;
      v1[u] = ok ? PbVal<VT>::template prod<AT>(val, kk, AT(1)) : AT(0);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int kk = k + u * kPbScatterThreads;
      const unsigned long long starts = __ballot((key[u] & kPbRunStart) != 0u);
      const int run = cb[u] + __popcll(starts & upto & ~1ull);
      sl[u] = kk < k1 ? kk + delta[run] : -1;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (sl[u] >= 0) prod[sl[u]] = v1[u] * static_cast<AT>(xs[key[u] & static_cast<unsigned>(kPbCols - 1)]);
  }
}

// FP64 runs: the far products of x.
template <typename VT, typename XT, typename AT>
__global__ __launch_bounds__(kPbScatterThreads) void pb_scatter(PbScatterArgs a, const VT *__restrict__ val,
                                                                const XT *__restrict__ x, AT *__restrict__ prod,
                                                                const PrState *st) {
  if (st->done) return;
  __shared__ XT xs[kPbCols];
  pb_scatter_sb<VT, XT, AT, false>(a, val, x, prod, xs, AnchorMove{});
}

// Mixed runs, one launch per iteration: in an iteration that moves the
// anchor, the anchor move and the FP64 anchor product's far products (values
// VA, FP64); otherwise the delta product's (values VD, delta XD, sums AD).
template <typename VA, typename VD, typename XD, typename AD>
__global__ __launch_bounds__(kPbScatterThreads) void pb_scatter_mixed(PbScatterArgs a, const VA *__restrict__ vala,
                                                                      const VD *__restrict__ vald,
                                                                      const XD *__restrict__ xd, void *prod,
                                                                      AnchorMove am, const PrState *st) {
  if (st->done) return;
  __shared__ double xs[kPbCols];
  if (st->need_anchor)
    pb_scatter_sb<VA, double, double, true>(a, vala, nullptr, static_cast<double *>(prod), xs, am);
  else
    pb_scatter_sb<VD, XD, AD, false>(a, vald, xd, static_cast<AD *>(prod), reinterpret_cast<XD *>(xs), am);
}

struct PbView {
  const int *bp, *lp;
  const unsigned short *frow, *lrow, *lcol;
  int n;
};

// One round of a bin's far products: U loads of (row, product) per thread,
// 1024 entries apart, then their LDS additions. MASKED checks each entry
// against the end of the bin's run.
template <int U, bool MASKED, typename AT>
__device__ inline void pb_far_round(const unsigned short *__restrict__ frow, const AT *__restrict__ prod, int k,
                                    int f1, double *ybin) {
  int r[U];
  double pv[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const int kk = k + u * 1024;
    if (MASKED) {
      r[u] = kk < f1 ? frow[kk] : -1;
      pv[u] = kk < f1 ? static_cast<double>(prod[kk]) : 0.0;
    } else {
      r[u] = frow[kk];
      pv[u] = static_cast<double>(prod[kk]);
    }
  }
#pragma unroll
  for (int u = 0; u < U; u++)
    if (!MASKED || r[u] >= 0) atomicAdd(&ybin[r[u]], pv[u]);
}

// The sums of one bin (blockIdx) and its row epilogue.
// MODE 0: b = beta*(yhat + s) + t (MIXED) or beta*s + t, block max.
// MODE 2: the anchor product with the update of an iteration that moves the
//         anchor (its delta is zero): yhat_out = s, b = beta*s + t, block max.
template <typename VT, typename XT, typename AT, bool MIXED, int TM, int MODE>
__device__ inline void pb_gather_bin(const PbView &v, const VT *__restrict__ lval, const XT *__restrict__ x,
                                     const AT *__restrict__ prod, const double *__restrict__ yhat, const Teleport &tp,
                                     double *__restrict__ out, const PrState *st, unsigned long long *slots,
                                     double *__restrict__ yhat_out, double *ybin, double *red) {
#pragma clang fp contract(off)
  static_assert(MODE == 0 || (MODE == 2 && !MIXED), "pb_gather modes: 0, and 2 without MIXED");
  const int bin = blockIdx.x, tid = threadIdx.x;
  const int base = bin << kPbShift;
  const int rows = min(kPbRows, v.n - base);
  for (int i = tid; i < kPbRows; i += 1024) ybin[i] = 0.0;
  __syncthreads();
  const int f1 = v.bp[bin + 1], l1 = v.lp[bin + 1];
  // far products: eight loads per thread issued before their LDS additions
  // (the compiler does not move global loads across LDS atomics); rounds
  // with all eight in range run unmasked, and one masked round ends the run
  constexpr int U = 8;
  int kf = v.bp[bin] + tid;
  for (; kf + (U - 1) * 1024 < f1; kf += U * 1024) pb_far_round<U, false>(v.frow, prod, kf, f1, ybin);
  if (kf < f1) pb_far_round<U, true>(v.frow, prod, kf, f1, ybin);
  // local entries are in row order and each bin's run is padded to groups
  // of four: one group per thread (aligned vector loads), runs of one row
  // summed and added once
  constexpr int LU = 2;  // groups per thread in flight (4 measured slower)
  for (int k = v.lp[bin] + kPbLocalGroup * tid; k < l1; k += LU * kPbLocalGroup * 1024) {
    int r[LU][4], c[LU][4];
    AT lv[LU][4], xv[LU][4];
#pragma unroll
    for (int g = 0; g < LU; g++) {
      const int kg = k + g * kPbLocalGroup * 1024;
      if (g == 0 || kg < l1) {
        const dev::v4us rr = *reinterpret_cast<const dev::v4us *>(v.lrow + kg);
        const dev::v4us cc = *reinterpret_cast<const dev::v4us *>(v.lcol + kg);
        PbVal<VT>::template get4<AT>(lval, kg, lv[g]);
        r[g][0] = rr.x, r[g][1] = rr.y, r[g][2] = rr.z, r[g][3] = rr.w;
        c[g][0] = cc.x, c[g][1] = cc.y, c[g][2] = cc.z, c[g][3] = cc.w;
      } else {  // past the bin's run: a group of padding
#pragma unroll
        for (int u = 0; u < 4; u++) r[g][u] = kPbPadRow, c[g][u] = 0, lv[g][u] = AT(0);
      }
    }
#pragma unroll
    for (int g = 0; g < LU; g++)
#pragma unroll
      for (int u = 0; u < 4; u++) xv[g][u] = static_cast<AT>(x[base + c[g][u]]);
#pragma unroll
    for (int g = 0; g < LU; g++) {
      double acc = 0.0;
      int row = r[g][0];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (r[g][u] == kPbPadRow) break;  // padding closes a bin's run
        if (r[g][u] != row) {
          atomicAdd(&ybin[row], acc);
          acc = 0.0;
          row = r[g][u];
        }
        acc += static_cast<double>(PbVal<VT>::template mul<AT>(lv[g][u], xv[g][u]));
      }
      if (row != kPbPadRow) atomicAdd(&ybin[row], acc);
    }
  }
  __syncthreads();
  double lmax = -INFINITY;
  const double beta = st->beta, tele = st->tele;
  for (int i = tid; i < rows; i += 1024) {
    const int r = base + i;
    if (MODE == 2) yhat_out[r] = static_cast<double>(ybin[i]);
    const double sv = MIXED ? yhat[r] + static_cast<double>(ybin[i]) : static_cast<double>(ybin[i]);
    const double bi = add_teleport<TM>(beta * sv, tele, tp, r);
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

// FP64 runs.
template <typename VT, typename XT, typename AT, int TM>
__global__ __launch_bounds__(1024) void pb_gather(PbView v, const VT *__restrict__ lval, const XT *__restrict__ x,
                                                  const AT *__restrict__ prod, Teleport tp, double *__restrict__ out,
                                                  PrState *st, unsigned long long *slots) {
  if (st->done) return;
  __shared__ double ybin[kPbRows];  // FP64 sums (products may be FP32)
  __shared__ double red[16];
  pb_gather_bin<VT, XT, AT, false, TM, 0>(v, lval, x, prod, nullptr, tp, out, st, slots, nullptr, ybin, red);
}

// Mixed runs, the partner of pb_scatter_mixed: the anchor product with the
// update (MODE 2, x = xa) or the delta product with the update (x = xd).
template <typename VA, typename VD, typename XD, typename AD, int TM>
__global__ __launch_bounds__(1024) void pb_gather_mixed(PbView v, const VA *__restrict__ lvala,
                                                        const double *__restrict__ xa, const VD *__restrict__ lvald,
                                                        const XD *__restrict__ xd, const void *prod, double *yhat,
                                                        Teleport tp, double *__restrict__ out, PrState *st,
                                                        unsigned long long *slots) {
  if (st->done) return;
  __shared__ double ybin[kPbRows];
  __shared__ double red[16];
  if (st->need_anchor)
    pb_gather_bin<VA, double, double, false, TM, 2>(v, lvala, xa, static_cast<const double *>(prod), nullptr, tp, out,
                                                    st, slots, yhat, ybin, red);
  else
    pb_gather_bin<VD, XD, AD, true, TM, 0>(v, lvald, xd, static_cast<const AD *>(prod), yhat, tp, out, st, slots,
                                           nullptr, ybin, red);
}

PbView pb_view(const PbMatrix &pb) {
  return PbView{pb.bp.as<int>(), pb.lp.as<int>(), pb.frow.as<unsigned short>(), pb.lrow.as<unsigned short>(),
                pb.lcol.as<unsigned short>(), pb.n};
}

PbScatterArgs pb_args(const PbMatrix &pb) {
  return PbScatterArgs{pb.sp.as<int>(),    pb.fkey.as<unsigned short>(), pb.rp.as<int>(), pb.rdelta.as<int>(),
                       pb.cbase.as<int>(), pb.cp.as<int>(),              pb.nsb,          pb.n};
}

// One propagation-blocked product with the update (FP64 runs): values of
// type VT from `pb` (dev::Pattern: value-free), x of type XT, sums of type AT.
template <typename VT, typename XT, typename AT, int TM>
void launch_pb(const PbMatrix &pb, int vslot, const XT *x, const Teleport &tp, double *out, PrState *st,
               unsigned long long *slots, hipStream_t s) {
  const VT *fval = vslot >= 0 ? pb.fval[vslot].as<VT>() : nullptr;
  const VT *lval = vslot >= 0 ? pb.lval[vslot].as<VT>() : nullptr;
  AT *prod = pb.prod.as<AT>();
  if (pb.n_far > 0)
    hipLaunchKernelGGL((pb_scatter<VT, XT, AT>), dim3((pb.nsb + 7) / 8 * 8), dim3(kPbScatterThreads), 0, s,
                       pb_args(pb), fval, x, prod, st);
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

// One iteration's product and update of a mixed run: the anchor move and
// the FP64 anchor product (values VA at slot va, x = z or xhat) when the
// state asks for it, else the delta product (values VD at slot vd, delta
// xd).  The scatter always runs: it also makes the anchor move.
template <typename VA, typename VD, typename XD, typename AD, int TM>
void launch_pb_mixed(const PbMatrix &pb, int va, int vd, const AnchorMove &am, const XD *xd, double *yhat,
                     const Teleport &tp, double *out, PrState *st, unsigned long long *slots, hipStream_t s) {
  const VA *fa = va >= 0 ? pb.fval[va].as<VA>() : nullptr;
  const VA *la = va >= 0 ? pb.lval[va].as<VA>() : nullptr;
  const VD *fd = vd >= 0 ? pb.fval[vd].as<VD>() : nullptr;
  const VD *ld = vd >= 0 ? pb.lval[vd].as<VD>() : nullptr;
  const double *xa = am.z ? am.z : am.xhat;
  hipLaunchKernelGGL((pb_scatter_mixed<VA, VD, XD, AD>), dim3((pb.nsb + 7) / 8 * 8), dim3(kPbScatterThreads), 0, s,
                     pb_args(pb), fa, fd, xd, pb.prod.get(), am, st);
  hipLaunchKernelGGL((pb_gather_mixed<VA, VD, XD, AD, TM>), dim3(pb.nbins), dim3(1024), 0, s, pb_view(pb), la, xa,
                     ld, xd, pb.prod.get(), yhat, tp, out, st, slots);
}

template <typename VA, typename VD, typename XD, typename AD>
void launch_pb_mixed_update(int tmode, const PbMatrix &pb, int va, int vd, const AnchorMove &am, const XD *xd,
                            double *yhat, const Teleport &tp, double *out, PrState *st, unsigned long long *slots,
                            hipStream_t s) {
  if (tmode == 1)
    launch_pb_mixed<VA, VD, XD, AD, 1>(pb, va, vd, am, xd, yhat, tp, out, st, slots, s);
  else if (tmode == 2)
    launch_pb_mixed<VA, VD, XD, AD, 2>(pb, va, vd, am, xd, yhat, tp, out, st, slots, s);
  else
    launch_pb_mixed<VA, VD, XD, AD, 0>(pb, va, vd, am, xd, yhat, tp, out, st, slots, s);
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

struct Scratch {
  DevBuf st, slots, perr, ps, pc, total, member;
  DevBuf X0, X1, Z, yhat, xhat, D0, D1, colscale, ZD0, ZD1, colscale32;
};

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

PageRankReport page_rank_device(const cpp::Matrix &m, const PageRankParams &p, double *out_dev, hipStream_t s) {
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
  const DeviceMatrix *dlo = nullptr;
  if (p.precision == PrPrecision::F32) dlo = &mirror(m, Dtype::F32);
  if (p.precision == PrPrecision::BF16) dlo = &mirror(m, Dtype::BF16);
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
  if (csr && !d64.noop && d64.rows != n)
    throw std::invalid_argument("tbsla: CSR rowptr does not cover every row");
  if (d64.x_need > n) throw std::invalid_argument("tbsla: column index beyond n_col");

  Scratch w;
  const unsigned gN = grid_for(n);
  // the mixed normalise pass (six streams) measured faster with twice the
  // blocks (four and eight times: slower); the FP64 one (three streams) slower
  const unsigned gM = mixed ? grid_for(n, 2L * kNormBlocks) : gN;
  w.st.alloc(sizeof(PrState));
  w.slots.alloc(sizeof(unsigned long long) * kMaxSlots);
  w.perr.alloc(sizeof(double) * std::max(gN, gM));
  w.ps.alloc(sizeof(double) * std::max(gN, gM));
  w.pc.alloc(sizeof(double) * std::max(gN, gM));
  w.total.alloc(sizeof(double));
  w.X0.alloc(sizeof(double) * n);
  w.X1.alloc(sizeof(double) * n);
  // teleport mode: 0 uniform, 1 node set (Matrix.cpp:63-114), 2 distribution vector
  const int tmode = p.teleport ? 2 : (p.personalized ? 1 : 0);
  Teleport tp{nullptr, p.teleport};
  const double b0 = p.personalized && !p.teleport ? 0.25 : 1.0;
  std::vector<unsigned char> mem;
  if (tmode == 1) {
    mem.assign(static_cast<std::size_t>(n), 0);
    for (int k = 0; k < p.n_nodes; k++)
      if (p.nodes[k] >= 0 && p.nodes[k] < n) mem[static_cast<std::size_t>(p.nodes[k])] = 1;
    w.member.upload(mem.data(), mem.size(), s);
    tp.member = w.member.as<unsigned char>();
  }
  const bool pattern = p.precision == PrPrecision::PATTERN;
  if (pattern || mixed) {
    // 1/count per column, read back from the stored values (every value of
    // column j equals fl(1/count_j) after fill_cqmat_stochastic)
    const auto &csrm = dynamic_cast<const cpp::MatrixCSR &>(m);
    std::vector<double> cs(static_cast<std::size_t>(n), 0.0);
    const auto &ci = csrm.get_colidx();
    const auto &va = csrm.get_values();
    const auto &rp = csrm.get_rowptr();
    // the stored entries the products read: [0, rowptr[rows] - rowptr[0])
    const std::size_t used = rp.size() < 2 ? 0 : static_cast<std::size_t>(rp.back() - rp.front());
    if (used > ci.size() || used > va.size())
      throw std::invalid_argument("tbsla: CSR colidx / values shorter than rowptr says");
    for (std::size_t k = 0; k < used; k++) {
      const int c = ci[k];
      if (c < 0 || c >= n) throw std::invalid_argument("tbsla: column index outside [0, n_col)");
      cs[static_cast<std::size_t>(c)] = va[k];
    }
    w.colscale.upload(cs.data(), cs.size() * sizeof(double), s);
    if (pat32) {
      std::vector<float> c32(cs.begin(), cs.end());
      w.colscale32.upload(c32.data(), c32.size() * sizeof(float), s);
    }
  }
  PrState h{};
  h.beta = p.beta;
  h.eps = p.epsilon;
  h.bound_target = p.epsilon * (1.0 - p.beta) / p.beta;
  h.tele_scale = tmode == 2   ? (1.0 - p.beta)
                 : tmode == 1 ? (1.0 - p.beta) / static_cast<double>(static_cast<std::size_t>(p.n_nodes))
                              : (1.0 - p.beta) / static_cast<double>(n);
  h.sum = b0 * static_cast<double>(n);
  h.tele = h.sum * h.tele_scale;
  h.bound = 2.0;
  h.bound_prev = 2.0;
  h.bound_anchor = 2.0;
  h.max_it = p.max_iteration;
  h.rule = static_cast<int>(p.rule);
  h.mixed = mixed ? 1 : 0;
  // mixed runs start from the anchor xhat = 0 (so yhat = A xhat = 0 with no
  // product) and delta = b0: the first phase carries the whole iterate in the
  // low precision, and the first anchor move comes when it has converged to
  // that precision's floor
  h.need_anchor = 0;
  h.anchors = 0;
  h.floor_eps = p.precision == PrPrecision::BF16 ? 4e-3 : 6e-8;  // FP32 and PATTERN_F32: 6e-8
  h.anchor_tau = p.anchor_tau > 0 ? p.anchor_tau : (p.precision == PrPrecision::BF16 ? 5e-2 : 3e-6);
  h.done = p.max_iteration <= 0 ? 1 : 0;
  TBSLA_HIP_CHECK(hipMemcpyAsync(w.st.get(), &h, sizeof(h), hipMemcpyHostToDevice, s));
  PrState *st = w.st.as<PrState>();
  double *X[2] = {w.X0.as<double>(), w.X1.as<double>()};
  hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, X[0], n, b0);
  unsigned long long *slots = w.slots.as<unsigned long long>();
  hipLaunchKernelGGL(zero_u64, dim3(1), dim3(kMaxSlots), 0, s, slots, kMaxSlots);
  float *D[2] = {nullptr, nullptr};
  if (mixed) {
    w.xhat.alloc(sizeof(double) * n);
    w.yhat.alloc(sizeof(double) * n);
    w.D0.alloc(sizeof(float) * n);
    w.D1.alloc(sizeof(float) * n);
    D[0] = w.D0.as<float>();
    D[1] = w.D1.as<float>();
    hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, w.xhat.as<double>(), n, 0.0);
    hipLaunchKernelGGL(fill_f64, dim3(gN), dim3(256), 0, s, w.yhat.as<double>(), n, 0.0);
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, D[0], n, static_cast<float>(b0));
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, D[1], n, 0.0f);
  }
  float *ZD[2] = {nullptr, nullptr};
  if (pat32) {
    w.ZD0.alloc(sizeof(float) * n);
    w.ZD1.alloc(sizeof(float) * n);
    ZD[0] = w.ZD0.as<float>();
    ZD[1] = w.ZD1.as<float>();
    hipLaunchKernelGGL(scale_into_f32, dim3(gN), dim3(256), 0, s, w.colscale.as<double>(), b0, ZD[0], n);
    hipLaunchKernelGGL(fill_f32, dim3(gN), dim3(256), 0, s, ZD[1], n, 0.0f);
  }
  if (pattern || (mixed && m.count_stochastic())) {
    w.Z.alloc(sizeof(double) * n);
    hipLaunchKernelGGL(scale_into, dim3(gN), dim3(256), 0, s, w.colscale.as<double>(), X[0], w.Z.as<double>(), n);
  }
  TBSLA_HIP_CHECK(hipGetLastError());
  // the set-up copies read pageable host memory (state, node set, 1/count)
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));

  hipEvent_t e0, e1;
  TBSLA_HIP_CHECK(hipEventCreate(&e0));
  TBSLA_HIP_CHECK(hipEventCreate(&e1));
  TBSLA_HIP_CHECK(hipEventRecord(e0, s));
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
  const double *colscale = w.colscale.as<double>();
  while (launched < p.max_iteration) {
    const int batch = std::min(next, p.max_iteration - launched);
    for (int k = 0; k < batch; k++) {
      const int it = launched + k;
      double *xin = X[it & 1], *xout = X[(it + 1) & 1];
      if (fused) {
        if (mixed) {
          float *dcur = D[it & 1], *dnext = D[(it + 1) & 1];
          float *zdcur = ZD[it & 1], *zdnext = ZD[(it + 1) & 1];
          double *xhat = w.xhat.as<double>(), *yhat = w.yhat.as<double>();
          if (pbm) {
            // one scatter and one gather per iteration: the anchor move and
            // the FP64 anchor product with the update when the state asks for
            // an anchor, else the delta product with the update
            const AnchorMove am{xhat, dcur, mixed_pattern ? Z : nullptr, colscale};
            if (pat32)
              launch_pb_mixed_update<dev::Pattern, dev::Pattern, float, float>(tmode, *pbm, -1, -1, am, zdcur, yhat,
                                                                                tp, xout, st, slots, s);
            else if (mixed_pattern && p.precision == PrPrecision::F32)
              launch_pb_mixed_update<dev::Pattern, float, float, float>(tmode, *pbm, -1, 1, am, dcur, yhat, tp, xout,
                                                                        st, slots, s);
            else if (mixed_pattern)
              launch_pb_mixed_update<dev::Pattern, dev::bf16, float, float>(tmode, *pbm, -1, 2, am, dcur, yhat, tp,
                                                                            xout, st, slots, s);
            else if (p.precision == PrPrecision::F32)
              launch_pb_mixed_update<double, float, float, float>(tmode, *pbm, 0, 1, am, dcur, yhat, tp, xout, st,
                                                                  slots, s);
            else
              launch_pb_mixed_update<double, dev::bf16, float, float>(tmode, *pbm, 0, 2, am, dcur, yhat, tp, xout,
                                                                      st, slots, s);
          } else {
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
                               nullptr, nullptr, zdnext, w.colscale32.as<float>(), n, st, w.perr.as<double>(),
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
    TBSLA_HIP_CHECK(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    if (hs.done) break;
    next = next_batch(hs, check);
  }
  TBSLA_HIP_CHECK(hipEventRecord(e1, s));
  const int it = hs.it;
  // final sum normalisation (Matrix.cpp:52-59), compensated
  const double *a = mixed ? w.xhat.as<double>() : X[it & 1];
  const float *dd = mixed ? D[it & 1] : nullptr;
  hipLaunchKernelGGL(pr_sum_kernel, dim3(gN), dim3(kTileThreads), 0, s, a, dd, n, w.ps.as<double>(),
                     w.pc.as<double>());
  hipLaunchKernelGGL(pr_sum_finalize, dim3(1), dim3(kTileThreads), 0, s, w.ps.as<double>(), w.pc.as<double>(),
                     static_cast<int>(gN), w.total.as<double>());
  hipLaunchKernelGGL(pr_scale_kernel, dim3(gN), dim3(256), 0, s, a, dd, n, w.total.as<double>(), out_dev);
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  TBSLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  rep.iterations = it;
  rep.anchors = hs.anchors;
  rep.last_error = hs.err;
  rep.bound = hs.bound;
  rep.seconds = ms * 1e-3;
  return rep;
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
