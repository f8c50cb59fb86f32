// Batched personalised PageRank engine (layout and launch sequence in
// batch.hpp).
#include <hip/hip_runtime.h>

#include <tbsla/cpp/MatrixCSR.hpp>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <sstream>
#include <cmath>
#include <cstring>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <type_traits>
#include <utility>

#include "batch.hpp"
#include "mirror.hpp"
#include "tile_csr.hpp"

namespace tbsla {
namespace hip {
namespace {

constexpr int kLanes = 64;           // RHS per SpMM group (one per lane)
// SpMM rows per wavefront (a work item is four wavefronts' rows): the
// autotuner picks one of these
constexpr int kSpmmRpw[3] = {8, 16, 32};
constexpr int kSlots = 16;           // atomic-max slots per RHS
constexpr int kChunkRows = 2048;     // rows per pass item, single-RHS graphs
// entries per tile of the single-RHS graphs, picked by the autotuner (the
// persistent step holds tile and SpMM items in one kernel, so the SpMV's
// 3072 is not always best).  Rows longer than kBatchLongRow entries are long
// rows at every tile size, summed in chunks of kBatchLongChunk, so every
// tile size computes the same bits.
constexpr int kBatchTiles[3] = {1024, 2048, 3072};
constexpr int kBatchLongRow = 512;
constexpr int kBatchLongChunk = 1024;
constexpr int kSpmmChunkRows = 1024;  // rows per pass item, SpMM groups
constexpr int kMaxGroups = 64;
constexpr int kQueues = 8;           // one per XCD

int tile_index(int tile_nnz) { return tile_nnz == 1024 ? 0 : (tile_nnz == 3072 ? 2 : 1); }
int rpw_index(int spmm_rows) { return spmm_rows == 32 ? 0 : (spmm_rows == 128 ? 2 : 1); }

struct RhsState {
  double tele_scale, tele, maxv, err, sum, sum_prev, bound, total;
  int it, done, pad0, pad1;
};

struct PassPart {
  double e, s, c;
};

struct RhsInfo {
  long base;  // first partial (index into the PassPart array)
  int count, stride;
};

struct PassItem {
  int kind;  // 0: rows [r0, r1) of single-RHS graph `id`; 1: rows [r0, r1) of SpMM group `id`
  int id, r0, r1;
  long part;  // first partial written by the item
};

template <typename VT>
struct SpmmGroup {
  const int *rowptr;
  const int *colidx;
  const VT *vals;
  const unsigned long long *member;  // bit j: row is in the node set of lane j
  int rows, rp0, f_col, n_rhs, rhs_base, item_base, nitems, pad;
  long xoff;  // first element of the group's n x 64 iterate
};

template <typename VT, typename XT>
struct BatchDev {
  dev::CsrView<VT> a;  // block-diagonal single-RHS graphs
  int a_ntiles, a_nlong, a_rows, n_groups;
  const int *row_graph;
  const int *row_base;  // first row of each single-RHS graph
  const unsigned char *a_member;
  XT *xa0, *xa1, *xb0, *xb1;  // iterate buffers by parity
  const SpmmGroup<VT> *g;     // n_groups descriptors (device memory)
  RhsState *rs;
  unsigned long long *slots;
  const int *q_items;
  const int *q_off;           // kQueues + 1 queue boundaries in q_items
  int xcd_map;
  int spmm_rpw;  // SpMM rows per wavefront
  double beta;
};

__device__ inline unsigned long long ord_encode(double v) {
  const unsigned long long u = static_cast<unsigned long long>(__double_as_longlong(v));
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ inline double ord_decode(unsigned long long u) {
  if (u == 0) return -INFINITY;
  return __longlong_as_double(static_cast<long long>((u >> 63) ? (u & 0x7fffffffffffffffull) : ~u));
}

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

// Row epilogue of the single-RHS tiles: b = beta*s (+ t on listed rows),
// running maxima of the tile's first and last graph (rows of any graph in
// between, possible only for graphs shorter than a tile, go to the slots).
template <typename XT>
struct TileEpi {
  XT *y;
  const int *row_graph;
  RhsState *rs;
  const unsigned char *member;
  unsigned long long *slots;
  double beta;
  int gA, gB, slot;
  double mA, mB;
  __device__ void operator()(int row, double s) {
#pragma clang fp contract(off)
    const int g = row_graph[row];
    if (rs[g].done) return;
    double bi = beta * s;
    if (member[row]) bi = bi + rs[g].tele;
    const XT bs = static_cast<XT>(bi);
    y[row] = bs;
    const double bv = static_cast<double>(bs);
    if (g == gA)
      mA = fmax(mA, bv);
    else if (g == gB)
      mB = fmax(mB, bv);
    else
      atomicMax(slots + g * kSlots + slot, ord_encode(bv));
  }
  __device__ void operator()(int row, float s) { (*this)(row, static_cast<double>(s)); }
};

// Four consecutive iterate values (the RHS quad of one lane), 16-byte loads.
template <typename XT>
struct Quad;
template <>
struct Quad<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static void load(const float *p, float (&o)[4]) {
    const v4 a = *reinterpret_cast<const v4 *>(p);
    o[0] = a.x;
    o[1] = a.y;
    o[2] = a.z;
    o[3] = a.w;
  }
  __device__ static void store(float *p, const float (&o)[4]) { *reinterpret_cast<v4 *>(p) = v4{o[0], o[1], o[2], o[3]}; }
};
template <>
struct Quad<double> {
  typedef double v2 __attribute__((ext_vector_type(2)));
  __device__ static void load(const double *p, double (&o)[4]) {
    const v2 a = reinterpret_cast<const v2 *>(p)[0], b = reinterpret_cast<const v2 *>(p)[1];
    o[0] = a.x;
    o[1] = a.y;
    o[2] = b.x;
    o[3] = b.y;
  }
  __device__ static void store(double *p, const double (&o)[4]) {
    reinterpret_cast<v2 *>(p)[0] = v2{o[0], o[1]};
    reinterpret_cast<v2 *>(p)[1] = v2{o[2], o[3]};
  }
};

// Lane L of each 16-lane row of the wavefront, broadcast to the whole row
// (a DPP row broadcast: a VALU move, no LDS round trip).
template <int L>
__device__ inline int row_bcast(int v) {
  static_assert(L >= 0 && L < 16, "lane within a 16-lane row");
  return __builtin_amdgcn_mov_dpp(v, 0x150 + L, 0xF, 0xF, false);
}
template <int L>
__device__ inline float row_bcast(float v) {
  return __int_as_float(row_bcast<L>(__float_as_int(v)));
}
template <int L>
__device__ inline double row_bcast(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = row_bcast<L>(static_cast<int>(b)), hi = row_bcast<L>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
// f(integral_constant<int, I>) for I = B .. E-1
template <int B, int E, typename F>
__device__ inline void static_for(F &&f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// SpMM rows of one work item (64 rows): each wavefront keeps four rows in
// flight, 16 lanes per row, lane quad q holding RHS 4q..4q+3.  A row's
// column indices and values are loaded 16 at a time (one entry per lane) and
// broadcast inside the 16-lane group; UNR gathers of 4 consecutive iterate
// values are issued before they are accumulated, in entry order, so every
// (row, RHS) sum is the reference's left-to-right sum.
template <typename VT, typename XT, int UNR>
__device__ void spmm_item(const BatchDev<VT, XT> &d, const SpmmGroup<VT> &G, int blk, const XT *__restrict__ xin,
                          XT *__restrict__ xout, double *red) {
#pragma clang fp contract(off)
  using AT = typename dev::Val<VT>::A;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int sub = lane >> 4, q = lane & 15;
  const int rows = G.rows, rp0 = G.rp0, f_col = G.f_col, n_rhs = G.n_rhs, rhs0 = G.rhs_base + 4 * q;
  const int *__restrict__ rowptr = G.rowptr;
  const int *__restrict__ colidx = G.colidx;
  const VT *__restrict__ vals = G.vals;
  const unsigned long long *__restrict__ member = G.member;
  bool act[4];
  XT lmax[4];  // maxima of the stored (XT) iterate values
  bool any = false;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    act[t] = 4 * q + t < n_rhs && !d.rs[rhs0 + t].done;
    lmax[t] = -INFINITY;
    any = any || act[t];
  }
  const XT *X = xin + G.xoff;
  XT *Y = xout + G.xoff;
  if (__ballot(any) != 0) {
    const int rpw = d.spmm_rpw;
    const int r0 = (blk * 4 + wave) * rpw;
    // bounds of the wave's rows, loaded once: lane l <= 16 holds rowptr[r0 + l]
    // (clamped to rows, an index rowptr has)
    const int rpl = lane <= rpw ? rowptr[min(r0 + lane, rows)] - rp0 : 0;
    // the next pass's row bounds, first 16 column indices and values, and
    // membership word, read while the current pass's gathers are in flight
    int a0n, a1n, cqn = 0;
    AT vqn = 0;
    unsigned long long mbn = 0;
    auto ahead = [&](int pass) {
      const int i = pass * 4 + sub, row = r0 + i;
      const bool rv = row < rows;
      a0n = __shfl(rpl, i, 64);
      a1n = __shfl(rpl, i + 1, 64);
      if (!rv) a0n = a1n = 0;
      cqn = 0;
      vqn = 0;
      if (q < min(16, a1n - a0n)) {
        cqn = colidx[a0n + q] - f_col;
        vqn = dev::Val<VT>::get(vals, a0n + q);
      }
      mbn = rv ? member[row] : 0ull;
    };
    ahead(0);
#pragma unroll 1
    for (int pass = 0; pass < rpw / 4; pass++) {
      const int row = r0 + pass * 4 + sub;
      const bool rv = row < rows;
      const int a0 = a0n, a1 = a1n;
      int cq = cqn;
      AT vq = vqn;
      const unsigned long long mbrow = mbn;
      if (pass + 1 < rpw / 4) ahead(pass + 1);
      AT acc[4] = {0, 0, 0, 0};
      for (int base = a0; __ballot(base < a1) != 0; base += 16) {
        const int cnt = min(16, a1 - base);  // <= 0 for a finished row
        if (base != a0) {  // rows longer than 16 entries: later chunks
          cq = 0;
          vq = 0;
          if (q < cnt) {
            cq = colidx[base + q] - f_col;
            vq = dev::Val<VT>::get(vals, base + q);
          }
        }
        // entry e of the chunk sits on lane e of the row's 16 lanes: its
        // column and value are row broadcasts; the UNR gathers of a group
        // are unconditional (entries past the row's end, and lanes whose
        // sets are all finished, read row 0) so they issue back to back
        static_for<0, 16 / UNR>([&](auto G) {
          AT xv[UNR][4], vv[UNR];
          static_for<0, UNR>([&](auto U) {
            constexpr int e = decltype(G)::value * UNR + decltype(U)::value;
            const int c = row_bcast<e>(cq);
            vv[U] = row_bcast<e>(vq);
            const int ce = e < cnt && any ? c : 0;
            XT t4[4];
            Quad<XT>::load(X + static_cast<long>(ce) * kLanes + 4 * q, t4);
#pragma unroll
            for (int t = 0; t < 4; t++) xv[U][t] = static_cast<AT>(t4[t]);
          });
          // (an entry past the row's end has value 0 -- its lane loaded no
          // value -- and adds a zero: the sums are unchanged)
#pragma unroll
          for (int u = 0; u < UNR; u++)
#pragma unroll
            for (int t = 0; t < 4; t++) acc[t] += vv[u] * xv[u][t];
        });
      }
      if (rv && any) {
        const unsigned long long mb = mbrow >> (4 * q);
        XT o[4];
        bool all = true;
#pragma unroll
        for (int t = 0; t < 4; t++) {
          double bi = d.beta * static_cast<double>(acc[t]);
          if ((mb >> t) & 1ull) bi = bi + d.rs[rhs0 + t].tele;  // teleport rows are few: read here
          o[t] = static_cast<XT>(bi);
          if (act[t]) lmax[t] = max(lmax[t], o[t]);
          all = all && act[t];
        }
        XT *yp = Y + static_cast<long>(row) * kLanes + 4 * q;
        if (all) {
          Quad<XT>::store(yp, o);
        } else {
#pragma unroll
          for (int t = 0; t < 4; t++)
            if (act[t]) yp[t] = o[t];  // a finished RHS keeps its final iterate
        }
      }
    }
  }
  // maxima: the four row groups of the wave, then the four wavefronts
#pragma unroll
  for (int t = 0; t < 4; t++) {
    lmax[t] = max(lmax[t], __shfl_xor(lmax[t], 16, 64));
    lmax[t] = max(lmax[t], __shfl_xor(lmax[t], 32, 64));
  }
  if (sub == 0)
#pragma unroll
    for (int t = 0; t < 4; t++) red[wave * 64 + 4 * q + t] = static_cast<double>(lmax[t]);
  __syncthreads();
  if (wave == 0) {
    const int j = lane;  // RHS j of the group
    const double m = fmax(fmax(red[j], red[64 + j]), fmax(red[128 + j], red[192 + j]));
    if (j < n_rhs && m > -INFINITY && !d.rs[G.rhs_base + j].done)
      atomicMax(d.slots + (G.rhs_base + j) * kSlots + (blk % kSlots), ord_encode(m));
  }
  __syncthreads();
}

// Persistent step: every workgroup walks its queue with a static stride
// (queue = XCD of the workgroup when xcd_map is set, else one shared list).
// (sixteen gathers in flight fit two waves per SIMD in FP32 only with the
// register budget held to 256)
template <typename VT, typename XT, int UNR, int TN>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(UNR == 16 ? 2 : 1)))
void batch_step(BatchDev<VT, XT> d, int par) {
  using AT = typename dev::Val<VT>::A;
  __shared__ dev::CsrTileLds<AT, TN> lds;
  __shared__ double red[kTileThreads];
  const XT *xa_in = par ? d.xa1 : d.xa0, *xb_in = par ? d.xb1 : d.xb0;
  XT *xa_out = par ? d.xa0 : d.xa1, *xb_out = par ? d.xb0 : d.xb1;
  int q, local, stride;
  if (d.xcd_map) {
    q = blockIdx.x % kQueues;
    local = blockIdx.x / kQueues;
    stride = gridDim.x / kQueues;
  } else {
    q = 0;
    local = blockIdx.x;
    stride = gridDim.x;
  }
  const int q_end = d.q_off[q + 1];
  for (int j = d.q_off[q] + local; j < q_end; j += stride) {
    const int w = d.q_items[j];
    if (w < d.a_ntiles + d.a_nlong) {
      const bool tile = w < d.a_ntiles;
      int gA, gB;
      if (tile) {
        const int rlo = d.a.tile_row[w], rhi = d.a.tile_row[w + 1];
        if (rlo >= rhi) continue;
        gA = d.row_graph[rlo];
        gB = d.row_graph[rhi - 1];
        if (gA == gB && d.rs[gA].done) continue;
      } else {
        gA = gB = d.row_graph[d.a.long_rows[w - d.a_ntiles]];
        if (d.rs[gA].done) continue;
      }
      TileEpi<XT> epi{xa_out, d.row_graph, d.rs, d.a_member, d.slots, d.beta, gA, gB, w % kSlots,
                      -INFINITY, -INFINITY};
      if (tile)
        dev::csr_tile_rows_at<VT, XT, TileEpi<XT>, TN, kBatchLongRow>(d.a, xa_in, lds, epi, w);
      else
        dev::csr_long_row_at<VT, XT, TileEpi<XT>, TN, kBatchLongChunk>(d.a, xa_in, lds, epi, w - d.a_ntiles);
      const double bA = dev::block_tree_max<double>(epi.mA, red);
      const double bB = dev::block_tree_max<double>(epi.mB, red);
      if (threadIdx.x == 0) {
        if (bA > -INFINITY) atomicMax(d.slots + gA * kSlots + (w % kSlots), ord_encode(bA));
        if (gB != gA && bB > -INFINITY) atomicMax(d.slots + gB * kSlots + (w % kSlots), ord_encode(bB));
      }
    } else {
      int s = 0;
      while (s + 1 < d.n_groups && w >= d.g[s + 1].item_base) s++;
      spmm_item<VT, XT, UNR>(d, d.g[s], w - d.g[s].item_base, xb_in, xb_out, red);
    }
  }
}

// One wavefront per RHS: max over its slots (cleared for the next step).
__global__ __launch_bounds__(64) void batch_finalize_max(RhsState *rs, unsigned long long *slots) {
  const int r = blockIdx.x, l = threadIdx.x;
  if (rs[r].done) return;
  unsigned long long u = 0;
  if (l < kSlots) {
    u = __hip_atomic_load(slots + r * kSlots + l, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(slots + r * kSlots + l, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(u, off, 64);
    u = o > u ? o : u;
  }
  if (l == 0) rs[r].maxv = ord_decode(u);
}

// MODE 0: normalise the step's output (b /= max, residual, compensated sum).
// MODE 1: compensated sum of each RHS's final iterate.
// MODE 2: out = final iterate / total.
template <typename VT, typename XT, int MODE>
__global__ __launch_bounds__(kTileThreads) void batch_pass(BatchDev<VT, XT> d, const PassItem *__restrict__ items,
                                                           PassPart *__restrict__ part,
                                                           const long *__restrict__ out_off,
                                                           double *__restrict__ out, int par) {
#pragma clang fp contract(off)
  __shared__ double red_e[kTileThreads], red_s[kTileThreads], red_c[kTileThreads];
  const PassItem it = items[blockIdx.x];
  if (it.kind == 0) {
    const int g = it.id;
    const RhsState &st = d.rs[g];
    if (MODE == 0 && st.done) return;
    XT *cur = MODE == 0 ? (par ? d.xa0 : d.xa1) : ((st.it & 1) ? d.xa1 : d.xa0);
    const XT *prev = par ? d.xa1 : d.xa0;
    const double m = st.maxv, S = st.total;
    double err = 0;
    CSum cs{0, 0};
    const long o = out_off[g] - d.row_base[g];
    for (int i = it.r0 + threadIdx.x; i < it.r1; i += kTileThreads) {
      if (MODE == 0) {
        const XT bs = static_cast<XT>(static_cast<double>(cur[i]) / m);
        const double bi = static_cast<double>(bs);
        err += fabs(bi - static_cast<double>(prev[i]));
        csum_add(cs, bi);
        cur[i] = bs;
      } else if (MODE == 1) {
        csum_add(cs, static_cast<double>(cur[i]));
      } else {
        out[o + i] = static_cast<double>(cur[i]) / S;
      }
    }
    if (MODE == 2) return;
    err = dev::block_tree_sum<double>(err, red_e);
    // compensated merge across the block, fixed order
    red_s[threadIdx.x] = cs.s;
    red_c[threadIdx.x] = cs.c;
    __syncthreads();
    for (int k = kTileThreads / 2; k > 0; k >>= 1) {
      if (static_cast<int>(threadIdx.x) < k) {
        const CSum mm = csum_merge(CSum{red_s[threadIdx.x], red_c[threadIdx.x]},
                                   CSum{red_s[threadIdx.x + k], red_c[threadIdx.x + k]});
        red_s[threadIdx.x] = mm.s;
        red_c[threadIdx.x] = mm.c;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[it.part] = PassPart{err, red_s[0], red_c[0]};
    return;
  }
  // SpMM group: wave w takes rows r0 + w, r0 + w + 4, ...; lane = RHS
  const SpmmGroup<VT> &G = d.g[it.id];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int rhs = G.rhs_base + lane;
  const bool valid = lane < G.n_rhs;
  const RhsState *st = valid ? &d.rs[rhs] : nullptr;
  const bool active = valid && (MODE != 0 || !st->done);
  double err = 0;
  CSum cs{0, 0};
  if (active) {
    XT *cur = (MODE == 0 ? (par ? d.xb0 : d.xb1) : ((st->it & 1) ? d.xb1 : d.xb0)) + G.xoff;
    const XT *prev = (par ? d.xb1 : d.xb0) + G.xoff;
    const double m = st->maxv, S = st->total;
    for (int i = it.r0 + wave; i < it.r1; i += 4) {
      const long k = static_cast<long>(i) * kLanes + lane;
      if (MODE == 0) {
        const XT bs = static_cast<XT>(static_cast<double>(cur[k]) / m);
        const double bi = static_cast<double>(bs);
        err += fabs(bi - static_cast<double>(prev[k]));
        csum_add(cs, bi);
        cur[k] = bs;
      } else if (MODE == 1) {
        csum_add(cs, static_cast<double>(cur[k]));
      } else {
        out[out_off[rhs] + i] = static_cast<double>(cur[k]) / S;
      }
    }
  }
  if (MODE == 2) return;
  red_e[threadIdx.x] = err;
  red_s[threadIdx.x] = cs.s;
  red_c[threadIdx.x] = cs.c;
  __syncthreads();
  if (wave == 0) {
    double e = red_e[lane];
    CSum t{red_s[lane], red_c[lane]};
    for (int w = 1; w < 4; w++) {
      e += red_e[w * 64 + lane];
      t = csum_merge(t, CSum{red_s[w * 64 + lane], red_c[w * 64 + lane]});
    }
    if (active) part[it.part + lane] = PassPart{e, t.s, t.c};
  }
}

// One wavefront per RHS: merge its partials in a fixed order.  FINAL = 0:
// the step decision (reference residual or sum-normalised bound) and the
// next teleport term; FINAL = 1: the total of the final iterate.
template <int FINAL>
__global__ __launch_bounds__(64) void batch_finalize_norm(RhsState *rs, const RhsInfo *__restrict__ info,
                                                          const PassPart *__restrict__ part, double eps,
                                                          double bound_target, int rule, int max_it) {
  const int r = blockIdx.x, l = threadIdx.x;
  RhsState &st = rs[r];
  if (!FINAL && st.done) return;
  const RhsInfo in = info[r];
  double e = 0;
  CSum t{0, 0};
  for (int k = l; k < in.count; k += 64) {
    const PassPart p = part[in.base + static_cast<long>(k) * in.stride];
    e += p.e;
    t = csum_merge(t, CSum{p.s, p.c});
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    const CSum o{__shfl_xor(t.s, off, 64), __shfl_xor(t.c, off, 64)};
    // the lower lane keeps the left operand so both lanes compute the same value
    t = (l & off) ? csum_merge(o, t) : csum_merge(t, o);
  }
  if (l != 0) return;
  const double S = t.s + t.c;
  if (FINAL) {
    st.total = S;
    return;
  }
  st.it += 1;
  st.err = e;
  st.sum_prev = st.sum;
  st.sum = S;
  st.bound = e / S + fabs(st.sum_prev / S - 1.0);
  const bool converged = rule == 0 ? (e < eps) : (st.bound <= bound_target);
  st.done = (converged || st.it >= max_it) ? 1 : 0;
  st.tele = S * st.tele_scale;
}

template <typename T>
__global__ void batch_fill(T *p, long n, T v) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * 256) p[i] = v;
}

unsigned fill_grid(long n) { return static_cast<unsigned>(std::max<long>(1, std::min<long>(4096, (n + 255) / 256))); }

}  // namespace

std::string BatchConfig::str() const {
  std::ostringstream o;
  o << "wg_per_cu=" << wg_per_cu << " unroll=" << unroll << " xcd_map=" << xcd_map << " tile_nnz=" << tile_nnz
    << " spmm_rows=" << spmm_rows;
  return o.str();
}

struct PageRankBatch::Impl {
  Dtype prec;
  struct Graph {
    const cpp::MatrixCSR *m;
    std::vector<std::vector<int>> rhs;
  };
  std::vector<Graph> graphs;
  bool prepared = false;
  int cus = 256;
  // single-RHS block-diagonal matrix
  DevBuf a_rowptr, a_colidx, a_vals, a_long, row_graph, row_base, a_member, xa[2];
  DevBuf a_tiles[3];  // tile tables by tile size (kBatchTiles)
  int a_ntiles[3] = {0, 0, 0};
  int a_rows = 0, a_nlong = 0, n_a = 0;
  long a_nnz = 0;
  // SpMM graphs and groups
  struct SGraph {
    DevBuf rowptr, colidx, vals;
    int rows = 0, rp0 = 0;
    long nnz = 0;
  };
  std::vector<std::unique_ptr<SGraph>> sgraphs;
  struct Group {
    int sgraph, n_rhs, rhs_base;
    long xoff;
    DevBuf member;
  };
  std::vector<std::unique_ptr<Group>> groups;
  DevBuf xb[2];
  long xb_len = 0;
  // RHS bookkeeping (device RHS order: single-RHS graphs, then SpMM lanes)
  int n_rhs_total = 0;
  std::vector<long> user_offsets;    // output offset per RHS in user order (+ total)
  std::vector<int> user_to_dev;      // device RHS id of each user RHS
  std::vector<int> dev_rows;         // rows of each device RHS's graph
  std::vector<double> dev_nodes;     // |nodes| of each device RHS
  DevBuf rs, slots, info, parts, items, out_off;
  int n_items = 0;
  long n_parts = 0;
  // The work items of one step for one (tile size, SpMM rows per wavefront):
  // the tiles, then the long rows, then each group's SpMM items; queued in
  // item order (xcd_map 0) or pinned to XCD queues (xcd_map 1).
  struct Work {
    DevBuf group_desc;  // SpmmGroup<VT> array with this layout's item bases
    DevBuf q_items[2], q_off_dev[2];
    int n_work = 0;
  };
  Work work[3][3];
  int n_work = 0;  // (of the default layout: zero means no work at all)
  BatchConfig cfg;
  bool tuned = false;
  double bytes_step = 0, bytes_pass = 0;
  // the key of the autotuner's cache: precision, device, the single-set
  // graphs' rows, entries and log2 row-length histogram, each multi-set
  // graph's rows, entries and set count
  std::string signature;

  std::size_t vsize() const { return prec == Dtype::F64 ? 8 : 4; }

  hipStream_t stream = nullptr;  // private stream: capture needs a non-default stream
  ~Impl() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <typename VT, typename XT>
  BatchDev<VT, XT> dev_view(const BatchConfig &c, double beta) const;
  template <typename VT, typename XT>
  void enqueue_iteration(const BatchDev<VT, XT> &d, int par, const BatchParams &p, hipStream_t s) const;
  template <typename VT, typename XT>
  void enqueue_rest(const BatchDev<VT, XT> &d, int par, const BatchParams &p, hipStream_t s) const;
  template <typename VT, typename XT>
  void launch_step(const BatchDev<VT, XT> &d, int par, const BatchConfig &c, hipStream_t s) const;
  template <typename VT, typename XT>
  BatchReport run_typed(const BatchParams &p, double *out, hipStream_t s);
  void build();
  void init_state(const BatchParams &p, hipStream_t s);
  void build_work(int ti, int ri, const std::vector<int> &rp_all, const std::vector<int> &rg);
  template <typename VT>
  void upload_groups(Work &w, const std::vector<int> &item_base, const std::vector<int> &nitems) {
    std::vector<SpmmGroup<VT>> gd;
    for (std::size_t k = 0; k < groups.size(); k++) {
      const Group &gp = *groups[k];
      const SGraph &sg = *sgraphs[static_cast<std::size_t>(gp.sgraph)];
      gd.push_back(SpmmGroup<VT>{sg.rowptr.as<int>(), sg.colidx.as<int>(), sg.vals.as<VT>(),
                                 gp.member.as<unsigned long long>(), sg.rows, sg.rp0, 0, gp.n_rhs, gp.rhs_base,
                                 item_base[k], nitems[k], 0, gp.xoff});
    }
    if (!gd.empty()) w.group_desc.upload(gd.data(), gd.size() * sizeof(SpmmGroup<VT>));
  }
};

namespace {

// The autotuner's choices by batch signature, for the life of the process.
std::mutex &tune_cache_mu() {
  static std::mutex mu;
  return mu;
}
std::map<std::string, BatchConfig> &tune_cache() {
  static std::map<std::string, BatchConfig> c;
  return c;
}

void upload_converted(DevBuf &dst, const double *src, std::size_t n, Dtype prec) {
  if (prec == Dtype::F64) {
    dst.upload(src, n * sizeof(double));
  } else {
    std::vector<float> t(n);
    for (std::size_t i = 0; i < n; i++) t[i] = static_cast<float>(src[i]);
    dst.upload(t.data(), n * sizeof(float));
  }
}

}  // namespace

// The work items and queues of one layout.  Weights are the bytes an item
// reads.  XCD-pinned queues: SpMM row ranges split evenly, single-RHS graphs
// whole, balanced by weight.
void PageRankBatch::Impl::build_work(int ti, int ri, const std::vector<int> &rp_all, const std::vector<int> &rg) {
  Work &w = work[ti][ri];
  const int tn = kBatchTiles[ti], rows_item = 4 * kSpmmRpw[ri];
  const int ntiles = a_rows > 0 ? a_ntiles[ti] : 0;
  const double vb = static_cast<double>(vsize());
  std::vector<int> item_base(groups.size()), nitems(groups.size());
  int nw = ntiles + a_nlong;
  for (std::size_t k = 0; k < groups.size(); k++) {
    const int rows = sgraphs[static_cast<std::size_t>(groups[k]->sgraph)]->rows;
    item_base[k] = nw;
    nitems[k] = (rows + rows_item - 1) / rows_item;
    nw += nitems[k];
  }
  w.n_work = nw;
  std::vector<double> wgt(static_cast<std::size_t>(nw), 0.0);
  std::vector<int> item_graph(static_cast<std::size_t>(nw), -1);
  if (a_rows > 0) {
    std::vector<int> t(static_cast<std::size_t>(ntiles) + 1);
    TBSLA_HIP_CHECK(hipMemcpy(t.data(), a_tiles[ti].get(), t.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < ntiles; k++) {
      const long e0 = static_cast<long>(k) * tn;
      wgt[static_cast<std::size_t>(k)] = static_cast<double>(std::min<long>(tn, a_nnz - e0)) * (vb + 4 + vb);
      const int r = std::min(t[static_cast<std::size_t>(k)], a_rows - 1);
      item_graph[static_cast<std::size_t>(k)] = rg[static_cast<std::size_t>(r)];
    }
    std::vector<int> lr(static_cast<std::size_t>(a_nlong));
    if (a_nlong) TBSLA_HIP_CHECK(hipMemcpy(lr.data(), a_long.get(), lr.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int k = 0; k < a_nlong; k++) {
      const int r = lr[static_cast<std::size_t>(k)];
      wgt[static_cast<std::size_t>(ntiles + k)] =
          static_cast<double>(rp_all[static_cast<std::size_t>(r) + 1] - rp_all[static_cast<std::size_t>(r)]) * (vb + 4 + vb);
      item_graph[static_cast<std::size_t>(ntiles + k)] = rg[static_cast<std::size_t>(r)];
    }
  }
  for (std::size_t k = 0; k < groups.size(); k++) {
    const SGraph &sg = *sgraphs[static_cast<std::size_t>(groups[k]->sgraph)];
    const double per_row = static_cast<double>(sg.nnz) / std::max(1, sg.rows) * (vb + 4 + kLanes * vb) + 2.0 * kLanes * vb;
    for (int b = 0; b < nitems[k]; b++) wgt[static_cast<std::size_t>(item_base[k] + b)] = per_row * rows_item;
  }
  std::vector<int> q_off[2];
  // interleaved: one list in item order
  {
    std::vector<int> all(static_cast<std::size_t>(nw));
    std::iota(all.begin(), all.end(), 0);
    q_off[0].assign(kQueues + 1, nw);
    q_off[0][0] = 0;
    if (nw) w.q_items[0].upload(all.data(), all.size() * sizeof(int));
  }
  // XCD-pinned
  {
    std::vector<std::vector<int>> q(kQueues);
    std::vector<double> load(kQueues, 0.0);
    for (std::size_t k = 0; k < groups.size(); k++)
      for (int x = 0; x < kQueues; x++) {
        const int b0 = static_cast<int>(static_cast<long>(nitems[k]) * x / kQueues);
        const int b1 = static_cast<int>(static_cast<long>(nitems[k]) * (x + 1) / kQueues);
        for (int b = b0; b < b1; b++) {
          q[static_cast<std::size_t>(x)].push_back(item_base[k] + b);
          load[static_cast<std::size_t>(x)] += wgt[static_cast<std::size_t>(item_base[k] + b)];
        }
      }
    std::vector<std::vector<int>> by_graph(static_cast<std::size_t>(std::max(1, n_a)));
    std::vector<double> gw(static_cast<std::size_t>(std::max(1, n_a)), 0.0);
    for (int k = 0; k < ntiles + a_nlong; k++) {
      const int g = item_graph[static_cast<std::size_t>(k)];
      by_graph[static_cast<std::size_t>(g)].push_back(k);
      gw[static_cast<std::size_t>(g)] += wgt[static_cast<std::size_t>(k)];
    }
    std::vector<int> order(static_cast<std::size_t>(n_a));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return gw[static_cast<std::size_t>(x)] > gw[static_cast<std::size_t>(y)]; });
    for (int g : order) {
      const int x = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
      for (int k : by_graph[static_cast<std::size_t>(g)]) q[static_cast<std::size_t>(x)].push_back(k);
      load[static_cast<std::size_t>(x)] += gw[static_cast<std::size_t>(g)];
    }
    std::vector<int> flat;
    q_off[1].assign(1, 0);
    for (int x = 0; x < kQueues; x++) {
      flat.insert(flat.end(), q[static_cast<std::size_t>(x)].begin(), q[static_cast<std::size_t>(x)].end());
      q_off[1].push_back(static_cast<int>(flat.size()));
    }
    if (!flat.empty()) w.q_items[1].upload(flat.data(), flat.size() * sizeof(int));
  }
  for (int k = 0; k < 2; k++) w.q_off_dev[k].upload(q_off[k].data(), q_off[k].size() * sizeof(int));
  if (prec == Dtype::F64)
    upload_groups<double>(w, item_base, nitems);
  else
    upload_groups<float>(w, item_base, nitems);
}

void PageRankBatch::Impl::build() {
  int dev = 0;
  TBSLA_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  TBSLA_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  cus = std::max(8, prop.multiProcessorCount);
  const std::size_t vb = vsize();
  // ---- device RHS numbering
  std::vector<int> single, multi;
  for (std::size_t g = 0; g < graphs.size(); g++) (graphs[g].rhs.size() == 1 ? single : multi).push_back(static_cast<int>(g));
  n_a = static_cast<int>(single.size());
  std::vector<std::vector<int>> user_rhs_dev(graphs.size());
  int next = 0;
  for (int g : single) user_rhs_dev[static_cast<std::size_t>(g)].push_back(next++);
  for (int g : multi)
    for (std::size_t k = 0; k < graphs[static_cast<std::size_t>(g)].rhs.size(); k++)
      user_rhs_dev[static_cast<std::size_t>(g)].push_back(next++);
  n_rhs_total = next;
  {
    std::ostringstream sig;
    sig << (prec == Dtype::F64 ? "f64" : "f32") << "/" << prop.gcnArchName << "/" << cus;
    long rows_a = 0, nnz_a = 0;
    std::vector<long> hist(32, 0);
    for (int g : single) {
      const auto &rp = graphs[static_cast<std::size_t>(g)].m->get_rowptr();
      rows_a += static_cast<long>(rp.size()) - 1;
      nnz_a += rp.empty() ? 0 : static_cast<long>(rp.back() - rp.front());
      for (std::size_t i = 0; i + 1 < rp.size(); i++) {
        int b = 0;
        for (unsigned len = static_cast<unsigned>(std::max(0, rp[i + 1] - rp[i])); len; len >>= 1) b++;
        hist[static_cast<std::size_t>(b)]++;
      }
    }
    sig << "|single " << single.size() << " " << rows_a << " " << nnz_a << " h";
    for (long h : hist) sig << " " << h;
    for (int g : multi) {
      const auto &rp = graphs[static_cast<std::size_t>(g)].m->get_rowptr();
      sig << "|multi " << (rp.size() - 1) << " " << (rp.empty() ? 0 : rp.back() - rp.front()) << " "
          << graphs[static_cast<std::size_t>(g)].rhs.size();
    }
    signature = sig.str();
  }
  dev_rows.assign(static_cast<std::size_t>(n_rhs_total), 0);
  dev_nodes.assign(static_cast<std::size_t>(n_rhs_total), 0);
  user_offsets.clear();
  user_to_dev.clear();
  long off = 0;
  std::vector<long> dev_out(static_cast<std::size_t>(n_rhs_total), 0);
  for (std::size_t g = 0; g < graphs.size(); g++) {
    const long n = graphs[g].m->get_n_row();
    for (std::size_t k = 0; k < graphs[g].rhs.size(); k++) {
      const int r = user_rhs_dev[g][k];
      user_offsets.push_back(off);
      user_to_dev.push_back(r);
      dev_out[static_cast<std::size_t>(r)] = off;
      dev_rows[static_cast<std::size_t>(r)] = static_cast<int>(n);
      dev_nodes[static_cast<std::size_t>(r)] = static_cast<double>(graphs[g].rhs[k].size());
      off += n;
    }
  }
  user_offsets.push_back(off);

  auto csr_of = [&](int g, std::vector<int> &rp, const int **ci, const double **va, long &nnz) {
    const cpp::MatrixCSR &m = *graphs[static_cast<std::size_t>(g)].m;
    const auto &r = m.get_rowptr();
    const int rows = m.get_n_row();
    if (static_cast<int>(r.size()) != rows + 1) throw std::invalid_argument("tbsla: batch graph rowptr does not cover every row");
    rp.assign(r.begin(), r.end());
    nnz = static_cast<long>(rp.back()) - rp[0];
    for (int i = 0; i < rows; i++)
      if (rp[static_cast<std::size_t>(i) + 1] < rp[static_cast<std::size_t>(i)]) throw std::invalid_argument("tbsla: CSR rowptr is not monotone");
    // entry k of the graph is colidx/values[rowptr[0] + k] (MatrixCSR.cpp:135)
    if (rp[0] < 0 || static_cast<long>(m.get_colidx().size()) < rp[0] + nnz ||
        static_cast<long>(m.get_values().size()) < rp[0] + nnz)
      throw std::invalid_argument("tbsla: CSR rowptr points past values/colidx");
    *ci = m.get_colidx().data() + rp[0];
    *va = m.get_values().data() + rp[0];
    for (long k = 0; k < nnz; k++) {
      const int c = (*ci)[k] - m.get_f_col();
      if (c < 0 || c >= m.get_n_col()) throw std::invalid_argument("tbsla: column index outside the graph");
    }
  };

  // ---- single-RHS graphs: one block-diagonal CSR
  {
    std::vector<int> rp_all{0}, ci_all, rg, rb;
    std::vector<double> va_all;
    std::vector<unsigned char> mem;
    for (int gi = 0; gi < n_a; gi++) {
      const int g = single[static_cast<std::size_t>(gi)];
      std::vector<int> rp;
      const int *ci;
      const double *va;
      long nnz;
      csr_of(g, rp, &ci, &va, nnz);
      const int rows = static_cast<int>(rp.size()) - 1;
      const long base_row = static_cast<long>(rg.size());
      const long base_e = static_cast<long>(ci_all.size());
      if (base_row + rows > (1L << 31) - 2 || base_e + nnz > (1L << 31) - 2)
        throw std::invalid_argument("tbsla: batch exceeds 2^31 rows or entries");
      rb.push_back(static_cast<int>(base_row));
      const int f_col = graphs[static_cast<std::size_t>(g)].m->get_f_col();
      for (long k = 0; k < nnz; k++) {
        ci_all.push_back(static_cast<int>(base_row + ci[k] - f_col));
        va_all.push_back(va[k]);
      }
      for (int i = 0; i < rows; i++) {
        rp_all.push_back(static_cast<int>(base_e + rp[static_cast<std::size_t>(i) + 1] - rp[0]));
        rg.push_back(gi);
      }
      std::vector<unsigned char> m(static_cast<std::size_t>(rows), 0);
      for (int node : graphs[static_cast<std::size_t>(g)].rhs[0])
        if (node >= 0 && node < rows) m[static_cast<std::size_t>(node)] = 1;
      mem.insert(mem.end(), m.begin(), m.end());
    }
    a_rows = static_cast<int>(rg.size());
    a_nnz = static_cast<long>(ci_all.size());
    if (a_rows > 0) {
      a_rowptr.upload(rp_all.data(), rp_all.size() * sizeof(int));
      a_colidx.upload(ci_all.data(), ci_all.size() * sizeof(int));
      upload_converted(a_vals, va_all.data(), va_all.size(), prec);
      row_graph.upload(rg.data(), rg.size() * sizeof(int));
      row_base.upload(rb.data(), rb.size() * sizeof(int));
      a_member.upload(mem.data(), mem.size());
      for (int ti = 0; ti < 3; ti++) {
        std::vector<int> t = csr_entry_tiles(rp_all.data(), a_rows, kBatchTiles[ti]);
        a_ntiles[ti] = static_cast<int>(t.size()) - 1;
        a_tiles[ti].upload(t.data(), t.size() * sizeof(int));
      }
      std::vector<int> lr;
      for (int i = 0; i < a_rows; i++)
        if (rp_all[static_cast<std::size_t>(i) + 1] - rp_all[static_cast<std::size_t>(i)] > kBatchLongRow) lr.push_back(i);
      a_nlong = static_cast<int>(lr.size());
      if (a_nlong) a_long.upload(lr.data(), lr.size() * sizeof(int));
      for (int k = 0; k < 2; k++) xa[k].alloc(vb * static_cast<std::size_t>(a_rows));
    }
  }

  // ---- multi-RHS graphs: one CSR each, groups of up to 64 RHS
  xb_len = 0;
  for (int g : multi) {
    auto sg = std::unique_ptr<SGraph>(new SGraph());
    std::vector<int> rp;
    const int *ci;
    const double *va;
    long nnz;
    csr_of(g, rp, &ci, &va, nnz);
    sg->rows = static_cast<int>(rp.size()) - 1;
    sg->rp0 = rp[0];
    sg->nnz = nnz;
    sg->rowptr.upload(rp.data(), rp.size() * sizeof(int));
    const int f_col = graphs[static_cast<std::size_t>(g)].m->get_f_col();
    std::vector<int> cc(static_cast<std::size_t>(nnz));
    for (long k = 0; k < nnz; k++) cc[static_cast<std::size_t>(k)] = ci[k] - f_col;
    sg->colidx.upload(cc.data(), cc.size() * sizeof(int));
    upload_converted(sg->vals, va, static_cast<std::size_t>(nnz), prec);
    const auto &rhs = graphs[static_cast<std::size_t>(g)].rhs;
    for (std::size_t k0 = 0; k0 < rhs.size(); k0 += kLanes) {
      if (static_cast<int>(groups.size()) >= kMaxGroups)
        throw std::invalid_argument("tbsla: batch supports at most 64 groups of 64 right-hand sides");
      auto gr = std::unique_ptr<Group>(new Group());
      gr->sgraph = static_cast<int>(sgraphs.size());
      gr->n_rhs = static_cast<int>(std::min<std::size_t>(kLanes, rhs.size() - k0));
      gr->rhs_base = user_rhs_dev[static_cast<std::size_t>(g)][k0];
      gr->xoff = xb_len;
      xb_len += static_cast<long>(sg->rows) * kLanes;
      std::vector<unsigned long long> mb(static_cast<std::size_t>(sg->rows), 0ull);
      for (int j = 0; j < gr->n_rhs; j++)
        for (int node : rhs[k0 + static_cast<std::size_t>(j)])
          if (node >= 0 && node < sg->rows) mb[static_cast<std::size_t>(node)] |= 1ull << j;
      gr->member.upload(mb.data(), mb.size() * sizeof(unsigned long long));
      groups.push_back(std::move(gr));
    }
    sgraphs.push_back(std::move(sg));
  }
  if (xb_len > 0)
    for (int k = 0; k < 2; k++) xb[k].alloc(vb * static_cast<std::size_t>(xb_len));

  // ---- pass items and partial layout
  std::vector<PassItem> it;
  std::vector<RhsInfo> inf(static_cast<std::size_t>(n_rhs_total));
  long np = 0;
  {
    std::vector<int> rbh;
    long base = 0;
    for (int gi = 0; gi < n_a; gi++) {
      const int rows = graphs[static_cast<std::size_t>(single[static_cast<std::size_t>(gi)])].m->get_n_row();
      inf[static_cast<std::size_t>(gi)] = RhsInfo{np, 0, 1};
      for (int r0 = 0; r0 < rows; r0 += kChunkRows) {
        it.push_back(PassItem{0, gi, static_cast<int>(base + r0), static_cast<int>(base + std::min(rows, r0 + kChunkRows)), np});
        np++;
        inf[static_cast<std::size_t>(gi)].count++;
      }
      base += rows;
    }
    for (std::size_t s = 0; s < groups.size(); s++) {
      const Group &gr = *groups[s];
      const int rows = sgraphs[static_cast<std::size_t>(gr.sgraph)]->rows;
      const long first = np;
      int cnt = 0;
      for (int r0 = 0; r0 < rows; r0 += kSpmmChunkRows) {
        it.push_back(PassItem{1, static_cast<int>(s), r0, std::min(rows, r0 + kSpmmChunkRows), np});
        np += kLanes;
        cnt++;
      }
      for (int j = 0; j < gr.n_rhs; j++) inf[static_cast<std::size_t>(gr.rhs_base + j)] = RhsInfo{first + j, cnt, kLanes};
    }
  }
  n_items = static_cast<int>(it.size());
  n_parts = np;
  if (n_items) items.upload(it.data(), it.size() * sizeof(PassItem));
  info.upload(inf.data(), inf.size() * sizeof(RhsInfo));
  parts.alloc(sizeof(PassPart) * static_cast<std::size_t>(std::max(1L, np)));
  out_off.upload(dev_out.data(), dev_out.size() * sizeof(long));
  rs.alloc(sizeof(RhsState) * static_cast<std::size_t>(std::max(1, n_rhs_total)));
  slots.alloc(sizeof(unsigned long long) * kSlots * static_cast<std::size_t>(std::max(1, n_rhs_total)));

  // ---- work items and queues of every (tile size, SpMM rows) layout
  {
    std::vector<int> rp_all, rg;
    if (a_rows > 0) {
      rp_all.resize(static_cast<std::size_t>(a_rows) + 1);
      rg.resize(static_cast<std::size_t>(a_rows));
      TBSLA_HIP_CHECK(hipMemcpy(rp_all.data(), a_rowptr.get(), rp_all.size() * sizeof(int), hipMemcpyDeviceToHost));
      TBSLA_HIP_CHECK(hipMemcpy(rg.data(), row_graph.get(), rg.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    for (int ti = 0; ti < 3; ti++)
      for (int ri = 0; ri < 3; ri++) build_work(ti, ri, rp_all, rg);
    n_work = work[1][1].n_work;
  }
  // algorithmic bytes of one iteration with every RHS active
  bytes_step = static_cast<double>(a_nnz) * (vb + 4) + static_cast<double>(a_rows) * (4 + 1 + 2 * vb);
  bytes_pass = 3.0 * vb * static_cast<double>(a_rows);
  for (const auto &gp : groups) {
    const SGraph &sg = *sgraphs[static_cast<std::size_t>(gp->sgraph)];
    bytes_step += static_cast<double>(sg.nnz) * (vb + 4) + static_cast<double>(sg.rows) * (4 + 8 + 2.0 * gp->n_rhs * vb);
    bytes_pass += 3.0 * vb * static_cast<double>(sg.rows) * gp->n_rhs;
  }
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  prepared = true;
}

template <typename VT, typename XT>
BatchDev<VT, XT> PageRankBatch::Impl::dev_view(const BatchConfig &c, double beta) const {
  const int ti = tile_index(c.tile_nnz), ri = rpw_index(c.spmm_rows);
  const Work &w = work[ti][ri];
  const int xcd_map = c.xcd_map ? 1 : 0;
  BatchDev<VT, XT> d;
  std::memset(&d, 0, sizeof(d));
  d.a = dev::CsrView<VT>{a_rowptr.as<int>(), a_colidx.as<int>(), a_vals.as<VT>(), a_tiles[ti].as<int>(),
                         a_long.as<int>(), a_nnz, 0, 0};
  d.a_ntiles = a_rows > 0 ? a_ntiles[ti] : 0;
  d.a_nlong = a_nlong;
  d.a_rows = a_rows;
  d.n_groups = static_cast<int>(groups.size());
  d.row_graph = row_graph.as<int>();
  d.row_base = row_base.as<int>();
  d.a_member = a_member.as<unsigned char>();
  d.xa0 = xa[0].as<XT>();
  d.xa1 = xa[1].as<XT>();
  d.xb0 = xb[0].as<XT>();
  d.xb1 = xb[1].as<XT>();
  d.g = w.group_desc.as<SpmmGroup<VT>>();
  d.rs = rs.as<RhsState>();
  d.slots = slots.as<unsigned long long>();
  d.q_items = w.q_items[xcd_map].as<int>();
  d.q_off = w.q_off_dev[xcd_map].as<int>();
  d.xcd_map = xcd_map;
  d.spmm_rpw = kSpmmRpw[ri];
  d.beta = beta;
  return d;
}

namespace {
template <typename VT, typename XT, int UNR>
void launch_step_tn(const BatchDev<VT, XT> &d, int par, int tn, unsigned grid, hipStream_t s) {
  if (tn == 1024)
    hipLaunchKernelGGL((batch_step<VT, XT, UNR, 1024>), dim3(grid), dim3(kTileThreads), 0, s, d, par);
  else if (tn == 3072)
    hipLaunchKernelGGL((batch_step<VT, XT, UNR, 3072>), dim3(grid), dim3(kTileThreads), 0, s, d, par);
  else
    hipLaunchKernelGGL((batch_step<VT, XT, UNR, 2048>), dim3(grid), dim3(kTileThreads), 0, s, d, par);
}
}  // namespace

template <typename VT, typename XT>
void PageRankBatch::Impl::launch_step(const BatchDev<VT, XT> &d, int par, const BatchConfig &c,
                                      hipStream_t s) const {
  if (n_work == 0) return;
  const unsigned grid = static_cast<unsigned>(std::max(kQueues, (cus * c.wg_per_cu) / kQueues * kQueues));
  const int tn = kBatchTiles[tile_index(c.tile_nnz)];
  if (c.unroll == 4)
    launch_step_tn<VT, XT, 4>(d, par, tn, grid, s);
  else if (c.unroll == 8)
    launch_step_tn<VT, XT, 8>(d, par, tn, grid, s);
  else
    launch_step_tn<VT, XT, 16>(d, par, tn, grid, s);
}

template <typename VT, typename XT>
void PageRankBatch::Impl::enqueue_iteration(const BatchDev<VT, XT> &d, int par, const BatchParams &p,
                                            hipStream_t s) const {
  launch_step<VT, XT>(d, par, cfg, s);
  enqueue_rest<VT, XT>(d, par, p, s);
}

template <typename VT, typename XT>
void PageRankBatch::Impl::enqueue_rest(const BatchDev<VT, XT> &d, int par, const BatchParams &p,
                                       hipStream_t s) const {
  hipLaunchKernelGGL(batch_finalize_max, dim3(n_rhs_total), dim3(64), 0, s, d.rs, d.slots);
  if (n_items)
    hipLaunchKernelGGL((batch_pass<VT, XT, 0>), dim3(n_items), dim3(kTileThreads), 0, s, d, items.as<PassItem>(),
                       parts.as<PassPart>(), out_off.as<long>(), static_cast<double *>(nullptr), par);
  hipLaunchKernelGGL(batch_finalize_norm<0>, dim3(n_rhs_total), dim3(64), 0, s, d.rs, info.as<RhsInfo>(),
                     parts.as<PassPart>(), p.epsilon, p.epsilon * (1.0 - p.beta) / p.beta, static_cast<int>(p.rule),
                     p.max_iteration);
}

void PageRankBatch::Impl::init_state(const BatchParams &p, hipStream_t s) {
  std::vector<RhsState> h(static_cast<std::size_t>(n_rhs_total));
  for (int r = 0; r < n_rhs_total; r++) {
    RhsState &st = h[static_cast<std::size_t>(r)];
    std::memset(&st, 0, sizeof(st));
    st.tele_scale = (1.0 - p.beta) / dev_nodes[static_cast<std::size_t>(r)];
    st.sum = 0.25 * static_cast<double>(dev_rows[static_cast<std::size_t>(r)]);
    st.tele = st.sum * st.tele_scale;
    st.bound = 2.0;
    st.done = p.max_iteration <= 0 ? 1 : 0;
  }
  TBSLA_HIP_CHECK(hipMemcpyAsync(rs.get(), h.data(), h.size() * sizeof(RhsState), hipMemcpyHostToDevice, s));
  if (prec == Dtype::F64) {
    if (a_rows) hipLaunchKernelGGL(batch_fill<double>, dim3(fill_grid(a_rows)), dim3(256), 0, s, xa[0].as<double>(), static_cast<long>(a_rows), 0.25);
    if (xb_len) hipLaunchKernelGGL(batch_fill<double>, dim3(fill_grid(xb_len)), dim3(256), 0, s, xb[0].as<double>(), xb_len, 0.25);
  } else {
    if (a_rows) hipLaunchKernelGGL(batch_fill<float>, dim3(fill_grid(a_rows)), dim3(256), 0, s, xa[0].as<float>(), static_cast<long>(a_rows), 0.25f);
    if (xb_len) hipLaunchKernelGGL(batch_fill<float>, dim3(fill_grid(xb_len)), dim3(256), 0, s, xb[0].as<float>(), xb_len, 0.25f);
  }
  const long ns = static_cast<long>(kSlots) * n_rhs_total;
  hipLaunchKernelGGL(batch_fill<unsigned long long>, dim3(fill_grid(ns)), dim3(256), 0, s, slots.as<unsigned long long>(), ns, 0ull);
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));  // the state copy reads host memory
}

template <typename VT, typename XT>
BatchReport PageRankBatch::Impl::run_typed(const BatchParams &p, double *out, hipStream_t s) {
  BatchReport rep;
  init_state(p, s);
  // The autotuner: a launch shape found for a batch of the same signature
  // (in this process) is reused; otherwise the first iterations of this run
  // are its trials, one candidate shape per iteration, each timed on its
  // step kernel.  Every shape computes the same bits, so the trial
  // iterations are ordinary iterations of the run.  The search is by
  // coordinates from the default shape: gathers in flight (4, 8, 16), then
  // workgroups per CU (2, 4, 8), then the queue mapping; the number of trial
  // iterations is made even so the iterate parity of the captured graph holds.
  int done_its = 0;
  if (p.autotune && !tuned && n_work > 0) {
    const auto t0 = std::chrono::steady_clock::now();
    {
      std::lock_guard<std::mutex> g(tune_cache_mu());
      auto it = tune_cache().find(signature);
      if (it != tune_cache().end()) {
        cfg = it->second;
        tuned = true;
        rep.tune_cached = true;
      }
    }
    if (!tuned) {
      hipEvent_t e0, e1;
      TBSLA_HIP_CHECK(hipEventCreate(&e0));
      TBSLA_HIP_CHECK(hipEventCreate(&e1));
      const BatchConfig base = cfg;
      BatchConfig best = base;
      float best_ms = 1e30f;
      std::vector<std::pair<BatchConfig, float>> tried;
      auto same = [](const BatchConfig &a, const BatchConfig &b) {
        return a.unroll == b.unroll && a.wg_per_cu == b.wg_per_cu && a.xcd_map == b.xcd_map &&
               a.tile_nnz == b.tile_nnz && a.spmm_rows == b.spmm_rows;
      };
      auto trial = [&](const BatchConfig &c, bool again) {
        if (!again)
          for (const auto &t : tried)
            if (same(t.first, c)) return;
        const BatchDev<VT, XT> d = dev_view<VT, XT>(c, p.beta);
        const int par = done_its & 1;
        TBSLA_HIP_CHECK(hipEventRecord(e0, s));
        launch_step<VT, XT>(d, par, c, s);
        TBSLA_HIP_CHECK(hipEventRecord(e1, s));
        enqueue_rest<VT, XT>(d, par, p, s);
        TBSLA_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        TBSLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        done_its++;
        tried.push_back({c, ms});
        if (ms < best_ms) {
          best_ms = ms;
          best = c;
        }
      };
      auto with = [](BatchConfig c, int BatchConfig::*f, int v) {
        c.*f = v;
        return c;
      };
      for (int un : {8, 4, 16}) trial(with(base, &BatchConfig::unroll, un), false);
      for (int wg : {2, 4, 8}) trial(with(best, &BatchConfig::wg_per_cu, wg), false);
      trial(with(best, &BatchConfig::xcd_map, 1 - base.xcd_map), false);
      if (a_rows > 0)
        for (int tn : {2048, 1024, 3072}) trial(with(best, &BatchConfig::tile_nnz, tn), false);
      if (!groups.empty())
        for (int sr : {64, 32, 128}) trial(with(best, &BatchConfig::spmm_rows, sr), false);
      if (done_its & 1) trial(best, true);  // one more iteration, with the chosen shape
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      cfg = best;
      tuned = true;
      std::lock_guard<std::mutex> g(tune_cache_mu());
      tune_cache()[signature] = cfg;
    }
    rep.tune_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rep.tune_iterations = done_its;
  }
  const BatchDev<VT, XT> d = dev_view<VT, XT>(cfg, p.beta);
  const int C = std::max(2, (p.check_every + 1) / 2 * 2);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (p.graph_capture && p.max_iteration > 0) {
    TBSLA_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < C; k++) enqueue_iteration<VT, XT>(d, k & 1, p, s);
    TBSLA_HIP_CHECK(hipStreamEndCapture(s, &graph));
    TBSLA_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  std::vector<RhsState> hs(static_cast<std::size_t>(n_rhs_total));
  // the state after the autotuner's trial iterations (the loop below may
  // have nothing left to run)
  TBSLA_HIP_CHECK(hipMemcpyAsync(hs.data(), rs.get(), hs.size() * sizeof(RhsState), hipMemcpyDeviceToHost, s));
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  bool all_done = true;
  for (const auto &st : hs) all_done = all_done && st.done;
  hipEvent_t e0, e1;
  TBSLA_HIP_CHECK(hipEventCreate(&e0));
  TBSLA_HIP_CHECK(hipEventCreate(&e1));
  TBSLA_HIP_CHECK(hipEventRecord(e0, s));
  for (int launched = done_its; !all_done && launched < p.max_iteration; launched += C) {
    if (exec)
      TBSLA_HIP_CHECK(hipGraphLaunch(exec, s));
    else
      for (int k = 0; k < C; k++) enqueue_iteration<VT, XT>(d, k & 1, p, s);
    TBSLA_HIP_CHECK(hipGetLastError());
    TBSLA_HIP_CHECK(hipMemcpyAsync(hs.data(), rs.get(), hs.size() * sizeof(RhsState), hipMemcpyDeviceToHost, s));
    TBSLA_HIP_CHECK(hipStreamSynchronize(s));
    bool all = true;
    for (const auto &st : hs) all = all && st.done;
    if (all) break;
  }
  TBSLA_HIP_CHECK(hipEventRecord(e1, s));
  // final sum normalisation of every RHS (Matrix.cpp:105-111), compensated
  if (n_items) {
    hipLaunchKernelGGL((batch_pass<VT, XT, 1>), dim3(n_items), dim3(kTileThreads), 0, s, d, items.as<PassItem>(),
                       parts.as<PassPart>(), out_off.as<long>(), out, 0);
    hipLaunchKernelGGL(batch_finalize_norm<1>, dim3(n_rhs_total), dim3(64), 0, s, d.rs, info.as<RhsInfo>(),
                       parts.as<PassPart>(), 0.0, 0.0, 0, 0);
    hipLaunchKernelGGL((batch_pass<VT, XT, 2>), dim3(n_items), dim3(kTileThreads), 0, s, d, items.as<PassItem>(),
                       parts.as<PassPart>(), out_off.as<long>(), out, 0);
  }
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  TBSLA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  // the trial iterations of the autotuner are iterations of the run: their
  // time counts in the run's (tune_seconds is the host time they took)
  rep.seconds = ms * 1e-3 + rep.tune_seconds;
  rep.config = cfg;
  rep.bytes_per_iteration = bytes_step + bytes_pass;
  for (int u = 0; u < n_rhs_total; u++) {
    const RhsState &st = hs[static_cast<std::size_t>(user_to_dev[static_cast<std::size_t>(u)])];
    rep.rhs_iterations.push_back(st.it);
    rep.rhs_error.push_back(st.err);
    rep.iterations = std::max(rep.iterations, st.it);
  }
  return rep;
}

PageRankBatch::PageRankBatch(Dtype precision) : impl_(new Impl()) {
  if (precision != Dtype::F64 && precision != Dtype::F32)
    throw std::invalid_argument("tbsla: batched PageRank runs in F64 or F32");
  impl_->prec = precision;
}

PageRankBatch::~PageRankBatch() = default;

void PageRankBatch::add_graph(const cpp::Matrix &m, const std::vector<std::vector<int>> &rhs_nodes) {
  if (impl_->prepared) throw std::invalid_argument("tbsla: batch already prepared");
  if (m.format() != cpp::FORMAT_CSR) throw std::invalid_argument("tbsla: batched PageRank needs CSR graphs");
  if (m.get_n_row() != m.get_n_col() || m.get_NR() != 1 || m.get_NC() != 1)
    throw std::invalid_argument("tbsla: page_rank needs a square matrix held in one block");
  if (rhs_nodes.empty()) throw std::invalid_argument("tbsla: a batch graph needs at least one teleport set");
  for (const auto &r : rhs_nodes)
    if (r.empty()) throw std::invalid_argument("tbsla: empty teleport node set");
  impl_->graphs.push_back(Impl::Graph{&dynamic_cast<const cpp::MatrixCSR &>(m), rhs_nodes});
}

void PageRankBatch::prepare() {
  if (!impl_->prepared) impl_->build();
}

BatchReport PageRankBatch::run(const BatchParams &p, double *out_dev, hipStream_t s) {
  prepare();
  if (impl_->n_rhs_total == 0) return BatchReport{};
  if (!impl_->stream) TBSLA_HIP_CHECK(hipStreamCreateWithFlags(&impl_->stream, hipStreamNonBlocking));
  // order the private stream after the caller's stream, and the caller's after the run
  hipEvent_t ev;
  TBSLA_HIP_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  TBSLA_HIP_CHECK(hipEventRecord(ev, s));
  TBSLA_HIP_CHECK(hipStreamWaitEvent(impl_->stream, ev, 0));
  BatchReport r = impl_->prec == Dtype::F64 ? impl_->run_typed<double, double>(p, out_dev, impl_->stream)
                                            : impl_->run_typed<float, float>(p, out_dev, impl_->stream);
  TBSLA_HIP_CHECK(hipEventRecord(ev, impl_->stream));
  TBSLA_HIP_CHECK(hipStreamWaitEvent(s, ev, 0));
  (void)hipEventDestroy(ev);
  return r;
}

int PageRankBatch::n_rhs() const {
  int n = 0;
  for (const auto &g : impl_->graphs) n += static_cast<int>(g.rhs.size());
  return n;
}

std::vector<long> PageRankBatch::rhs_offsets() const {
  if (impl_->prepared) return impl_->user_offsets;
  std::vector<long> o{0};
  for (const auto &g : impl_->graphs)
    for (std::size_t k = 0; k < g.rhs.size(); k++) o.push_back(o.back() + g.m->get_n_row());
  return o;
}

const BatchConfig &PageRankBatch::config() const { return impl_->cfg; }

}  // namespace hip
}  // namespace tbsla
