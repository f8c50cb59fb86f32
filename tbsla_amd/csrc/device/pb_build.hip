// GPU build of the propagation-blocked PageRank layout (PbMatrix,
// mirror.hpp) from the FP64 CSR mirror: the arrays of the host build
// (pb_build_host, mirror.cpp), from device sorts and scans.
//
//   far entries   compacted in CSR order (a scan of the far flags), then a
//                 stable radix sort by column gives the CSC order (rows
//                 ascending within a column); a stable sort of that by bin
//                 gives the bin-grouped slots, and stable sorts by
//                 (superblock, bin) give the two scatter orders, each with
//                 its run starts, run offsets (slot - entry) and chunk runs.
//   local entries compacted in CSR order (row order), then one wavefront
//                 per bin follows the bin's chain of groups of four (a group
//                 ends early where the next row is kPbLocalPad or more rows
//                 past its base), and one thread per entry places it.
// Range boundaries (superblocks, bins) come from binary searches over the
// sorted keys.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mirror.hpp"
#include "sort.hpp"

namespace tbsla {
namespace hip {
namespace {

constexpr int kT = 256;

unsigned blocks_for(long n) { return static_cast<unsigned>(std::max<long>(1, std::min<long>(1L << 16, (n + kT - 1) / kT))); }

int bits_for(long v) {  // bits that hold every value in [0, v]
  int b = 0;
  while (b < 62 && (1L << b) <= v) b++;
  return b;
}

[[noreturn]] void bad(const std::string &what) { throw std::invalid_argument("tbsla: " + what); }

// row of every used entry
__global__ __launch_bounds__(kT) void pbb_rows(const int *__restrict__ rowptr, int rp0, int n, int *__restrict__ row_of) {
  for (long r = blockIdx.x * static_cast<long>(kT) + threadIdx.x; r < n; r += static_cast<long>(gridDim.x) * kT)
    for (int e = rowptr[r] - rp0; e < rowptr[r + 1] - rp0; e++) row_of[e] = static_cast<int>(r);
}

// far flag of every entry (its column outside its row's bin); err[0] for a
// column outside [0, n)
__global__ __launch_bounds__(kT) void pbb_far(const int *__restrict__ colidx, const int *__restrict__ row_of, long nnz,
                                              int f_col, int n, int *__restrict__ far, int *__restrict__ err) {
  for (long e = blockIdx.x * static_cast<long>(kT) + threadIdx.x; e < nnz; e += static_cast<long>(gridDim.x) * kT) {
    const int c = colidx[e] - f_col;
    if (c < 0 || c >= n) {
      err[0] = 1;
      far[e] = 0;
    } else {
      far[e] = (c >> kPbShift) != (row_of[e] >> kPbShift) ? 1 : 0;
    }
  }
}

// the far entries in CSR order: key = column, val = entry (fpos: exclusive
// scan of the far flags); the local ones: row, column in the bin, entry
__global__ __launch_bounds__(kT) void pbb_compact(const int *__restrict__ far, const int *__restrict__ fpos,
                                                  const int *__restrict__ colidx, const int *__restrict__ row_of,
                                                  long nnz, int f_col, unsigned long long *__restrict__ fkey,
                                                  int *__restrict__ fval, int *__restrict__ lrow,
                                                  int *__restrict__ lcol, int *__restrict__ le) {
  for (long e = blockIdx.x * static_cast<long>(kT) + threadIdx.x; e < nnz; e += static_cast<long>(gridDim.x) * kT) {
    const int c = colidx[e] - f_col;
    if (far[e]) {
      const int p = fpos[e];
      fkey[p] = static_cast<unsigned long long>(c);
      fval[p] = static_cast<int>(e);
    } else {
      const int li = static_cast<int>(e) - fpos[e];
      lrow[li] = row_of[e];
      lcol[li] = c & (kPbRows - 1);
      le[li] = static_cast<int>(e);
    }
  }
}

// key = bin of the entry's row, val = its CSC position
__global__ __launch_bounds__(kT) void pbb_bin_keys(const int *__restrict__ csc_e, const int *__restrict__ row_of,
                                                   long nf, unsigned long long *__restrict__ key, int *__restrict__ val) {
  for (long k = blockIdx.x * static_cast<long>(kT) + threadIdx.x; k < nf; k += static_cast<long>(gridDim.x) * kT) {
    key[k] = static_cast<unsigned long long>(row_of[csc_e[k]] >> kPbShift);
    val[k] = static_cast<int>(k);
  }
}

// out[p] = the first k with key[k] >= p << shift (n if none), p = 0..np, for
// ascending keys: the boundaries of the key ranges (a binary search per p;
// counting with atomics contends on one address per range)
__global__ __launch_bounds__(kT) void pbb_bounds(const unsigned long long *__restrict__ key, long n, int np, int shift,
                                                 int *__restrict__ out) {
  for (long p = blockIdx.x * static_cast<long>(kT) + threadIdx.x; p <= np; p += static_cast<long>(gridDim.x) * kT) {
    const unsigned long long target = static_cast<unsigned long long>(p) << shift;
    long lo = 0, hi = n;
    while (lo < hi) {
      const long mid = (lo + hi) >> 1;
      if (key[mid] < target)
        lo = mid + 1;
      else
        hi = mid;
    }
    out[p] = static_cast<int>(lo);
  }
}

// slot s holds CSC position slot_list[s]: its inverse and the row in the bin
__global__ __launch_bounds__(kT) void pbb_slots(const int *__restrict__ slot_list, const int *__restrict__ csc_e,
                                                const int *__restrict__ row_of, long nf, int *__restrict__ slot_of,
                                                unsigned short *__restrict__ frow) {
  for (long s = blockIdx.x * static_cast<long>(kT) + threadIdx.x; s < nf; s += static_cast<long>(gridDim.x) * kT) {
    const int k = slot_list[s];
    slot_of[k] = static_cast<int>(s);
    frow[s] = static_cast<unsigned short>(row_of[csc_e[k]] & (kPbRows - 1));
  }
}

// key = superblock << binbits | bin, val = CSC position
__global__ __launch_bounds__(kT) void pbb_sc_keys(const unsigned long long *__restrict__ csc_col,
                                                  const int *__restrict__ csc_e, const int *__restrict__ row_of, long nf,
                                                  int colshift, int binbits, unsigned long long *__restrict__ key,
                                                  int *__restrict__ val) {
  for (long k = blockIdx.x * static_cast<long>(kT) + threadIdx.x; k < nf; k += static_cast<long>(gridDim.x) * kT) {
    const int sb = static_cast<int>(csc_col[k] >> colshift);
    const int b = row_of[csc_e[k]] >> kPbShift;
    key[k] = (static_cast<unsigned long long>(sb) << binbits) | static_cast<unsigned long long>(b);
    val[k] = static_cast<int>(k);
  }
}

// scatter entry o (CSC position sc[o], sorted key key[o]): its key with the
// run-start flag, its CSR entry, the run-start flag and slot - o
__global__ __launch_bounds__(kT) void pbb_sc_entries(const unsigned long long *__restrict__ key,
                                                     const int *__restrict__ sc,
                                                     const unsigned long long *__restrict__ csc_col,
                                                     const int *__restrict__ csc_e, const int *__restrict__ slot_of,
                                                     long nf, int cols, unsigned short *__restrict__ fkey,
                                                     int *__restrict__ far_src, int *__restrict__ start,
                                                     int *__restrict__ dd) {
  for (long o = blockIdx.x * static_cast<long>(kT) + threadIdx.x; o < nf; o += static_cast<long>(gridDim.x) * kT) {
    const int k = sc[o];
    const bool st = o == 0 || key[o] != key[o - 1];
    fkey[o] = static_cast<unsigned short>(static_cast<unsigned>(csc_col[k] & static_cast<unsigned long long>(cols - 1)) |
                                          (st ? kPbRunStart : 0u));
    far_src[o] = csc_e[k];
    start[o] = st ? 1 : 0;
    dd[o] = slot_of[k] - static_cast<int>(o);
  }
}

// rpos: exclusive scan of the run starts, so a start's run is rpos[o] and
// any entry's is rpos[o] + start[o] - 1
__global__ __launch_bounds__(kT) void pbb_rdelta(const int *__restrict__ start, const int *__restrict__ rpos,
                                                 const int *__restrict__ dd, long nf, int *__restrict__ rdelta) {
  for (long o = blockIdx.x * static_cast<long>(kT) + threadIdx.x; o < nf; o += static_cast<long>(gridDim.x) * kT)
    if (start[o]) rdelta[rpos[o]] = dd[o];
}
// every entry of a run lands at the run's offset (err[1] otherwise)
__global__ __launch_bounds__(kT) void pbb_rdelta_check(const int *__restrict__ start, const int *__restrict__ rpos,
                                                       const int *__restrict__ dd, const int *__restrict__ rdelta,
                                                       long nf, int *__restrict__ err) {
  for (long o = blockIdx.x * static_cast<long>(kT) + threadIdx.x; o < nf; o += static_cast<long>(gridDim.x) * kT)
    if (rdelta[rpos[o] + start[o] - 1] != dd[o]) err[1] = 1;
}
// first run of each superblock with entries (-1: none)
__global__ __launch_bounds__(kT) void pbb_first_run(const int *__restrict__ sp, const int *__restrict__ start,
                                                    const int *__restrict__ rpos, int ns, int *__restrict__ rfirst) {
  for (long sb = blockIdx.x * static_cast<long>(kT) + threadIdx.x; sb < ns; sb += static_cast<long>(gridDim.x) * kT) {
    const int o = sp[sb];
    rfirst[sb] = sp[sb + 1] > o ? rpos[o] + start[o] - 1 : -1;
  }
}
// the run (relative to its superblock's first) of each chunk's first entry
__global__ __launch_bounds__(kT) void pbb_cbase(const unsigned long long *__restrict__ key, const int *__restrict__ start,
                                                const int *__restrict__ rpos, const int *__restrict__ sp,
                                                const int *__restrict__ cp, const int *__restrict__ rp, long nf,
                                                int binbits, int *__restrict__ cbase) {
  for (long o = blockIdx.x * static_cast<long>(kT) + threadIdx.x; o < nf; o += static_cast<long>(gridDim.x) * kT) {
    const int sb = static_cast<int>(key[o] >> binbits);
    const int j0 = static_cast<int>(o) - sp[sb];
    if ((j0 & (kPbChunk - 1)) == 0) cbase[cp[sb] + j0 / kPbChunk] = rpos[o] + start[o] - 1 - rp[sb];
  }
}

// local entries [lbeg[b], lbeg[b + 1]) of bin b (in row order)
__global__ __launch_bounds__(kT) void pbb_local_begin(const int *__restrict__ rowptr, int rp0,
                                                      const int *__restrict__ fpos, long nnz, int n, int nb,
                                                      int nloc, int *__restrict__ lbeg) {
  for (long b = blockIdx.x * static_cast<long>(kT) + threadIdx.x; b <= nb; b += static_cast<long>(gridDim.x) * kT) {
    const long r = b << kPbShift;
    int v = nloc;
    if (r < n) {
      const long e = rowptr[r] - rp0;
      if (e < nnz) v = static_cast<int>(e) - fpos[e];
    }
    lbeg[b] = v;
  }
}
// Groups of each bin, one wavefront per bin: an entry opens a group when the
// open one is full or its row is kPbLocalPad or more rows past the group's
// base.  So a group that opens at entry s ends before the first entry j > s
// with j = s + kPbLocalGroup or row(j) - row(s) >= kPbLocalPad: each lane
// finds that end for a group opening at its own entry, and the wavefront
// follows the chain of groups through each 64-entry window (a lane shuffle
// per group).  rel[li] = the slot of local entry li relative to its bin's
// first slot (kPbLocalGroup * group + position in the group).
__global__ __launch_bounds__(kT) void pbb_local_walk(const int *__restrict__ lbeg, const int *__restrict__ lrow, int nb,
                                                     int *__restrict__ rel, int *__restrict__ ngroups) {
  const long wave = (blockIdx.x * static_cast<long>(kT) + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= nb) return;  // (uniform over the wavefront)
  const int b = static_cast<int>(wave);
  const int l0 = lbeg[b], l1 = lbeg[b + 1];
  int g = 0;              // groups opened so far
  int s = l0, sp = l0;    // the next group's first entry, the last opened group's
  for (int w = l0; w < l1; w += 64) {
    const int i = w + lane;
    int nx = i + kPbLocalGroup;  // end of a group opening at entry i
    if (i < l1) {
      const int r = lrow[i];
#pragma unroll
      for (int k = kPbLocalGroup - 1; k >= 1; k--)
        if (i + k < l1 && lrow[i + k] - r >= kPbLocalPad) nx = i + k;
    }
    nx = min(nx, l1);
    // entries before s in this window belong to the group opened at sp
    int mine = i < s ? kPbLocalGroup * (g - 1) + (i - sp) : -1;
    while (s < w + 64 && s < l1) {
      const int e = __shfl(nx, s - w, 64);
      if (i >= s && i < e) mine = kPbLocalGroup * g + (i - s);
      sp = s;
      s = e;
      g++;
    }
    if (i < l1) rel[i] = mine;
  }
  if (lane == 0) ngroups[b] = g;
}
// Local entry li into its slot: key, CSR entry, and the group's base row at
// the group's first entry (padding slots keep their fill)
__global__ __launch_bounds__(kT) void pbb_local_place(const int *__restrict__ rel, const int *__restrict__ lrow,
                                                      const int *__restrict__ lcol, const int *__restrict__ le,
                                                      const int *__restrict__ lp, long nloc,
                                                      unsigned short *__restrict__ lkey, unsigned short *__restrict__ lbase,
                                                      int *__restrict__ loc_src) {
  for (long li = blockIdx.x * static_cast<long>(kT) + threadIdx.x; li < nloc; li += static_cast<long>(gridDim.x) * kT) {
    const int r = lrow[li], q = rel[li];
    const int pos = q % kPbLocalGroup;
    const int base = lrow[li - pos];  // the group's first entry
    const int k = lp[r >> kPbShift] + q;
    loc_src[k] = le[li];
    lkey[k] = static_cast<unsigned short>(lcol[li] | ((r - base) << kPbShift));
    if (pos == 0) lbase[k / kPbLocalGroup] = static_cast<unsigned short>(base & (kPbRows - 1));
  }
}

__global__ __launch_bounds__(kT) void pbb_fill_u16(unsigned short *__restrict__ a, long n, unsigned short v) {
  for (long i = blockIdx.x * static_cast<long>(kT) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * kT) a[i] = v;
}
__global__ __launch_bounds__(kT) void pbb_fill_i32(int *__restrict__ a, long n, int v) {
  for (long i = blockIdx.x * static_cast<long>(kT) + threadIdx.x; i < n; i += static_cast<long>(gridDim.x) * kT) a[i] = v;
}

// value of entry src[k] (0 for padding, src < 0) in the stored type
__device__ inline unsigned short bf16_of(double d) {  // round to nearest even via FP32, as the host build
  const float f = static_cast<float>(d);
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}
template <typename T>
__device__ inline T conv(double d);
template <>
__device__ inline double conv<double>(double d) {
  return d;
}
template <>
__device__ inline float conv<float>(double d) {
  return static_cast<float>(d);
}
template <>
__device__ inline unsigned short conv<unsigned short>(double d) {
  return bf16_of(d);
}
template <typename T>
__global__ __launch_bounds__(kT) void pbb_values(const double *__restrict__ vals, const int *__restrict__ src, long m,
                                                 T *__restrict__ out) {
  for (long k = blockIdx.x * static_cast<long>(kT) + threadIdx.x; k < m; k += static_cast<long>(gridDim.x) * kT) {
    const int e = src[k];
    out[k] = e >= 0 ? conv<T>(vals[e]) : T(0);
  }
}

int read_int(const int *p) {
  int v = 0;
  TBSLA_HIP_CHECK(hipMemcpy(&v, p, sizeof(int), hipMemcpyDeviceToHost));
  return v;
}

template <typename T>
std::vector<T> download(const DevBuf &b, std::size_t n) {
  std::vector<T> v(n);
  if (n) TBSLA_HIP_CHECK(hipMemcpy(v.data(), b.get(), n * sizeof(T), hipMemcpyDeviceToHost));
  return v;
}

// One scatter order (superblocks of `cols` columns) of the CSC-ordered far
// entries: sp, fkey, rp, rdelta, cbase, cp and far_src.
void scatter_order(int cols, int n, int nb, long nf, const DevBuf &csc_col, const DevBuf &csc_e, const DevBuf &row_of,
                   const DevBuf &slot_of, int &nsb, DevBuf &sp_d, DevBuf &fkey_d, DevBuf &rp_d, DevBuf &rdelta_d,
                   DevBuf &cbase_d, DevBuf &cp_d, DevBuf &far_src_d, DevBuf &err) {
  const hipStream_t s = nullptr;
  int colshift = 0;
  while ((1 << colshift) < cols) colshift++;
  const int ns = (n + cols - 1) / cols;
  nsb = ns;
  const int binbits = bits_for(nb - 1), sbbits = bits_for(ns - 1);
  DevBuf key, val, k2, v2, spb;
  spb.alloc(sizeof(int) * (static_cast<std::size_t>(ns) + 1));
  fkey_d.alloc(sizeof(unsigned short) * static_cast<std::size_t>(std::max(1L, nf)));
  far_src_d.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nf)));
  DevBuf start, rpos, dd;
  if (nf > 0) {
    key.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(nf));
    val.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    k2.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(nf));
    v2.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    hipLaunchKernelGGL(pbb_sc_keys, dim3(blocks_for(nf)), dim3(kT), 0, s, csc_col.as<unsigned long long>(),
                       csc_e.as<int>(), row_of.as<int>(), nf, colshift, binbits, key.as<unsigned long long>(),
                       val.as<int>());
    radix_sort_pairs(key.as<unsigned long long>(), val.as<int>(), k2.as<unsigned long long>(), v2.as<int>(), nf,
                     binbits + sbbits, s);
    hipLaunchKernelGGL(pbb_bounds, dim3(blocks_for(ns + 1L)), dim3(kT), 0, s, key.as<unsigned long long>(), nf, ns,
                       binbits, spb.as<int>());
    start.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    rpos.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    dd.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    hipLaunchKernelGGL(pbb_sc_entries, dim3(blocks_for(nf)), dim3(kT), 0, s, key.as<unsigned long long>(),
                       val.as<int>(), csc_col.as<unsigned long long>(), csc_e.as<int>(), slot_of.as<int>(), nf, cols,
                       fkey_d.as<unsigned short>(), far_src_d.as<int>(), start.as<int>(), dd.as<int>());
    TBSLA_HIP_CHECK(hipMemcpyAsync(rpos.get(), start.get(), sizeof(int) * static_cast<std::size_t>(nf),
                                   hipMemcpyDeviceToDevice, s));
    exclusive_scan(rpos.as<int>(), nf, s);
  }
  const int runs = nf > 0 ? read_int(rpos.as<int>() + nf - 1) + read_int(start.as<int>() + nf - 1) : 0;
  // superblock boundaries, first runs and chunk offsets (small, on the host)
  std::vector<int> sp(static_cast<std::size_t>(ns) + 1, 0), cp(static_cast<std::size_t>(ns) + 1, 0);
  if (nf > 0) sp = download<int>(spb, static_cast<std::size_t>(ns) + 1);
  for (int b = 0; b < ns; b++)
    cp[static_cast<std::size_t>(b) + 1] =
        cp[static_cast<std::size_t>(b)] +
        (sp[static_cast<std::size_t>(b) + 1] - sp[static_cast<std::size_t>(b)] + kPbChunk - 1) / kPbChunk;
  sp_d.upload(sp.data(), sp.size() * sizeof(int));
  cp_d.upload(cp.data(), cp.size() * sizeof(int));
  std::vector<int> rp(static_cast<std::size_t>(ns) + 1, runs);
  if (nf > 0) {
    DevBuf rfirst;
    rfirst.alloc(sizeof(int) * static_cast<std::size_t>(ns));
    hipLaunchKernelGGL(pbb_first_run, dim3(blocks_for(ns)), dim3(kT), 0, s, sp_d.as<int>(), start.as<int>(),
                       rpos.as<int>(), ns, rfirst.as<int>());
    std::vector<int> rf = download<int>(rfirst, static_cast<std::size_t>(ns));
    for (int b = ns - 1; b >= 0; b--)  // a superblock with no entries starts where the next one does
      rp[static_cast<std::size_t>(b)] = rf[static_cast<std::size_t>(b)] >= 0 ? rf[static_cast<std::size_t>(b)]
                                                                             : rp[static_cast<std::size_t>(b) + 1];
  }
  rp_d.upload(rp.data(), rp.size() * sizeof(int));
  rdelta_d.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1, runs)));
  const int nck = cp[static_cast<std::size_t>(ns)];
  cbase_d.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1, nck)));
  if (nf > 0) {
    hipLaunchKernelGGL(pbb_rdelta, dim3(blocks_for(nf)), dim3(kT), 0, s, start.as<int>(), rpos.as<int>(),
                       dd.as<int>(), nf, rdelta_d.as<int>());
    hipLaunchKernelGGL(pbb_rdelta_check, dim3(blocks_for(nf)), dim3(kT), 0, s, start.as<int>(), rpos.as<int>(),
                       dd.as<int>(), rdelta_d.as<int>(), nf, err.as<int>());
    hipLaunchKernelGGL(pbb_cbase, dim3(blocks_for(nf)), dim3(kT), 0, s, key.as<unsigned long long>(),
                       start.as<int>(), rpos.as<int>(), sp_d.as<int>(), cp_d.as<int>(), rp_d.as<int>(), nf, binbits,
                       cbase_d.as<int>());
  }
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
}

}  // namespace

std::unique_ptr<PbMatrix> pb_build_gpu(const DeviceMatrix &d, int n) {
  const hipStream_t s = nullptr;
  if (d.format != cpp::FORMAT_CSR || d.vdt != Dtype::F64) bad("the propagation-blocked layout is built from the FP64 CSR mirror");
  if (!d.noop && d.rows != n) bad("CSR rowptr does not cover every row");
  const long nnz = d.noop ? 0 : d.entries;  // no values: no entries (MatrixCSR.cpp:130-131)
  if (nnz >= (1L << 31)) bad("propagation-blocked layout limited to 2^31 entries");
  auto pb = std::unique_ptr<PbMatrix>(new PbMatrix());
  pb->n = n;
  pb->nbins = (n + kPbRows - 1) / kPbRows;
  pb->nsb = (n + kPbCols - 1) / kPbCols;
  const int nb = pb->nbins;
  DevBuf err, row_of, far, fpos;
  err.alloc(4 * sizeof(int));
  const std::size_t e1 = static_cast<std::size_t>(std::max(1L, nnz));
  row_of.alloc(sizeof(int) * e1);
  far.alloc(sizeof(int) * e1);
  fpos.alloc(sizeof(int) * e1);
  long nf = 0;
  if (nnz > 0) {
    hipLaunchKernelGGL(pbb_rows, dim3(blocks_for(n)), dim3(kT), 0, s, d.rowptr.as<int>(), d.rp0, n, row_of.as<int>());
    hipLaunchKernelGGL(pbb_far, dim3(blocks_for(nnz)), dim3(kT), 0, s, d.colidx.as<int>(), row_of.as<int>(), nnz,
                       d.f_col, n, far.as<int>(), err.as<int>());
    TBSLA_HIP_CHECK(hipMemcpyAsync(fpos.get(), far.get(), sizeof(int) * static_cast<std::size_t>(nnz),
                                   hipMemcpyDeviceToDevice, s));
    exclusive_scan(fpos.as<int>(), nnz, s);
    nf = static_cast<long>(read_int(fpos.as<int>() + nnz - 1)) + read_int(far.as<int>() + nnz - 1);
    if (read_int(err.as<int>())) bad("CSR column index outside the matrix");
  }
  const long nloc = nnz - nf;
  pb->n_far = nf;
  // far entries in CSR order, then CSC order; local entries in CSR order
  DevBuf csc_col, csc_e, k2, v2, lrow, lcol, le;
  csc_col.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(std::max(1L, nf)));
  csc_e.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nf)));
  lrow.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nloc)));
  lcol.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nloc)));
  le.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nloc)));
  if (nnz > 0)
    hipLaunchKernelGGL(pbb_compact, dim3(blocks_for(nnz)), dim3(kT), 0, s, far.as<int>(), fpos.as<int>(),
                       d.colidx.as<int>(), row_of.as<int>(), nnz, d.f_col, csc_col.as<unsigned long long>(),
                       csc_e.as<int>(), lrow.as<int>(), lcol.as<int>(), le.as<int>());
  if (nf > 0) {
    k2.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(nf));
    v2.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    radix_sort_pairs(csc_col.as<unsigned long long>(), csc_e.as<int>(), k2.as<unsigned long long>(), v2.as<int>(), nf,
                     bits_for(n - 1), s);
  }
  // bin-grouped slots
  DevBuf bkey, slot_list, bpb, slot_of;
  bpb.alloc(sizeof(int) * (static_cast<std::size_t>(nb) + 1));
  pb->frow.alloc(sizeof(unsigned short) * static_cast<std::size_t>(std::max(1L, nf)));
  slot_of.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nf)));
  if (nf > 0) {
    bkey.alloc(sizeof(unsigned long long) * static_cast<std::size_t>(nf));
    slot_list.alloc(sizeof(int) * static_cast<std::size_t>(nf));
    hipLaunchKernelGGL(pbb_bin_keys, dim3(blocks_for(nf)), dim3(kT), 0, s, csc_e.as<int>(), row_of.as<int>(), nf,
                       bkey.as<unsigned long long>(), slot_list.as<int>());
    radix_sort_pairs(bkey.as<unsigned long long>(), slot_list.as<int>(), k2.as<unsigned long long>(), v2.as<int>(), nf,
                     bits_for(nb - 1), s);
    hipLaunchKernelGGL(pbb_bounds, dim3(blocks_for(nb + 1L)), dim3(kT), 0, s, bkey.as<unsigned long long>(), nf, nb, 0,
                       bpb.as<int>());
    hipLaunchKernelGGL(pbb_slots, dim3(blocks_for(nf)), dim3(kT), 0, s, slot_list.as<int>(), csc_e.as<int>(),
                       row_of.as<int>(), nf, slot_of.as<int>(), pb->frow.as<unsigned short>());
  }
  if (nf > 0) {
    pb->bp = std::move(bpb);
  } else {
    std::vector<int> bp(static_cast<std::size_t>(nb) + 1, 0);
    pb->bp.upload(bp.data(), bp.size() * sizeof(int));
  }
  k2.release();
  v2.release();
  bkey.release();
  slot_list.release();
  // the two scatter orders
  int ns_check = 0;
  scatter_order(kPbCols, n, nb, nf, csc_col, csc_e, row_of, slot_of, ns_check, pb->sp, pb->fkey, pb->rp, pb->rdelta,
                pb->cbase, pb->cp, pb->far_src, err);
  if (ns_check != pb->nsb) bad("propagation-blocked layout: superblock count");
  scatter_order(kPbColsWide, n, nb, nf, csc_col, csc_e, row_of, slot_of, pb->nsb_w, pb->sp_w, pb->fkey_w, pb->rp_w,
                pb->rdelta_w, pb->cbase_w, pb->cp_w, pb->far_src_w, err);
  if (read_int(err.as<int>() + 1)) bad("propagation-blocked layout: run slots are not consecutive");
  // local groups
  DevBuf lbeg, ngroups, rel;
  lbeg.alloc(sizeof(int) * (static_cast<std::size_t>(nb) + 1));
  ngroups.alloc(sizeof(int) * static_cast<std::size_t>(nb));
  rel.alloc(sizeof(int) * static_cast<std::size_t>(std::max(1L, nloc)));
  if (nloc > 0) {
    hipLaunchKernelGGL(pbb_local_begin, dim3(blocks_for(nb + 1)), dim3(kT), 0, s, d.rowptr.as<int>(), d.rp0,
                       fpos.as<int>(), nnz, n, nb, static_cast<int>(nloc), lbeg.as<int>());
    hipLaunchKernelGGL(pbb_local_walk, dim3((nb + kT / 64 - 1) / (kT / 64)), dim3(kT), 0, s, lbeg.as<int>(),
                       lrow.as<int>(), nb, rel.as<int>(), ngroups.as<int>());
  }
  std::vector<int> g = download<int>(ngroups, static_cast<std::size_t>(nb));
  std::vector<int> lp(static_cast<std::size_t>(nb) + 1, 0);
  for (int b = 0; b < nb; b++) {
    const long v = static_cast<long>(lp[static_cast<std::size_t>(b)]) + static_cast<long>(g[static_cast<std::size_t>(b)]) * kPbLocalGroup;
    if (v >= (1L << 31)) bad("propagation-blocked layout limited to 2^31 entries");
    lp[static_cast<std::size_t>(b) + 1] = static_cast<int>(v);
  }
  const long nl = lp[static_cast<std::size_t>(nb)];
  pb->n_loc = nl;
  pb->lp.upload(lp.data(), lp.size() * sizeof(int));
  const std::size_t nl1 = static_cast<std::size_t>(std::max(1L, nl));
  pb->lkey.alloc(sizeof(unsigned short) * nl1);
  pb->lbase.alloc(sizeof(unsigned short) * std::max<std::size_t>(1, static_cast<std::size_t>(nl) / kPbLocalGroup));
  pb->loc_src.alloc(sizeof(int) * nl1);
  if (nl > 0) {
    hipLaunchKernelGGL(pbb_fill_u16, dim3(blocks_for(nl)), dim3(kT), 0, s, pb->lkey.as<unsigned short>(), nl,
                       static_cast<unsigned short>(kPbLocalPad << kPbShift));
    hipLaunchKernelGGL(pbb_fill_i32, dim3(blocks_for(nl)), dim3(kT), 0, s, pb->loc_src.as<int>(), nl, -1);
    hipLaunchKernelGGL(pbb_local_place, dim3(blocks_for(nloc)), dim3(kT), 0, s, rel.as<int>(), lrow.as<int>(),
                       lcol.as<int>(), le.as<int>(), pb->lp.as<int>(), nloc, pb->lkey.as<unsigned short>(),
                       pb->lbase.as<unsigned short>(), pb->loc_src.as<int>());
  }
  pb->prod.alloc(sizeof(double) * (static_cast<std::size_t>(nf) + kPbChunk));  // + spare slots
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  return pb;
}

void pb_build_values(PbMatrix &pb, const DeviceMatrix &d, Dtype vdt) {
  const hipStream_t s = nullptr;
  const int slot = static_cast<int>(vdt);
  const long nf = pb.n_far, nl = pb.n_loc;
  // far values in the scatter order that reads them: F64 (anchor and FP64
  // products) and F32 (delta products) narrow, BF16 (delta products) wide
  const DevBuf &fsrc = vdt == Dtype::BF16 ? pb.far_src_w : pb.far_src;
  const std::size_t esz = vdt == Dtype::F64 ? 8 : (vdt == Dtype::F32 ? 4 : 2);
  pb.fval[slot].alloc(esz * static_cast<std::size_t>(std::max(1L, nf)));
  pb.lval[slot].alloc(esz * static_cast<std::size_t>(std::max(1L, nl)));
  auto run = [&](auto tag) {
    using T = decltype(tag);
    if (nf > 0)
      hipLaunchKernelGGL(pbb_values<T>, dim3(blocks_for(nf)), dim3(kT), 0, s, d.vals.as<double>(), fsrc.as<int>(), nf,
                         pb.fval[slot].as<T>());
    if (nl > 0)
      hipLaunchKernelGGL(pbb_values<T>, dim3(blocks_for(nl)), dim3(kT), 0, s, d.vals.as<double>(),
                         pb.loc_src.as<int>(), nl, pb.lval[slot].as<T>());
  };
  if (vdt == Dtype::F64)
    run(double());
  else if (vdt == Dtype::F32)
    run(float());
  else if (vdt == Dtype::BF16)
    run(static_cast<unsigned short>(0));
  TBSLA_HIP_CHECK(hipGetLastError());
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
}

}  // namespace hip
}  // namespace tbsla
