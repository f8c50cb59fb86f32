// Device mirror of one host matrix: the HBM layout the kernels read.
//
//   CSR   rowptr/colidx as stored, values in the requested type, plus the
//         entry-tile table: tile t = entries [t*kCsrTileNnz, (t+1)*kCsrTileNnz)
//         owns rows [tile_row[t], tile_row[t+1]) (the rows whose first entry
//         lies in the tile), and the list of rows longer than kCsrLongRow.
//   ELL   sliced column-major: slice b = rows [b*kEllSlice, (b+1)*kEllSlice)
//         holds its K slots contiguously, slot j of row i at
//         ((i / kEllSlice) * K + j) * kEllSlice + i % kEllSlice, so one
//         workgroup streams one contiguous block.  Columns are stored local
//         and clamped, max(col - f_col, 0) (MatrixELL.cpp:110); the slice's
//         rows past the last hold column 0, value 0.
//   COO / SCOO
//         entries in a row-sorted view (a stable counting sort by row when the
//         input is not already sorted), compressed to row segments: segment s
//         = entries [rowptr[s], rowptr[s+1]) of row segrow[s].  The CSR tile
//         kernels run on it, each segment's sum starting from y[row], so every
//         row is accumulated in entry order.
//   DENSE row-major, as stored.
//
// Every index the kernels will read is bounded here on the host (x_need,
// y range); the launchers refuse vectors shorter than that.
#pragma once

#include <tbsla/cpp/Matrix.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>

#include <memory>
#include <string>
#include <mutex>
#include <vector>

#include "common.hpp"

namespace tbsla {
namespace hip {

constexpr int kTileThreads = 256;
constexpr int kTilePer = 12;  // 3072-entry tiles: 2048 and 4096 measured slower on the cqmat block
constexpr int kCsrTileNnz = kTileThreads * kTilePer;  // entries staged in LDS per tile
constexpr int kCsrLongRow = kCsrTileNnz / 2;          // longer rows: one workgroup each
constexpr int kEllSlice = 128;  // ELL rows per slice (one row per thread, one workgroup per slice)

struct DeviceMatrix {
  cpp::Format format = cpp::FORMAT_CSR;
  Dtype vdt = Dtype::F64;
  bool noop = false;       // the reference returns before touching r
  int rows = 0;            // CSR/ELL/DENSE: rows written (assign) / accumulated (DENSE)
  long entries = 0;        // stored entries / slots read by one product
  int f_row = 0, f_col = 0;
  long x_need = 0;         // x must hold at least this many elements
  long y_lo = 0, y_hi = 0; // COO/SCOO: rows touched are [y_lo, y_hi) before vect_incr
  // CSR
  DevBuf rowptr, colidx, vals, tiles, long_rows;
  int ntiles = 0, nlong = 0, rp0 = 0;
  // ELL (sliced column-major)
  int K = 0;
  DevBuf ecols, evals;
  // COO / SCOO: segments of the row-sorted view (rowptr, colidx, vals, tiles
  // and long_rows above hold the segment CSR)
  DevBuf segrow;
  int nseg = 0;
  // DENSE
  int dcols = 0;
  DevBuf dvals;
  // count-stochastic CSR: 1/count of each local column
  DevBuf colscale;
  // transposed CSR structure: CSR entry index of each transposed entry
  DevBuf perm;
  std::size_t algorithmic_bytes(std::size_t x_elem, std::size_t y_elem) const;
};

// Propagation-blocked layout of a square CSR matrix for PageRank.  Rows are
// grouped in bins of kPbRows; an entry whose column lies in its row's bin is
// "local", every other entry is "far".  Columns are grouped in superblocks of
// kPbCols.
//   far    stored twice over: in scatter order -- by (column superblock,
//          row bin, column) -- with the column offset in the superblock, the
//          value and the slot of its product (sp: superblock boundaries), and
//          in bin-grouped order -- by (row bin, column) -- with the row offset
//          in the bin (bp: bin boundaries).  Within a bin the products of one
//          superblock are contiguous, so the scatter pass writes runs.
//   local  bin-grouped in row order (lp: bin boundaries), summed by the
//          gather pass with x read from the bin's own slice, in groups of
//          kPbLocalGroup entries: each group has a base row (lbase, offset in
//          the bin) and each entry a 16-bit key lkey = column offset in the
//          bin | (row - base) << kPbShift, with row - base < kPbLocalPad; a
//          key whose top field is kPbLocalPad is padding (column 0, value 0).
//          A group ends early, padded, where the next entry's row is
//          kPbLocalPad or more rows past its base, and each bin's run is
//          padded to a whole number of groups
constexpr int kPbShift = 13;
constexpr int kPbRows = 1 << kPbShift;
constexpr int kPbColShift = 13;
constexpr int kPbLocalGroup = 4;
constexpr int kPbLocalPad = 7;  // the row field of a padding key (3 bits above the column)
constexpr int kPbCols = 1 << kPbColShift;
// the wide scatter order: superblocks of kPbColsWide columns, for the FP32
// delta products of mixed runs (a 4-byte x slice of that width fills the
// same 64 KB of LDS as an 8-byte slice of kPbCols)
constexpr int kPbColShiftWide = 14;
constexpr int kPbColsWide = 1 << kPbColShiftWide;
// scatter keys: the column within the superblock, and this flag on the first
// entry of each run; the run of an entry is found per wavefront-sized chunk
constexpr unsigned kPbRunStart = 0x8000u;
constexpr int kPbChunk = 64;
static_assert(kPbShift + 3 <= 16, "local keys hold the column and a 3-bit row offset in 16 bits");
static_assert(kPbCols <= 0x2000, "scatter keys hold the column in 13 bits");
static_assert(kPbColsWide <= kPbRunStart, "wide scatter keys hold the column in 15 bits");

struct PbMatrix {
  int n = 0, nbins = 0, nsb = 0;
  long n_far = 0, n_loc = 0;  // n_loc counts the padding entries too
  // far entries in scatter order: fkey (16 bits) = column within the
  // superblock, | kPbRunStart on the first entry of a run. A run is the far
  // entries of one superblock that go to one bin; entry o of run r of
  // superblock s goes to product slot o + rdelta[rp[s] + r]. Chunk j of
  // superblock s (entries sp[s] + 64j ...) begins in run cbase[cp[s] + j],
  // so the run of an entry is cbase plus the run starts after the chunk's
  // first entry up to the entry itself.
  DevBuf sp, fkey, rp, rdelta, cbase, cp, frow, bp, lp, lkey, lbase;
  // values by Dtype F64, F32, BF16 (built on demand); the far values in the
  // scatter order that reads them: F64 and F32 narrow, BF16 wide
  DevBuf fval[3], lval[3];
  // The far entries in the wide scatter order (superblocks of kPbColsWide
  // columns, the same product slots), read by the delta products of BF16 and
  // value-free mixed runs.
  int nsb_w = 0;
  DevBuf sp_w, fkey_w, rp_w, rdelta_w, cbase_w, cp_w;
  DevBuf prod;              // 8-byte product slots, n_far + kPbChunk (spare slots that
                            // the scatter's lanes past a superblock's end write)
  // CSR entry of each far entry (narrow and wide scatter order) and of each
  // local slot (-1: padding), int, on the device (the value builds read them)
  DevBuf far_src, far_src_w, loc_src;
};

struct DeviceCache {
  std::mutex mu;
  std::unique_ptr<DeviceMatrix> mirrors[4];
  std::unique_ptr<DeviceMatrix> transposed;
  std::unique_ptr<PbMatrix> pb;
  DevBuf x_stage, y_stage;
  // count-stochastic CSR: the value of each column (fl(1/count)) in FP64 and
  // FP32, 0 for a column with no entries (built by the PageRank engine)
  DevBuf colscale64, colscale32;
};

// Build (or fetch) the mirror of `m` with values of type vdt.
const DeviceMatrix &mirror(const cpp::Matrix &m, Dtype vdt);
// The mirror if it was built already, else nullptr (builds nothing).
const DeviceMatrix *mirror_if_built(const cpp::Matrix &m, Dtype vdt);

// Value-free CSR structure of A^T for a CSR matrix A (rows = A's local
// columns, columns = A's local rows): row j lists the entries of column j of
// A in increasing row order, and perm[k] is the CSR entry index (relative to
// rowptr[0]) of transposed entry k, so values_T[k] = values[perm[k]].
const DeviceMatrix &transpose_mirror(const cpp::Matrix &m);

// The propagation-blocked layout of a square single-block CSR matrix, with
// values of type vdt (PATTERN: indices only), built on the GPU once per
// matrix (values once per type).
const PbMatrix &pb_mirror(const cpp::Matrix &m, Dtype vdt);

// Builds of the layout's index arrays (no values): on the GPU from the FP64
// mirror d64 of an n x n CSR matrix (device sorts), and on the host from
// the matrix's arrays; both give the same arrays.
std::unique_ptr<PbMatrix> pb_build_gpu(const DeviceMatrix &d64, int n);
std::unique_ptr<PbMatrix> pb_build_host(const cpp::MatrixCSR &c);
// The far and local values of type vdt, gathered on the GPU from d64.
void pb_build_values(PbMatrix &pb, const DeviceMatrix &d64, Dtype vdt);
// "" when the index arrays of a and b are equal, else the first that differs.
std::string pb_layout_compare(const PbMatrix &a, const PbMatrix &b);

// Host tile tables (exposed for tests and the autotuner).
// tile_row of the CSR entry tiles (ntiles + 1 entries, ntiles = nnz / T + 1).
std::vector<int> csr_entry_tiles(const int *rowptr, int rows, int tile_nnz);

}  // namespace hip
}  // namespace tbsla
