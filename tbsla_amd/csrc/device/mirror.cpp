#include "mirror.hpp"

#include <tbsla/cpp/MatrixCOO.hpp>
#include <tbsla/cpp/MatrixCSR.hpp>
#include <tbsla/cpp/MatrixDENSE.hpp>
#include <tbsla/cpp/MatrixELL.hpp>
#include <tbsla/cpp/MatrixSCOO.hpp>

#include <algorithm>
#include <cstring>
#include <limits>

#include "../host/gen.hpp"

namespace tbsla {
namespace hip {
namespace {

// Round-to-nearest-even double -> float -> bfloat16 (upper 16 bits).
unsigned short to_bf16(double d) {
  float f = static_cast<float>(d);
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<unsigned short>(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}

// Upload `n` values from `src` converted to vdt (PATTERN uploads nothing).
void upload_values(DevBuf &dst, const double *src, std::size_t n, Dtype vdt) {
  switch (vdt) {
    case Dtype::F64:
      dst.upload(src, n * sizeof(double));
      break;
    case Dtype::F32: {
      std::vector<float> t(n);
      gen::parallel_for(static_cast<long>(n), [&](long b, long e) {
        for (long i = b; i < e; i++) t[static_cast<std::size_t>(i)] = static_cast<float>(src[i]);
      });
      dst.upload(t.data(), n * sizeof(float));
      break;
    }
    case Dtype::BF16: {
      std::vector<unsigned short> t(n);
      gen::parallel_for(static_cast<long>(n), [&](long b, long e) {
        for (long i = b; i < e; i++) t[static_cast<std::size_t>(i)] = to_bf16(src[i]);
      });
      dst.upload(t.data(), n * sizeof(unsigned short));
      break;
    }
    case Dtype::PATTERN:
      break;
  }
}

struct MinMax {
  long lo = std::numeric_limits<long>::max(), hi = std::numeric_limits<long>::min();
};
MinMax minmax_offset(const int *a, std::size_t n, long off) {
  std::vector<MinMax> res;
  std::mutex mu;
  gen::parallel_for(static_cast<long>(n), [&](long b, long e) {
    MinMax m;
    for (long i = b; i < e; i++) {
      const long v = static_cast<long>(a[i]) + off;
      m.lo = std::min(m.lo, v);
      m.hi = std::max(m.hi, v);
    }
    std::lock_guard<std::mutex> g(mu);
    res.push_back(m);
  }, 1 << 16);
  MinMax out;
  for (auto &m : res) {
    out.lo = std::min(out.lo, m.lo);
    out.hi = std::max(out.hi, m.hi);
  }
  return out;
}

[[noreturn]] void bad(const char *what) { throw std::invalid_argument(std::string("tbsla: ") + what); }

void build_csr(DeviceMatrix &d, const cpp::MatrixCSR &m, Dtype vdt) {
  const auto &vals = m.get_values();
  const auto &rp = m.get_rowptr();
  const auto &ci = m.get_colidx();
  if (vals.empty() || rp.size() < 2) {  // MatrixCSR.cpp:130-131
    d.noop = true;
    return;
  }
  d.rows = static_cast<int>(rp.size() - 1);
  d.rp0 = rp[0];
  for (int i = 0; i < d.rows; i++)
    if (rp[static_cast<std::size_t>(i) + 1] < rp[static_cast<std::size_t>(i)]) bad("CSR rowptr is not monotone");
  const long used = static_cast<long>(rp.back()) - rp[0];
  if (used > static_cast<long>(vals.size()) || used > static_cast<long>(ci.size()))
    bad("CSR rowptr points past values/colidx");
  d.entries = used;
  d.f_col = m.get_f_col();
  if (used > 0) {
    const MinMax mm = minmax_offset(ci.data(), static_cast<std::size_t>(used), -static_cast<long>(d.f_col));
    if (mm.lo < 0) bad("CSR column index below the block's first column");
    d.x_need = mm.hi + 1;
  }
  d.rowptr.upload(rp.data(), rp.size() * sizeof(int));
  // entry j of row i is values[rowptr[i] - rowptr[0] + k] (MatrixCSR.cpp:135)
  d.colidx.upload(ci.data(), static_cast<std::size_t>(used) * sizeof(int));
  upload_values(d.vals, vals.data(), static_cast<std::size_t>(used), vdt);
  std::vector<int> t = csr_entry_tiles(rp.data(), d.rows, kCsrTileNnz);
  d.ntiles = static_cast<int>(t.size()) - 1;
  d.tiles.upload(t.data(), t.size() * sizeof(int));
  std::vector<int> lr;
  for (int i = 0; i < d.rows; i++)
    if (rp[static_cast<std::size_t>(i) + 1] - rp[static_cast<std::size_t>(i)] > kCsrLongRow) lr.push_back(i);
  d.nlong = static_cast<int>(lr.size());
  if (d.nlong) d.long_rows.upload(lr.data(), lr.size() * sizeof(int));
}

void build_ell(DeviceMatrix &d, const cpp::MatrixELL &m, Dtype vdt) {
  const auto &vals = m.get_values();
  const auto &cols = m.get_columns();
  const int K = m.get_max_col();
  if (m.get_nnz() == 0 || K == 0) {  // MatrixELL.cpp:102-103
    d.noop = true;
    return;
  }
  const long rows = std::min<long>(m.get_ln_row(), static_cast<long>(vals.size() / static_cast<std::size_t>(K)));
  if (static_cast<long>(cols.size()) < rows * K) bad("ELL columns shorter than values");
  d.rows = static_cast<int>(rows);
  d.K = K;
  d.entries = rows * K;
  d.f_col = m.get_f_col();
  const long slots = (rows + kEllSlice - 1) / kEllSlice * kEllSlice * K;
  std::vector<int> ccm(static_cast<std::size_t>(slots), 0);
  std::vector<double> vcm(static_cast<std::size_t>(slots), 0.0);
  long hi = 0;
  std::mutex mu;
  gen::parallel_for(rows, [&](long b, long e) {
    long h = 0;
    for (long i = b; i < e; i++) {
      const long o = (i / kEllSlice) * K * kEllSlice + i % kEllSlice;
      for (long j = 0; j < K; j++) {
        // MatrixELL.cpp:110 reads x[max(col - f_col, 0)]
        const long c = std::max<long>(static_cast<long>(cols[static_cast<std::size_t>(i * K + j)]) - d.f_col, 0);
        ccm[static_cast<std::size_t>(o + j * kEllSlice)] = static_cast<int>(c);
        vcm[static_cast<std::size_t>(o + j * kEllSlice)] = vals[static_cast<std::size_t>(i * K + j)];
        h = std::max(h, c);
      }
    }
    std::lock_guard<std::mutex> g(mu);
    hi = std::max(hi, h);
  });
  d.x_need = hi + 1;
  d.ecols.upload(ccm.data(), ccm.size() * sizeof(int));
  upload_values(d.evals, vcm.data(), vcm.size(), vdt);
}

template <class M>
void build_coo_like(DeviceMatrix &d, const M &m, Dtype vdt, bool scoo) {
  const auto &vals = m.get_values();
  const auto &row = m.get_row();
  const auto &col = m.get_col();
  const std::size_t n = vals.size();
  if (row.size() < n || col.size() < n) bad("COO row/col shorter than values");
  d.entries = static_cast<long>(n);
  d.f_row = scoo ? m.get_f_row() : 0;
  d.f_col = scoo ? m.get_f_col() : 0;
  if (n == 0) {
    d.noop = true;
    return;
  }
  const MinMax cx = minmax_offset(col.data(), n, -static_cast<long>(d.f_col));
  if (cx.lo < 0) bad("COO column index below the block's first column");
  d.x_need = cx.hi + 1;
  const MinMax ry = minmax_offset(row.data(), n, -static_cast<long>(d.f_row));
  d.y_lo = ry.lo;
  d.y_hi = ry.hi + 1;
  bool sorted = true;
  for (std::size_t k = 1; k < n && sorted; k++)
    if (row[k] < row[k - 1]) sorted = false;
  std::vector<int> srow, scol;
  std::vector<double> sval;
  const int *prow = row.data(), *pcol = col.data();
  const double *pval = vals.data();
  if (!sorted) {
    // stable counting sort by row: each row keeps its entries in input order
    const long lo = ry.lo + d.f_row, span = ry.hi - ry.lo + 1;
    std::vector<long> start(static_cast<std::size_t>(span) + 1, 0);
    for (std::size_t k = 0; k < n; k++) start[static_cast<std::size_t>(row[k] - lo) + 1]++;
    for (long r = 0; r < span; r++) start[static_cast<std::size_t>(r) + 1] += start[static_cast<std::size_t>(r)];
    srow.resize(n);
    scol.resize(n);
    sval.resize(n);
    for (std::size_t k = 0; k < n; k++) {
      const std::size_t p = static_cast<std::size_t>(start[static_cast<std::size_t>(row[k] - lo)]++);
      srow[p] = row[k];
      scol[p] = col[k];
      sval[p] = vals[k];
    }
    prow = srow.data();
    pcol = scol.data();
    pval = sval.data();
  }
  if (n >= (1UL << 31)) bad("COO mirrors are limited to 2^31 entries");
  // segments: maximal runs of one row in the sorted view
  std::vector<int> segptr{0}, segrow;
  for (std::size_t k = 0; k < n; k++) {
    if (k > 0 && prow[k] != prow[k - 1]) segptr.push_back(static_cast<int>(k));
    if (k == 0 || prow[k] != prow[k - 1]) segrow.push_back(prow[k]);
  }
  segptr.push_back(static_cast<int>(n));
  d.nseg = static_cast<int>(segrow.size());
  d.rows = d.nseg;
  d.rp0 = 0;
  d.rowptr.upload(segptr.data(), segptr.size() * sizeof(int));
  d.segrow.upload(segrow.data(), segrow.size() * sizeof(int));
  d.colidx.upload(pcol, n * sizeof(int));
  upload_values(d.vals, pval, n, vdt);
  std::vector<int> t = csr_entry_tiles(segptr.data(), d.nseg, kCsrTileNnz);
  d.ntiles = static_cast<int>(t.size()) - 1;
  d.tiles.upload(t.data(), t.size() * sizeof(int));
  std::vector<int> lr;
  for (int sidx = 0; sidx < d.nseg; sidx++)
    if (segptr[static_cast<std::size_t>(sidx) + 1] - segptr[static_cast<std::size_t>(sidx)] > kCsrLongRow) lr.push_back(sidx);
  d.nlong = static_cast<int>(lr.size());
  if (d.nlong) d.long_rows.upload(lr.data(), lr.size() * sizeof(int));
}

void build_dense(DeviceMatrix &d, const cpp::MatrixDENSE &m, Dtype vdt) {
  const auto &vals = m.get_values();
  if (m.get_nnz() == 0 || vals.empty()) {  // MatrixDENSE.cpp:65-66
    d.noop = true;
    return;
  }
  if (vdt != Dtype::F64) bad("DENSE mirrors are FP64 only");
  const long rows = m.get_ln_row(), cols = m.get_ln_col();
  if (static_cast<long>(vals.size()) < rows * cols) bad("DENSE values shorter than ln_row * ln_col");
  d.rows = static_cast<int>(rows);
  d.dcols = static_cast<int>(cols);
  d.entries = rows * cols;
  d.x_need = cols;
  d.dvals.upload(vals.data(), static_cast<std::size_t>(rows * cols) * sizeof(double));
}

}  // namespace

std::vector<int> csr_entry_tiles(const int *rowptr, int rows, int tile_nnz) {
  const long rp0 = rowptr[0];
  const long nnz = static_cast<long>(rowptr[rows]) - rp0;
  const long ntiles = nnz / tile_nnz + 1;
  std::vector<int> t(static_cast<std::size_t>(ntiles) + 1);
  // tile_row[k] = first row whose first entry is at or after k * tile_nnz
  int r = 0;
  for (long k = 0; k < ntiles; k++) {
    const long e = k * tile_nnz;
    while (r < rows && static_cast<long>(rowptr[r]) - rp0 < e) r++;
    t[static_cast<std::size_t>(k)] = r;
  }
  t[static_cast<std::size_t>(ntiles)] = rows;
  return t;
}


std::size_t DeviceMatrix::algorithmic_bytes(std::size_t xe, std::size_t ye) const {
  const std::size_t ve = dtype_size(vdt);
  switch (format) {
    case cpp::FORMAT_CSR:
      return static_cast<std::size_t>(entries) * (ve + 4) + 4 * (static_cast<std::size_t>(rows) + 1) +
             xe * static_cast<std::size_t>(x_need) + ye * static_cast<std::size_t>(rows);
    case cpp::FORMAT_ELL:
      return static_cast<std::size_t>(entries) * (ve + 4) + xe * static_cast<std::size_t>(x_need) +
             ye * static_cast<std::size_t>(rows);
    case cpp::FORMAT_COO:
    case cpp::FORMAT_SCOO:
      return static_cast<std::size_t>(entries) * (ve + 8) + xe * static_cast<std::size_t>(x_need) +
             ye * static_cast<std::size_t>(y_hi - y_lo);
    case cpp::FORMAT_DENSE:
      return static_cast<std::size_t>(entries) * 8 + xe * static_cast<std::size_t>(dcols) +
             ye * static_cast<std::size_t>(rows);
  }
  return 0;
}

const DeviceMatrix &transpose_mirror(const cpp::Matrix &m) {
  if (m.format() != cpp::FORMAT_CSR) bad("transposed products need a CSR matrix");
  const auto &c = dynamic_cast<const cpp::MatrixCSR &>(m);
  DeviceCache &cache = m.device_cache();
  std::lock_guard<std::mutex> g(cache.mu);
  if (cache.transposed) return *cache.transposed;
  const auto &rp = c.get_rowptr();
  const auto &ci = c.get_colidx();
  auto d = std::unique_ptr<DeviceMatrix>(new DeviceMatrix());
  d->format = cpp::FORMAT_CSR;
  d->vdt = Dtype::PATTERN;
  const int rows = rp.size() < 2 ? 0 : static_cast<int>(rp.size() - 1);
  const int cols = c.get_ln_col();
  const long used = rows ? static_cast<long>(rp.back()) - rp[0] : 0;
  if (used > static_cast<long>(ci.size())) bad("CSR rowptr points past colidx");
  std::vector<int> ptr(static_cast<std::size_t>(cols) + 1, 0);
  for (long k = 0; k < used; k++) {
    const long j = static_cast<long>(ci[static_cast<std::size_t>(k)]) - c.get_f_col();
    if (j < 0 || j >= cols) bad("CSR column index outside the block");
    ptr[static_cast<std::size_t>(j) + 1]++;
  }
  for (int j = 0; j < cols; j++) ptr[static_cast<std::size_t>(j) + 1] += ptr[static_cast<std::size_t>(j)];
  std::vector<int> next(ptr.begin(), ptr.end() - 1), ridx(static_cast<std::size_t>(used)),
      perm(static_cast<std::size_t>(used));
  for (int i = 0; i < rows; i++)
    for (long k = rp[static_cast<std::size_t>(i)] - rp[0]; k < rp[static_cast<std::size_t>(i) + 1] - rp[0]; k++) {
      const int j = ci[static_cast<std::size_t>(k)] - c.get_f_col();
      const int pos = next[static_cast<std::size_t>(j)]++;
      ridx[static_cast<std::size_t>(pos)] = i;
      perm[static_cast<std::size_t>(pos)] = static_cast<int>(k);
    }
  d->rows = cols;
  d->entries = used;
  d->x_need = rows;
  d->rowptr.upload(ptr.data(), ptr.size() * sizeof(int));
  d->colidx.upload(ridx.data(), ridx.size() * sizeof(int));
  d->perm.upload(perm.data(), perm.size() * sizeof(int));
  std::vector<int> t = csr_entry_tiles(ptr.data(), cols, kCsrTileNnz);
  d->ntiles = static_cast<int>(t.size()) - 1;
  d->tiles.upload(t.data(), t.size() * sizeof(int));
  std::vector<int> lr;
  for (int j = 0; j < cols; j++)
    if (ptr[static_cast<std::size_t>(j) + 1] - ptr[static_cast<std::size_t>(j)] > kCsrLongRow) lr.push_back(j);
  d->nlong = static_cast<int>(lr.size());
  if (d->nlong) d->long_rows.upload(lr.data(), lr.size() * sizeof(int));
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  cache.transposed = std::move(d);
  return *cache.transposed;
}

// The host build of the layout (pb_build_gpu builds the same arrays on the
// device; this one is the check of that build, tbsla_hip_pb_layout_check).
std::unique_ptr<PbMatrix> pb_build_host(const cpp::MatrixCSR &c) {
  const auto &rp = c.get_rowptr();
  const auto &ci = c.get_colidx();
  const int n = c.get_n_col();
  if (c.get_n_row() != n || static_cast<int>(rp.size()) != n + 1) bad("the propagation-blocked layout needs a square matrix in one block");
  const long r0 = rp[0];
  const long nnz = static_cast<long>(rp.back()) - r0;
  // entry e of the used range is colidx[e], e = rowptr[i] - rowptr[0] + k (MatrixCSR.cpp:135)
  if (static_cast<long>(ci.size()) < nnz) bad("CSR rowptr points past colidx");
  auto pb = std::unique_ptr<PbMatrix>(new PbMatrix());
  pb->n = n;
  pb->nbins = (n + kPbRows - 1) / kPbRows;
  pb->nsb = (n + kPbCols - 1) / kPbCols;
  const int nb = pb->nbins, ns = pb->nsb;
  // pass 1: counts per bin (far; local groups) and per column (far).  The
  // local groups of a bin are formed in row order: an entry opens a new
  // group when the current one is full or its row is kPbLocalPad or more
  // rows past the group's base row
  std::vector<int> colcnt(static_cast<std::size_t>(n) + 1, 0), bincnt(static_cast<std::size_t>(nb) + 1, 0),
      loccnt(static_cast<std::size_t>(nb) + 1, 0);
  int gfill = kPbLocalGroup, gbase = 0;  // the current bin's open group
  for (int r = 0; r < n; r++) {
    const int b = r >> kPbShift;
    if ((r & (kPbRows - 1)) == 0) gfill = kPbLocalGroup;  // a new bin
    for (long e = rp[static_cast<std::size_t>(r)] - r0; e < rp[static_cast<std::size_t>(r) + 1] - r0; e++) {
      const int col = ci[static_cast<std::size_t>(e)] - c.get_f_col();
      if (col < 0 || col >= n) bad("CSR column index outside the matrix");
      if ((col >> kPbShift) == b) {
        if (gfill == kPbLocalGroup || r - gbase >= kPbLocalPad) {
          loccnt[static_cast<std::size_t>(b) + 1] += kPbLocalGroup;
          gfill = 0;
          gbase = r;
        }
        gfill++;
      } else {
        colcnt[static_cast<std::size_t>(col) + 1]++;
        bincnt[static_cast<std::size_t>(b) + 1]++;
      }
    }
  }
  for (int j = 0; j < n; j++) colcnt[static_cast<std::size_t>(j) + 1] += colcnt[static_cast<std::size_t>(j)];
  for (int b = 0; b < nb; b++) {
    bincnt[static_cast<std::size_t>(b) + 1] += bincnt[static_cast<std::size_t>(b)];
    loccnt[static_cast<std::size_t>(b) + 1] += loccnt[static_cast<std::size_t>(b)];
  }
  pb->n_far = colcnt[static_cast<std::size_t>(n)];
  pb->n_loc = loccnt[static_cast<std::size_t>(nb)];
  if (pb->n_far >= (1L << 31) || pb->n_loc >= (1L << 31)) bad("propagation-blocked layout limited to 2^31 entries");
  const std::size_t nf = static_cast<std::size_t>(pb->n_far), nl = static_cast<std::size_t>(pb->n_loc);
  // pass 2: far entries in column order (rows ascending), local entries
  // bin-grouped (the groups of pass 1; unused slots stay padding)
  std::vector<int> csc_src(nf), csc_row(nf);
  std::vector<int> loc_src_h(nl, -1), far_src_h, far_src_w_h;
  std::vector<unsigned short> lkey(nl, static_cast<unsigned short>(kPbLocalPad << kPbShift));
  std::vector<unsigned short> lbase(nl / kPbLocalGroup, 0);
  std::vector<int> ccur(colcnt.begin(), colcnt.end() - 1);
  long lnext = 0;  // first slot of the next local group
  int lslot = 0;   // next slot in the open group (kPbLocalGroup: none open)
  gfill = kPbLocalGroup;
  for (int r = 0; r < n; r++) {
    const int b = r >> kPbShift;
    if ((r & (kPbRows - 1)) == 0) {
      gfill = kPbLocalGroup;
      lnext = loccnt[static_cast<std::size_t>(b)];
    }
    for (long e = rp[static_cast<std::size_t>(r)] - r0; e < rp[static_cast<std::size_t>(r) + 1] - r0; e++) {
      const int col = ci[static_cast<std::size_t>(e)] - c.get_f_col();
      if ((col >> kPbShift) == b) {
        if (gfill == kPbLocalGroup || r - gbase >= kPbLocalPad) {
          lslot = static_cast<int>(lnext);
          lnext += kPbLocalGroup;
          gfill = 0;
          gbase = r;
          lbase[static_cast<std::size_t>(lslot / kPbLocalGroup)] = static_cast<unsigned short>(r & (kPbRows - 1));
        }
        const std::size_t k = static_cast<std::size_t>(lslot + gfill);
        gfill++;
        loc_src_h[k] = static_cast<int>(e);
        lkey[k] = static_cast<unsigned short>((col & (kPbRows - 1)) | ((r - gbase) << kPbShift));
      } else {
        const int k = ccur[static_cast<std::size_t>(col)]++;
        csc_src[static_cast<std::size_t>(k)] = static_cast<int>(e);
        csc_row[static_cast<std::size_t>(k)] = r;
      }
    }
  }
  // bin-grouped slot of every far entry: bins in order, each in column order
  std::vector<int> slot_csc(nf), bcur(bincnt.begin(), bincnt.end() - 1);
  std::vector<unsigned short> frow(nf);
  for (std::size_t k = 0; k < nf; k++) {
    const int r = csc_row[k];
    const int sl = bcur[static_cast<std::size_t>(r >> kPbShift)]++;
    slot_csc[k] = sl;
    frow[static_cast<std::size_t>(sl)] = static_cast<unsigned short>(r & (kPbRows - 1));
  }
  // scatter order over superblocks of `cols` columns: superblocks in
  // order, within a superblock by (bin, column).  Per superblock, the runs
  // of entries bound for one bin, in bin order: entry o of a run lands in
  // slot o + delta, with delta the run's first slot minus its first entry.
  constexpr int kNoDelta = std::numeric_limits<int>::min();
  auto scatter_side = [&](int cols, int &nsb_out, std::vector<int> &far_src, DevBuf &sp_d, DevBuf &fkey_d,
                          DevBuf &rp_d, DevBuf &rdelta_d, DevBuf &cbase_d, DevBuf &cp_d) {
    const int ns2 = (n + cols - 1) / cols;
    nsb_out = ns2;
    std::vector<int> spc(static_cast<std::size_t>(ns2) + 1, 0);
    for (int s2 = 0; s2 < ns2; s2++)
      spc[static_cast<std::size_t>(s2) + 1] =
          colcnt[static_cast<std::size_t>(std::min(static_cast<long>(n), static_cast<long>(s2 + 1) * cols))];
    far_src.resize(nf);
    std::vector<unsigned short> fkey(nf);
    std::vector<int> runp(static_cast<std::size_t>(ns2) + 1, 0), rdelta;
    std::vector<int> chp(static_cast<std::size_t>(ns2) + 1, 0), cbase;
    std::vector<int> cnt(static_cast<std::size_t>(nb) + 1), erun, run_of(static_cast<std::size_t>(nb));
    for (int s2 = 0; s2 < ns2; s2++) {
      const int k0 = spc[static_cast<std::size_t>(s2)], k1 = spc[static_cast<std::size_t>(s2) + 1];
      std::fill(cnt.begin(), cnt.end(), 0);
      for (int k = k0; k < k1; k++) cnt[static_cast<std::size_t>(csc_row[static_cast<std::size_t>(k)] >> kPbShift) + 1]++;
      // run index of each non-empty bin, and its slot delta (the run's
      // entries are consecutive in CSC order within the bin's slots)
      std::fill(run_of.begin(), run_of.end(), -1);
      int nr = 0;
      for (int b = 0; b < nb; b++)
        if (cnt[static_cast<std::size_t>(b) + 1] > 0) run_of[static_cast<std::size_t>(b)] = nr++;
      for (int b = 0; b < nb; b++) cnt[static_cast<std::size_t>(b) + 1] += cnt[static_cast<std::size_t>(b)];
      const std::vector<int> first(cnt);  // first entry of each bin's run, from k0
      const std::size_t rbase = rdelta.size();
      rdelta.resize(rbase + static_cast<std::size_t>(nr), kNoDelta);
      erun.assign(static_cast<std::size_t>(k1 - k0), -1);
      for (int k = k0; k < k1; k++) {  // column order, so each bin's run stays in column order
        const int b = csc_row[static_cast<std::size_t>(k)] >> kPbShift;
        const int o = k0 + cnt[static_cast<std::size_t>(b)]++;
        const int r = run_of[static_cast<std::size_t>(b)];
        far_src[static_cast<std::size_t>(o)] = csc_src[static_cast<std::size_t>(k)];
        const int col = (ci[static_cast<std::size_t>(csc_src[static_cast<std::size_t>(k)])] - c.get_f_col()) &
                        (cols - 1);
        const bool starts = o - k0 == first[static_cast<std::size_t>(b)];
        fkey[static_cast<std::size_t>(o)] = static_cast<unsigned short>(col | (starts ? kPbRunStart : 0u));
        erun[static_cast<std::size_t>(o - k0)] = r;
        // a bin's slots follow CSC order and a run is one superblock's
        // column range of that order, so slot - o is the same for the
        // whole run
        const int d = slot_csc[static_cast<std::size_t>(k)] - o;
        int &rd = rdelta[rbase + static_cast<std::size_t>(r)];
        if (rd != kNoDelta && rd != d) bad("propagation-blocked layout: run slots are not consecutive");
        rd = d;
      }
      runp[static_cast<std::size_t>(s2) + 1] = static_cast<int>(rdelta.size());
      // run of the first entry of each chunk, checked against the run
      // every entry gets from it and the run-start flags
      const int nck = (k1 - k0 + kPbChunk - 1) / kPbChunk;
      for (int j = 0; j < nck; j++) {
        const int e0 = j * kPbChunk, e1 = std::min(k1 - k0, e0 + kPbChunk);
        int run = erun[static_cast<std::size_t>(e0)];
        cbase.push_back(run);
        for (int e = e0; e < e1; e++) {
          if (e > e0 && (fkey[static_cast<std::size_t>(k0 + e)] & kPbRunStart)) run++;
          if (run != erun[static_cast<std::size_t>(e)]) bad("propagation-blocked layout: run starts do not give the runs");
        }
      }
      chp[static_cast<std::size_t>(s2) + 1] = chp[static_cast<std::size_t>(s2)] + nck;
    }
    sp_d.upload(spc.data(), spc.size() * sizeof(int));
    fkey_d.upload(fkey.data(), fkey.size() * sizeof(unsigned short));
    rp_d.upload(runp.data(), runp.size() * sizeof(int));
    if (rdelta.empty()) rdelta.push_back(0);
    rdelta_d.upload(rdelta.data(), rdelta.size() * sizeof(int));
    if (cbase.empty()) cbase.push_back(0);
    cbase_d.upload(cbase.data(), cbase.size() * sizeof(int));
    cp_d.upload(chp.data(), chp.size() * sizeof(int));
  };
  int ns_check = 0;
  scatter_side(kPbCols, ns_check, far_src_h, pb->sp, pb->fkey, pb->rp, pb->rdelta, pb->cbase, pb->cp);
  if (ns_check != ns) bad("propagation-blocked layout: superblock count");
  scatter_side(kPbColsWide, pb->nsb_w, far_src_w_h, pb->sp_w, pb->fkey_w, pb->rp_w, pb->rdelta_w, pb->cbase_w,
               pb->cp_w);
  pb->frow.upload(frow.data(), frow.size() * sizeof(unsigned short));
  pb->bp.upload(bincnt.data(), bincnt.size() * sizeof(int));
  pb->lp.upload(loccnt.data(), loccnt.size() * sizeof(int));
  pb->lkey.upload(lkey.data(), lkey.size() * sizeof(unsigned short));
  pb->lbase.upload(lbase.data(), lbase.size() * sizeof(unsigned short));
  pb->prod.alloc(sizeof(double) * (nf + kPbChunk));  // + spare slots
  pb->far_src.upload(far_src_h.data(), far_src_h.size() * sizeof(int));
  pb->far_src_w.upload(far_src_w_h.data(), far_src_w_h.size() * sizeof(int));
  pb->loc_src.upload(loc_src_h.data(), loc_src_h.size() * sizeof(int));
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  return pb;
}

std::string pb_layout_compare(const PbMatrix &a, const PbMatrix &b) {
  if (a.n != b.n || a.nbins != b.nbins || a.nsb != b.nsb || a.nsb_w != b.nsb_w || a.n_far != b.n_far ||
      a.n_loc != b.n_loc)
    return "sizes";
  auto dl = [](const DevBuf &d, std::size_t bytes) {
    std::vector<unsigned char> v(bytes);
    if (bytes) TBSLA_HIP_CHECK(hipMemcpy(v.data(), d.get(), bytes, hipMemcpyDeviceToHost));
    return v;
  };
  const std::size_t nf = static_cast<std::size_t>(a.n_far), nl = static_cast<std::size_t>(a.n_loc);
  const std::size_t i4 = sizeof(int), u2 = sizeof(unsigned short);
  auto last = [&](const DevBuf &d, int count) {  // the total an offset table ends with
    int v = 0;
    TBSLA_HIP_CHECK(hipMemcpy(&v, d.as<int>() + count, sizeof(int), hipMemcpyDeviceToHost));
    return static_cast<std::size_t>(v);
  };
  struct Item {
    const char *name;
    const DevBuf &x, &y;
    std::size_t bytes;
  };
  const std::size_t ra = last(a.rp, a.nsb), rb = last(b.rp, b.nsb), ca = last(a.cp, a.nsb), cb = last(b.cp, b.nsb);
  const std::size_t raw = last(a.rp_w, a.nsb_w), rbw = last(b.rp_w, b.nsb_w);
  const std::size_t caw = last(a.cp_w, a.nsb_w), cbw = last(b.cp_w, b.nsb_w);
  if (ra != rb || ca != cb || raw != rbw || caw != cbw) return "run or chunk counts";
  const std::size_t ns1 = static_cast<std::size_t>(a.nsb) + 1, nw1 = static_cast<std::size_t>(a.nsb_w) + 1;
  const std::size_t nb1 = static_cast<std::size_t>(a.nbins) + 1;
  const Item items[] = {
      {"sp", a.sp, b.sp, ns1 * i4},           {"fkey", a.fkey, b.fkey, nf * u2},
      {"rp", a.rp, b.rp, ns1 * i4},           {"rdelta", a.rdelta, b.rdelta, ra * i4},
      {"cbase", a.cbase, b.cbase, ca * i4},   {"cp", a.cp, b.cp, ns1 * i4},
      {"sp_w", a.sp_w, b.sp_w, nw1 * i4},     {"fkey_w", a.fkey_w, b.fkey_w, nf * u2},
      {"rp_w", a.rp_w, b.rp_w, nw1 * i4},     {"rdelta_w", a.rdelta_w, b.rdelta_w, raw * i4},
      {"cbase_w", a.cbase_w, b.cbase_w, caw * i4}, {"cp_w", a.cp_w, b.cp_w, nw1 * i4},
      {"frow", a.frow, b.frow, nf * u2},      {"bp", a.bp, b.bp, nb1 * i4},
      {"lp", a.lp, b.lp, nb1 * i4},           {"lkey", a.lkey, b.lkey, nl * u2},
      {"lbase", a.lbase, b.lbase, nl / kPbLocalGroup * u2},
      {"far_src", a.far_src, b.far_src, nf * i4}, {"far_src_w", a.far_src_w, b.far_src_w, nf * i4},
      {"loc_src", a.loc_src, b.loc_src, nl * i4}};
  for (const Item &it : items)
    if (dl(it.x, it.bytes) != dl(it.y, it.bytes)) return it.name;
  return "";
}

const PbMatrix &pb_mirror(const cpp::Matrix &m, Dtype vdt) {
  if (m.format() != cpp::FORMAT_CSR) bad("the propagation-blocked layout needs a CSR matrix");
  const auto &c = dynamic_cast<const cpp::MatrixCSR &>(m);
  const int n = c.get_n_col();
  if (c.get_n_row() != n || static_cast<int>(c.get_rowptr().size()) != n + 1)
    bad("the propagation-blocked layout needs a square matrix in one block");
  // built from the FP64 mirror (taken before the cache lock: mirror() locks it)
  const DeviceMatrix &d64 = mirror(m, Dtype::F64);
  DeviceCache &cache = m.device_cache();
  std::lock_guard<std::mutex> g(cache.mu);
  if (!cache.pb) cache.pb = pb_build_gpu(d64, n);
  PbMatrix &pb = *cache.pb;
  if (vdt != Dtype::PATTERN && !pb.fval[static_cast<int>(vdt)].get()) pb_build_values(pb, d64, vdt);
  return pb;
}

const DeviceMatrix *mirror_if_built(const cpp::Matrix &m, Dtype vdt) {
  DeviceCache &cache = m.device_cache();
  std::lock_guard<std::mutex> g(cache.mu);
  return cache.mirrors[static_cast<int>(vdt)].get();
}

const DeviceMatrix &mirror(const cpp::Matrix &m, Dtype vdt) {
  DeviceCache &cache = m.device_cache();
  std::lock_guard<std::mutex> g(cache.mu);
  auto &slot = cache.mirrors[static_cast<int>(vdt)];
  if (slot) return *slot;
  auto d = std::unique_ptr<DeviceMatrix>(new DeviceMatrix());
  d->format = m.format();
  d->vdt = vdt;
  d->f_row = m.get_f_row();
  d->f_col = m.get_f_col();
  switch (m.format()) {
    case cpp::FORMAT_CSR: build_csr(*d, dynamic_cast<const cpp::MatrixCSR &>(m), vdt); break;
    case cpp::FORMAT_ELL: build_ell(*d, dynamic_cast<const cpp::MatrixELL &>(m), vdt); break;
    case cpp::FORMAT_COO: build_coo_like(*d, dynamic_cast<const cpp::MatrixCOO &>(m), vdt, false); break;
    case cpp::FORMAT_SCOO: build_coo_like(*d, dynamic_cast<const cpp::MatrixSCOO &>(m), vdt, true); break;
    case cpp::FORMAT_DENSE: build_dense(*d, dynamic_cast<const cpp::MatrixDENSE &>(m), vdt); break;
  }
  TBSLA_HIP_CHECK(hipDeviceSynchronize());
  slot = std::move(d);
  return *slot;
}

}  // namespace hip
}  // namespace tbsla
