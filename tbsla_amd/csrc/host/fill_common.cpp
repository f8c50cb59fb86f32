#include "fill_common.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace tbsla {
namespace host {

RowEntries generate_rows(const CqmatBlock &b, long r0, long r1, bool want_rows) {
  RowEntries out;
  const long nrows = std::max(0L, r1 - r0);
  out.counts.assign(static_cast<std::size_t>(nrows), 0);
  if (nrows == 0) return out;
  const int nt = gen::host_threads();
  const long chunks = std::max(1L, std::min<long>(nt * 4, nrows / 1024 + 1));
  struct Part {
    std::vector<int> rows, cols;
    std::vector<double> vals;
  };
  std::vector<Part> parts(static_cast<std::size_t>(chunks));
  gen::parallel_for(chunks, [&](long cb, long ce) {
    for (long ch = cb; ch < ce; ch++) {
      const long ra = r0 + nrows * ch / chunks, rb = r0 + nrows * (ch + 1) / chunks;
      Part &p = parts[static_cast<std::size_t>(ch)];
      long est = 0;
      for (long r = ra; r < rb; r++) est += b.s.candidates(r);
      p.cols.reserve(static_cast<std::size_t>(est));
      p.vals.reserve(static_cast<std::size_t>(est));
      if (want_rows) p.rows.reserve(static_cast<std::size_t>(est));
      for (long r = ra; r < rb; r++) {
        const long i0 = b.s.base(r), nc = b.s.candidates(r);
        int kept = 0;
        for (long k = 0; k < nc; k++) {
          long ii, jj;
          b.entry(i0 + k, ii, jj);
          if (b.in_block(ii, jj)) {
            p.cols.push_back(static_cast<int>(jj));
            p.vals.push_back(b.value(i0 + k));
            if (want_rows) p.rows.push_back(static_cast<int>(ii));
            kept++;
          }
        }
        out.counts[static_cast<std::size_t>(r - r0)] = kept;
      }
    }
  }, 1);
  std::vector<std::size_t> off(parts.size() + 1, 0);
  for (std::size_t k = 0; k < parts.size(); k++) off[k + 1] = off[k] + parts[k].cols.size();
  const std::size_t tot = off.back();
  out.cols.resize(tot);
  out.vals.resize(tot);
  if (want_rows) out.rows.resize(tot);
  gen::parallel_for(static_cast<long>(parts.size()), [&](long cb, long ce) {
    for (long k = cb; k < ce; k++) {
      const Part &p = parts[static_cast<std::size_t>(k)];
      const std::size_t o = off[static_cast<std::size_t>(k)];
      if (p.cols.empty()) continue;
      std::memcpy(out.cols.data() + o, p.cols.data(), p.cols.size() * sizeof(int));
      std::memcpy(out.vals.data() + o, p.vals.data(), p.vals.size() * sizeof(double));
      if (want_rows) std::memcpy(out.rows.data() + o, p.rows.data(), p.rows.size() * sizeof(int));
    }
  }, 1);
  return out;
}

long overcount_nnz(const CqmatBlock &b, long kept_full_rows) {
  // rows [f_row, lim3) walked with stride cc from base(f_row)
  const long lim3 = b.lim3();
  const long first_partial = std::max(b.f_row, b.s.min_);
  long nnz = kept_full_rows;
  const long incr_save = b.s.base(b.f_row);
  for (long r = first_partial; r < lim3; r++) {
    const long i0 = incr_save + (r - b.f_row) * b.s.cc;
    for (long j = 0; j < b.s.cc; j++) {
      long ii, jj;
      b.entry(i0 + j, ii, jj);
      if (b.in_block(ii, jj)) nnz++;
    }
  }
  return nnz;
}

long cdiag_block_count(long f_row, long ln_row, long f_col, long ln_col, int cdiag) {
  auto inb = [&](long ii, long jj) {
    return ii >= f_row && ii < f_row + ln_row && jj >= f_col && jj < f_col + ln_col;
  };
  long nv = 0;
  for (long i = f_row; i < f_row + ln_row; i++) {
    if (inb(i, i - cdiag)) nv++;
    if (cdiag != 0 && inb(i, i + cdiag)) nv++;
  }
  return nv;
}

void print_dense_matrix(int nr, int nc, const std::vector<double> &m, std::ostream &os) {
  int max_p = 1;
  for (long i = 0; i < static_cast<long>(nr) * nc; i++) {
    int log_r = static_cast<int>(std::log10(std::fabs(m[static_cast<std::size_t>(i)])) + 1);
    if (log_r > max_p) max_p = log_r;
  }
  os << std::fixed;
  os.precision(max_p);
  for (int i = 0; i < nr; i++) {
    for (int j = 0; j < nc; j++) os << std::setw(max_p + 3) << m[static_cast<std::size_t>(i) * nc + j] << "  ";
    os << std::endl;
  }
}

std::ostream &print_header(std::ostream &os, const char *title, int n_row, int n_col, int ln_row, int ln_col,
                           int f_row, int f_col, long nnz, int pr, int pc, int NR, int NC) {
  os << "-----------------" << std::endl;
  os << title << std::endl;
  os << "-----------------" << std::endl;
  os << "number of rows             -  n_row   : " << n_row << std::endl;
  os << "number of columns          -  n_col   : " << n_col << std::endl;
  os << "local number of rows       -  ln_row  : " << ln_row << std::endl;
  os << "local number of columns    -  ln_col  : " << ln_col << std::endl;
  os << "first row                  -  f_row   : " << f_row << std::endl;
  os << "first column               -  f_col   : " << f_col << std::endl;
  os << "number of non-zero elts    -  nnz     : " << nnz << std::endl;
  os << "block position (row)       -  pr      : " << pr << std::endl;
  os << "block position (column)    -  pc      : " << pc << std::endl;
  os << "number of blocks (row)     -  NR      : " << NR << std::endl;
  os << "number of blocks (column)  -  NC      : " << NC << std::endl;
  return os;
}

std::vector<int> row_col_order(int n_row, const std::vector<int> &row, const std::vector<int> &col) {
  // counting sort by row (stable), then a stable sort by column inside each row
  const std::size_t n = row.size();
  std::vector<int> start(static_cast<std::size_t>(n_row) + 1, 0);
  for (std::size_t k = 0; k < n; k++) start[static_cast<std::size_t>(row[k]) + 1]++;
  for (int r = 0; r < n_row; r++) start[static_cast<std::size_t>(r) + 1] += start[static_cast<std::size_t>(r)];
  std::vector<int> pos(start.begin(), start.end() - 1), perm(n);
  for (std::size_t k = 0; k < n; k++) perm[static_cast<std::size_t>(pos[static_cast<std::size_t>(row[k])]++)] = static_cast<int>(k);
  gen::parallel_for(n_row, [&](long rb, long re) {
    for (long r = rb; r < re; r++) {
      auto b = perm.begin() + start[static_cast<std::size_t>(r)], e = perm.begin() + start[static_cast<std::size_t>(r) + 1];
      std::stable_sort(b, e, [&](int i, int j) { return col[static_cast<std::size_t>(i)] < col[static_cast<std::size_t>(j)]; });
    }
  }, 1024);
  return perm;
}

void write_header(std::ostream &os, int n_row, int n_col, int ln_row, int ln_col, int f_row, int f_col, long nnz,
                  int pr, int pc, int NR, int NC) {
  const int h1[6] = {n_row, n_col, ln_row, ln_col, f_row, f_col};
  const int h2[4] = {pr, pc, NR, NC};
  os.write(reinterpret_cast<const char *>(h1), sizeof(h1));
  os.write(reinterpret_cast<const char *>(&nnz), sizeof(nnz));
  os.write(reinterpret_cast<const char *>(h2), sizeof(h2));
}

void read_header(std::istream &is, int &n_row, int &n_col, int &ln_row, int &ln_col, int &f_row, int &f_col,
                 long &nnz, int &pr, int &pc, int &NR, int &NC) {
  int h1[6] = {0, 0, 0, 0, 0, 0}, h2[4] = {0, 0, 1, 1};
  is.read(reinterpret_cast<char *>(h1), sizeof(h1));
  is.read(reinterpret_cast<char *>(&nnz), sizeof(nnz));
  is.read(reinterpret_cast<char *>(h2), sizeof(h2));
  if (!is) throw std::runtime_error("tbsla: truncated binary file (header)");
  n_row = h1[0]; n_col = h1[1]; ln_row = h1[2]; ln_col = h1[3]; f_row = h1[4]; f_col = h1[5];
  pr = h2[0]; pc = h2[1]; NR = h2[2]; NC = h2[3];
}

}  // namespace host
}  // namespace tbsla
