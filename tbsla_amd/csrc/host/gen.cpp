#include "gen.hpp"

#include <atomic>
#include <thread>

namespace tbsla {
namespace gen {

// range.cpp:7-15
std::size_t lnv(std::size_t size, int local_pos, int number_pos) {
  std::size_t n = size / static_cast<std::size_t>(number_pos);
  std::size_t mod = size % static_cast<std::size_t>(number_pos);
  if (static_cast<std::size_t>(local_pos) < mod) n++;
  return n;
}

// range.cpp:21-26
std::size_t pflv(std::size_t size, int local_pos, int number_pos) {
  std::size_t mod = size % static_cast<std::size_t>(number_pos);
  std::size_t n = lnv(size, local_pos, number_pos) * static_cast<std::size_t>(local_pos);
  if (static_cast<std::size_t>(local_pos) >= mod) n += mod;
  return n;
}

// values_generation.cpp:4-24
void cdiag_value(std::size_t i, std::size_t, std::size_t nc, std::size_t cdiag, std::uint64_t &row,
                 std::uint64_t &col) {
  if (cdiag == 0) {
    row = i;
    col = i;
    return;
  }
  const long snc = static_cast<long>(nc), sc = static_cast<long>(cdiag);
  const long t1 = std::max(std::min(snc - sc, sc), 0L);
  if (i < static_cast<std::size_t>(t1)) {
    row = i;
    col = i + cdiag;
    return;
  }
  const long t2 = std::max(sc + 2 * (snc - 2 * sc), 0L);
  if (i < static_cast<std::size_t>(t2)) {
    const std::uint64_t it = (i + cdiag) / 2;
    row = it;
    col = (i % 2 == 0) ? it - cdiag : it + cdiag;
    return;
  }
  std::uint64_t it = i - static_cast<std::size_t>(snc - 2 * sc);
  if (cdiag > nc) it -= cdiag - nc;
  row = it;
  col = it - cdiag;
}

std::uint64_t splitmix64(std::uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int host_threads() {
  unsigned n = std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > 64) n = 64;
  return static_cast<int>(n);
}

void parallel_for(long n, const std::function<void(long, long)> &f, long min_chunk) {
  if (n <= 0) return;
  int nt = host_threads();
  long chunks = std::min<long>(nt, (n + min_chunk - 1) / min_chunk);
  if (chunks <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(static_cast<std::size_t>(chunks));
  for (long t = 0; t < chunks; t++) {
    long b = n * t / chunks, e = n * (t + 1) / chunks;
    th.emplace_back([&f, b, e] { f(b, e); });
  }
  for (auto &t : th) t.join();
}

std::vector<double> cqmat_column_sums(std::size_t nr, std::size_t nc, std::size_t c, double q,
                                      unsigned seed_mult) {
  std::size_t gnv = 0;
  const std::size_t cc = std::min(c, nc);
  // values_generation.cpp:48-55; the second loop's bound is unsigned there
  const std::size_t nfull = std::min(nc - cc + 1, nr);
  gnv = cc * nfull;
  const std::size_t lim = std::min(nr, nc);
  if (lim > nfull)
    for (std::size_t i = 0; i < lim - nfull; i++) gnv += cc - i - 1;
  std::vector<std::atomic<std::uint32_t>> cnt(nc);
  for (auto &a : cnt) a.store(0, std::memory_order_relaxed);
  parallel_for(static_cast<long>(gnv), [&](long b, long e) {
    for (long i = b; i < e; i++) {
      std::uint64_t r, cl;
      cqmat_value(static_cast<std::size_t>(i), nr, nc, c, q, seed_mult, r, cl);
      cnt[cl].fetch_add(1, std::memory_order_relaxed);
    }
  }, 1 << 16);
  std::vector<double> sum(nc);
  for (std::size_t j = 0; j < nc; j++) sum[j] = static_cast<double>(cnt[j].load(std::memory_order_relaxed));
  return sum;
}

}  // namespace gen
}  // namespace tbsla
