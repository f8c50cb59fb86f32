// GPU PageRank engine behind tbsla::cpp::Matrix::page_rank and
// personalized_page_rank (reference: src/cpp/matrix/Matrix.cpp:13-114).
//
// Per iteration (reference order): b_t = b; b = A b_t; tele = (1-beta)/n *
// sum(b_t) (personalised: (1-beta)/|nodes|, added to listed nodes only);
// b = beta*b + tele; b /= max(b); err = sum|b - b_t|; stop when err < eps.
// At the end b /= sum(b).
//
// Precision modes (CSR matrices):
//   F64      values and vectors in FP64 (the plain path).
//   F32/BF16 values stored in FP32 / BF16, products and row sums in FP32, on
//            a correction delta to an FP64 anchor: A b = A64 xhat + A32 delta.
//            The anchor product A64 xhat, the update, max, L1 residual and the
//            compensated sum normalisation are FP64.  The anchor is moved
//            (xhat += delta, delta = 0) when the residual has fallen by a
//            factor tau since the last anchor, so the FP32 error, which scales
//            with |delta|, stays below the target.
//   PATTERN  FP64 without a value array for count-stochastic matrices
//            (value(i,j) = 1/count_j): the normalise pass writes
//            z = colscale .* b, the next product sums gathered z.  Bit-identical
//            products to F64.
//   PATTERN_F32  the mixed scheme of F32 without a value array, for
//            count-stochastic matrices: the normalise pass writes, beside the
//            FP32 delta, zd = colscale .* delta rounded to FP32, and the FP32
//            product sums gathered zd (no value bytes, FP32 products); the
//            anchor product is the value-free FP64 one.
// Stopping rules: REFERENCE (err < epsilon, the reference's quantity) or
// SUMNORM (a bound on the L1 distance between successive sum-normalised
// iterates, err/S_k + |S_{k-1}/S_k - 1| <= epsilon * (1-beta)/beta).
#pragma once

#include <tbsla/cpp/Matrix.hpp>

#include "common.hpp"

namespace tbsla {
namespace hip {

enum class PrPrecision : int { F64 = 0, F32 = 1, BF16 = 2, PATTERN = 3, PATTERN_F32 = 4 };
enum class PrRule : int { REFERENCE = 0, SUMNORM = 1 };

struct PageRankParams {
  double epsilon = 1e-10;
  double beta = 0.85;
  int max_iteration = 100;
  bool personalized = false;
  const int *nodes = nullptr;
  int n_nodes = 0;
  PrPrecision precision = PrPrecision::F64;
  PrRule rule = PrRule::REFERENCE;
  int check_every = 8;      // iterations launched between convergence reads
  double anchor_tau = 0;    // 0: default per precision
  // Optional device inputs (FP64, CSR only):
  //   values    edge values in CSR entry order, used instead of the stored
  //             ones (precision F64)
  //   teleport  teleport distribution v (n_col entries): the update adds
  //             (1-beta) * sum(b_t) * v[i], i.e. M = beta*A + (1-beta) v 1^T
  const double *values = nullptr;
  const double *teleport = nullptr;
  // CSR product structure: 0 = propagation-blocked on n >= 2^16 (else
  // tiles), 1 = fixed-order entry tiles, 2 = propagation-blocked
  int gather = 0;
};

struct PageRankReport {
  int iterations = 0;
  int anchors = 0;          // FP64 anchor products (mixed precision)
  double last_error = 0;    // reference quantity of the last iteration
  double bound = 0;         // SUMNORM bound of the last iteration
  double seconds = 0;       // device time of the iteration loop (hipEvents)
};

// Result (sum-normalised, FP64) copied to host memory `out` (n_col entries).
PageRankReport page_rank(const cpp::Matrix &m, const PageRankParams &p, double *out);
// Result left in device memory `out_dev` (n_col doubles) on `stream`.
PageRankReport page_rank_device(const cpp::Matrix &m, const PageRankParams &p, double *out_dev, hipStream_t stream);

// page_rank_device with no device allocation: every buffer of the run is a
// slice of a fixed static scratch buffer of kNoAllocScratchBytes per GPU.
// FP64 with the fixed-order tile products only; the FP64 mirror must exist
// (built by an earlier product or tbsla_hip_prepare).  Throws
// std::invalid_argument when the run needs more scratch than the buffer
// holds (page_rank_noalloc_bytes).  Calls that overlap on one GPU share the
// buffer and race.
constexpr std::size_t kNoAllocScratchBytes = std::size_t(40) << 20;
std::size_t page_rank_noalloc_bytes(long n, int n_nodes);
PageRankReport page_rank_device_noalloc(const cpp::Matrix &m, const PageRankParams &p, double *out_dev,
                                        hipStream_t stream);

}  // namespace hip
}  // namespace tbsla
