// Adjoint (reverse-mode derivative) of PageRank on a CSR matrix.  The
// reference has no counterpart; the forward quantity is the fixed point of
// the reference's iteration (src/cpp/matrix/Matrix.cpp:13-61):
//
//   p = M p / (1^T M p),   M = beta*A + (1-beta) v 1^T,   1^T p = 1
//
// (v = 1/n for the reference's uniform teleport).  For L(p) with w = dL/dp,
// write s = 1^T M p.  Differentiating the fixed point gives
//   dL = g^T dM p,   g = (lambda - (p^T lambda) 1) / s,
// where lambda solves the adjoint fixed point
//   lambda = M^T (lambda - (p^T lambda) 1) / s + w.
// Hence
//   dL/dA_e = beta * g_{row(e)} * p_{col(e)}       (every stored entry e)
//   dL/dv_k = (1-beta) * (1^T p) * g_k.
// The iteration contracts like the forward one (rate ~ beta): the
// p-direction is annihilated by (I - 1 p^T) and the rest of M^T / s has
// spectral radius < 1.
//
// Device work per iteration: u = (lambda - a) / s in the product precision,
// y = A^T u on the transposed CSR structure (same tile kernel as the forward
// product), lambda' = beta*y + (1-beta)*(v^T u) + w fused with the partial
// sums p^T lambda', v^T lambda', |lambda' - lambda|_1 and |lambda'|_1.
#pragma once

#include <tbsla/cpp/Matrix.hpp>

#include "common.hpp"

namespace tbsla {
namespace hip {

struct AdjointParams {
  double beta = 0.85;
  // stop when |lambda_k - lambda_{k-1}|_1 <= epsilon * |lambda_k|_1
  double epsilon = 1e-7;
  int max_iteration = 1000;
  Dtype precision = Dtype::F32;  // transposed products in F32 (values, u, sums) or F64
  int check_every = 8;
  const double *values = nullptr;    // device, CSR entry order; nullptr: stored values
  const double *teleport = nullptr;  // device, n entries; nullptr: uniform 1/n
};

struct AdjointReport {
  int iterations = 0;
  double last_change = 0;  // |lambda_k - lambda_{k-1}|_1 / |lambda_k|_1
  double s = 0;            // 1^T M p
  double seconds = 0;      // device time of the whole adjoint (hipEvents)
};

// pr, w: device FP64 vectors of n_col entries (the forward fixed point,
// sum 1, and dL/dp).  grad_values: device FP64, one per stored CSR entry
// (rowptr[rows] - rowptr[0] entries); grad_teleport: device FP64, n_col.
// Either output may be nullptr.
AdjointReport page_rank_adjoint(const cpp::Matrix &m, const AdjointParams &p, const double *pr, const double *w,
                                double *grad_values, double *grad_teleport, hipStream_t stream);

}  // namespace hip
}  // namespace tbsla
