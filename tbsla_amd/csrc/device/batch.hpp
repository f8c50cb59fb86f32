// Batched personalised PageRank: many graphs of uneven size and many
// teleport vectors per graph, advanced together.  Each right-hand side
// (RHS) follows Matrix::personalized_page_rank (src/cpp/matrix/
// Matrix.cpp:63-114): b = 0.25, then per iteration b_t = b; b = A b_t;
// t = (1-beta) * sum(b_t) / |nodes|, added to the listed nodes; b /= max(b);
// err = sum|b - b_t|; stop when err < epsilon (or the sum-normalised bound);
// finally b /= sum(b).  Every RHS keeps its own state and stops on its own.
//
// Device layout:
//   single-RHS graphs  one block-diagonal CSR (graphs concatenated, columns
//                      offset), entry tiles as in the single-graph product
//                      but of 2048 entries; iterate vectors concatenated.
//   multi-RHS graphs   one CSR per graph with up to 64 RHS; iterates stored
//                      row-major n x 64 so a gathered row of X is 64
//                      consecutive values (one coalesced load per wavefront).
//
// One iteration is four launches, captured in a hipGraph and replayed:
//   batch_step          persistent work-queue kernel over every tile of the
//                       block-diagonal matrix and every 64-row block of the
//                       multi-RHS graphs (SpMM: lane = RHS), fused with
//                       b = beta*(A b_t) + t and per-RHS maxima (atomic max
//                       slots); per-XCD queues keep a graph's tiles on one
//                       XCD so its iterate stays in that XCD's L2
//   batch_finalize_max  per-RHS maximum
//   batch_pass<0>       b /= max, residual and compensated sum partials
//   batch_finalize_norm per-RHS decision, teleport term for the next step
// The launch shape (workgroups per CU, queue mapping, gathers in flight)
// is chosen by an autotuner that times candidate shapes on the first
// iterations of the batch's own run (each candidate runs one real
// iteration), and is cached by batch signature for later batches.
#pragma once

#include <tbsla/cpp/Matrix.hpp>

#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "pagerank.hpp"

namespace tbsla {
namespace hip {

struct BatchParams {
  double epsilon = 1e-10;
  double beta = 0.85;
  int max_iteration = 1000;
  PrRule rule = PrRule::REFERENCE;
  int check_every = 16;      // iterations per captured graph / between state reads
  bool graph_capture = true; // replay a hipGraph of check_every iterations
  bool autotune = true;      // time candidate launch shapes on first use
};

struct BatchConfig {
  int wg_per_cu = 4;  // persistent workgroups per compute unit
  int unroll = 8;     // gathers in flight per lane in the SpMM (4 or 8)
  int xcd_map = 1;    // 1: graphs / row ranges pinned to XCD queues; 0: interleaved
  int tile_nnz = 2048;  // entries per tile of the single-set graphs (1024, 2048, 3072)
  int spmm_rows = 64;   // rows per SpMM work item (32, 64, 128)
  std::string str() const;
};

struct BatchReport {
  int iterations = 0;           // the most iterations any RHS ran
  double seconds = 0;           // device time of the iteration loop
  double tune_seconds = 0;      // host time of the autotuner's trial iterations (first run), part of seconds
  int tune_iterations = 0;      // iterations of this run spent as autotuner trials
  bool tune_cached = false;     // the launch shape came from the cache (a batch of the same signature)
  BatchConfig config;
  std::vector<int> rhs_iterations;
  std::vector<double> rhs_error;  // last reference residual of each RHS
  double bytes_per_iteration = 0; // algorithmic bytes of one step with every RHS active
};

class PageRankBatch {
 public:
  explicit PageRankBatch(Dtype precision);  // F64 or F32 (values, iterates, products)
  ~PageRankBatch();
  PageRankBatch(const PageRankBatch &) = delete;
  PageRankBatch &operator=(const PageRankBatch &) = delete;

  // A square single-block CSR graph with one teleport node set per RHS
  // (one set: the block-diagonal product; several: the SpMM product, in
  // groups of up to 64).  The matrix is read by prepare() only.
  void add_graph(const cpp::Matrix &m, const std::vector<std::vector<int>> &rhs_nodes);
  // Build the device structures (idempotent; called by run()).
  void prepare();
  // Results, sum-normalised FP64, RHS after RHS in add order, into out_dev.
  BatchReport run(const BatchParams &p, double *out_dev, hipStream_t s);

  int n_rhs() const;
  // offset of each RHS's vector in the output (n_rhs + 1 entries)
  std::vector<long> rhs_offsets() const;
  const BatchConfig &config() const;

  struct Impl;

 private:
  std::unique_ptr<Impl> impl_;
};

}  // namespace hip
}  // namespace tbsla
