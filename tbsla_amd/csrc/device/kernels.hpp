// Host-callable launchers of the SpMV kernels (spmv.hip).
#pragma once

#include "mirror.hpp"

namespace tbsla {
namespace hip {

// One product with the reference's Ax semantics on device vectors:
//   CSR, ELL   y[i]  = sum_j A(i,j) x(j)   for the rows the matrix defines
//   COO, SCOO  y[r + vect_incr (- f_row)] += A(r,c) x(c (- f_col)), entry order per row
//   DENSE      y[i] += sum_j A(i,j) x(j)
// Supported element types: values F64 with x, y F64; values F32/BF16 with
// x, y F32 (float accumulation).  Throws on an unsupported combination.
void launch_spmv(const DeviceMatrix &d, const void *x, Dtype xdt, void *y, Dtype ydt, int vect_incr,
                 hipStream_t s);

// y[0:n) = value
void launch_fill(void *y, Dtype ydt, long n, double value, hipStream_t s);

// Kernel symbol names (for profiles): one per format.
const char *spmv_kernel_name(cpp::Format f);

}  // namespace hip
}  // namespace tbsla
