// Host-vector products on the GPU: the bodies of tbsla::cpp::Matrix::Ax /
// spmv for every format (x copied to HBM, kernel, result copied back).
#pragma once

#include <tbsla/cpp/Matrix.hpp>

#include <vector>

#include "common.hpp"

namespace tbsla {
namespace hip {

// Reference Ax semantics (assign for CSR/ELL, accumulate for COO/SCOO/DENSE).
// Throws std::invalid_argument when v or r is too short for the indices the
// matrix holds (the reference would read or write out of bounds).
void Ax_host(const cpp::Matrix &m, std::vector<double> &r, const std::vector<double> &v, int vect_incr);

// Device-vector product on `stream` (Ax semantics); sizes in elements.
void Ax_device(const cpp::Matrix &m, Dtype vdt, const void *x, std::size_t nx, Dtype xdt, void *y, std::size_t ny,
               Dtype ydt, int vect_incr, hipStream_t stream);

}  // namespace hip
}  // namespace tbsla
