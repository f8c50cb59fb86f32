#include "products.hpp"

#include <mutex>
#include <stdexcept>
#include <string>

#include "kernels.hpp"
#include "mirror.hpp"

namespace tbsla {
namespace hip {
namespace {

// Rows of r the product reads or writes: [lo, hi).
void needed_rows(const DeviceMatrix &d, int vect_incr, long &lo, long &hi, bool &accumulate) {
  switch (d.format) {
    case cpp::FORMAT_CSR:
    case cpp::FORMAT_ELL:
      lo = 0;
      hi = d.rows;
      accumulate = false;
      return;
    case cpp::FORMAT_DENSE:
      lo = 0;
      hi = d.rows;
      accumulate = true;
      return;
    case cpp::FORMAT_COO:
    case cpp::FORMAT_SCOO:
      lo = d.y_lo + vect_incr;
      hi = d.y_hi + vect_incr;
      accumulate = true;
      return;
  }
}

void check_sizes(const DeviceMatrix &d, std::size_t nx, std::size_t ny, int vect_incr) {
  if (static_cast<long>(nx) < d.x_need)
    throw std::invalid_argument("tbsla: x has " + std::to_string(nx) + " entries, the matrix reads " +
                                std::to_string(d.x_need));
  long lo, hi;
  bool acc;
  needed_rows(d, vect_incr, lo, hi, acc);
  if (lo < 0 || hi > static_cast<long>(ny))
    throw std::invalid_argument("tbsla: result vector has " + std::to_string(ny) + " entries, the product writes rows [" +
                                std::to_string(lo) + ", " + std::to_string(hi) + ")");
}

}  // namespace

void Ax_host(const cpp::Matrix &m, std::vector<double> &r, const std::vector<double> &v, int vect_incr) {
  const DeviceMatrix &d = mirror(m, Dtype::F64);
  if (d.noop) return;
  check_sizes(d, v.size(), r.size(), vect_incr);
  long lo, hi;
  bool acc;
  needed_rows(d, vect_incr, lo, hi, acc);
  if (hi <= lo) return;
  DeviceCache &c = m.device_cache();
  std::lock_guard<std::mutex> g(c.mu);
  c.x_stage.alloc(v.size() * sizeof(double));
  c.y_stage.alloc(r.size() * sizeof(double));
  double *dx = c.x_stage.as<double>(), *dy = c.y_stage.as<double>();
  TBSLA_HIP_CHECK(hipMemcpy(dx, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  if (acc)
    TBSLA_HIP_CHECK(hipMemcpy(dy + lo, r.data() + lo, static_cast<std::size_t>(hi - lo) * sizeof(double),
                              hipMemcpyHostToDevice));
  launch_spmv(d, dx, Dtype::F64, dy, Dtype::F64, vect_incr, nullptr);
  TBSLA_HIP_CHECK(hipMemcpy(r.data() + lo, dy + lo, static_cast<std::size_t>(hi - lo) * sizeof(double),
                            hipMemcpyDeviceToHost));
}

void Ax_device(const cpp::Matrix &m, Dtype vdt, const void *x, std::size_t nx, Dtype xdt, void *y, std::size_t ny,
               Dtype ydt, int vect_incr, hipStream_t stream) {
  const DeviceMatrix &d = mirror(m, vdt);
  if (d.noop) return;
  check_sizes(d, nx, ny, vect_incr);
  launch_spmv(d, x, xdt, y, ydt, vect_incr, stream);
}

}  // namespace hip
}  // namespace tbsla
