// Device-side plumbing shared by the mirrors and the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace tbsla {
namespace hip {

// Value / vector element types understood by the kernels.
enum class Dtype : int {
  F64 = 0,      // double
  F32 = 1,      // float
  BF16 = 2,     // bfloat16 storage (values only), float arithmetic
  PATTERN = 3,  // no value array: value(i, j) = colscale[j] (count-stochastic matrices)
};

inline std::size_t dtype_size(Dtype d) {
  switch (d) {
    case Dtype::F64: return 8;
    case Dtype::F32: return 4;
    case Dtype::BF16: return 2;
    case Dtype::PATTERN: return 0;
  }
  return 0;
}

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};

#define TBSLA_HIP_CHECK(expr)                                                                        \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess)                                                                            \
      throw ::tbsla::hip::HipError(_e, std::string(#expr) + ": " + hipGetErrorString(_e) + " at " + \
                                           __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

// Owning device allocation, padded to 256 bytes so vector loads that round
// an index range up never leave the allocation.
class DevBuf {
 public:
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) {
    o.p_ = nullptr;
    o.bytes_ = 0;
  }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_;
      bytes_ = o.bytes_;
      o.p_ = nullptr;
      o.bytes_ = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }

  void alloc(std::size_t bytes) {
    if (bytes <= bytes_ && p_) return;
    release();
    std::size_t padded = ((bytes ? bytes : 1) + 255 + 256) & ~std::size_t(255);
    TBSLA_HIP_CHECK(hipMalloc(&p_, padded));
    TBSLA_HIP_CHECK(hipMemset(p_, 0, padded));
    bytes_ = padded - 256;
  }
  void upload(const void *src, std::size_t bytes, hipStream_t s = nullptr) {
    alloc(bytes);
    if (bytes) TBSLA_HIP_CHECK(hipMemcpyAsync(p_, src, bytes, hipMemcpyHostToDevice, s));
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    bytes_ = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(p_);
  }
  void *get() const { return p_; }
  std::size_t bytes() const { return bytes_; }

 private:
  void *p_ = nullptr;
  std::size_t bytes_ = 0;
};

}  // namespace hip
}  // namespace tbsla
