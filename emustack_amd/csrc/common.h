// Shared HIP helpers: error checking, complex arithmetic on double2.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace emu {

struct HipError : std::runtime_error {
  explicit HipError(const std::string &s) : std::runtime_error(s) {}
};

#define EMU_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      throw ::emu::HipError(std::string(#call) + ": " + hipGetErrorString(e_));     \
  } while (0)

#define EMU_LAUNCH_CHECK() EMU_HIP(hipGetLastError())

typedef double2 zd;

__host__ __device__ __forceinline__ zd zmk(double r, double i) { zd z; z.x = r; z.y = i; return z; }
__host__ __device__ __forceinline__ zd zadd(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ zd zsub(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ zd zmul(zd a, zd b) {
  return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ zd zconj(zd a) { return zmk(a.x, -a.y); }
__host__ __device__ __forceinline__ zd zscale(zd a, double s) { return zmk(a.x * s, a.y * s); }
// c + a*b
__host__ __device__ __forceinline__ zd zfma(zd a, zd b, zd c) {
  return zmk(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
// c - a*b
__host__ __device__ __forceinline__ zd zfms(zd a, zd b, zd c) {
  return zmk(fma(-a.x, b.x, fma(a.y, b.y, c.x)), fma(-a.x, b.y, fma(-a.y, b.x, c.y)));
}
// conj(a)*b + c
__host__ __device__ __forceinline__ zd zcfma(zd a, zd b, zd c) {
  return zmk(fma(a.x, b.x, fma(a.y, b.y, c.x)), fma(a.x, b.y, fma(-a.y, b.x, c.y)));
}
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
  // Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, d = b.x + b.y * r;
    return zmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    double r = b.x / b.y, d = b.y + b.x * r;
    return zmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
__host__ __device__ __forceinline__ double zabs1(zd a) { return fabs(a.x) + fabs(a.y); }

}  // namespace emu
