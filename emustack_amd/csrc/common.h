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

// Load of a value that is read once (a streamed matrix): the nontemporal hint
// makes its cache line the first to go, so that it does not push reused lines
// (gathered vector entries) out of L2.
__device__ __forceinline__ zd ld_stream(const zd *p) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_nontemporal_load(reinterpret_cast<const d2 *>(p));
  return zmk(v.x, v.y);
}
__device__ __forceinline__ int32_t ld_stream(const int32_t *p) { return __builtin_nontemporal_load(p); }

// Raw buffer view of `bytes` bytes at p.  Its loads take the base from SGPRs,
// a 32-bit per-lane byte offset and a wave-uniform byte offset (SGPR): a
// batch of loads that differ only in the uniform offset shares one VGPR of
// address, where 64-bit pointer loads would hold a VGPR pair each.  Reads
// at or beyond `bytes` return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_view(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ zd buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
  return zmk(v.x, v.y);
}
// the same for data read once (cache policy nt, as ld_stream)
__device__ __forceinline__ zd buf_ld_stream(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const d2 v = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 2));
  return zmk(v.x, v.y);
}
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
