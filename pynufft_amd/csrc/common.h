// Shared device/host helpers for the MI355X (gfx950) NUFFT kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PNH_MAXD 3

// Complex arithmetic on float2 / double2.  Products use the textbook formula
// (ar*br - ai*bi, ar*bi + ai*br) with contraction disabled at the call sites
// that must reproduce scipy's complex128 CSR arithmetic bit-for-bit.
template <typename T> struct cplx;
template <> struct cplx<float> { typedef float2 t; };
template <> struct cplx<double> { typedef double2 t; };

template <typename T>
__device__ __forceinline__ typename cplx<T>::t mk(T re, T im) {
  typename cplx<T>::t z; z.x = re; z.y = im; return z;
}

#pragma clang fp contract(off)
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C z; z.x = a.x * b.x - a.y * b.y; z.y = a.x * b.y + a.y * b.x; return z;
}
// conj(a) * b
template <typename C>
__device__ __forceinline__ C cmulc(C a, C b) {
  C z; z.x = a.x * b.x + a.y * b.y; z.y = a.x * b.y - a.y * b.x; return z;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) { C z; z.x = a.x + b.x; z.y = a.y + b.y; return z; }
template <typename C>
__device__ __forceinline__ C csub(C a, C b) { C z; z.x = a.x - b.x; z.y = a.y - b.y; return z; }
template <typename C, typename T>
__device__ __forceinline__ C cscale(C a, T s) { C z; z.x = a.x * s; z.y = a.y * s; return z; }
template <typename C>
__device__ __forceinline__ C cconj(C a) { a.y = -a.y; return a; }
#pragma clang fp contract(on)

// Error reporting shared by every entry point (thread-local last error).
void pnh_set_error(const char* fmt, ...);

#define PNH_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      pnh_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),    \
                    __FILE__, __LINE__);                                      \
      return -3;                                                              \
    }                                                                         \
  } while (0)

#define PNH_LAUNCH_CHECK()                                                    \
  do {                                                                        \
    hipError_t _e = hipGetLastError();                                        \
    if (_e != hipSuccess) {                                                   \
      pnh_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(_e),\
                    __FILE__, __LINE__);                                      \
      return -3;                                                              \
    }                                                                         \
  } while (0)
