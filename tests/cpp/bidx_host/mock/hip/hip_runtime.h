// Host-only stand-in of the HIP runtime for tests/cpp/bidx_host: device
// memory is host memory, copies are memcpy, streams are synchronous and a
// kernel launch calls the CPU emulation of that kernel.  Test code only.
#pragma once
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdexcept>
typedef int hipError_t;
enum { hipSuccess = 0 };
typedef void *hipStream_t;
typedef void *hipEvent_t;
struct int2 { int x, y; };
struct int4 { int x, y, z, w; };
inline int2 make_int2(int a, int b) { return int2{a, b}; }
struct dim3 { unsigned x, y, z; dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {} };
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
enum { hipHostMallocDefault = 0 };
inline hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n ? n : 1); return hipSuccess; }
inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipHostMalloc(void **p, size_t n, int) { *p = malloc(n ? n : 1); return hipSuccess; }
template <class T> inline hipError_t hipHostMalloc(T **p, size_t n, int f) { return hipHostMalloc((void **)p, n, f); }
inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) { memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline const char *hipGetErrorString(hipError_t) { return "mock"; }
#define HIP_KERNEL_NAME(...) __VA_ARGS__
#define hipLaunchKernelGGL(k, g, b, sh, s, ...) k(dim3(g), dim3(b), __VA_ARGS__)
