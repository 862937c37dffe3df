/*
 * Host stand-in for <hip/hip_runtime.h> -- TEST INFRASTRUCTURE ONLY.
 *
 * Lets the solver's own device headers (jetshow_amd/csrc) compile with g++ so
 * that one lane of the state machine can be stepped on the CPU, pixel by
 * pixel, and compared with the oracle.  Kernels are never launched here
 * (hipLaunchKernelGGL is a no-op); device memory is host memory; wavefront
 * intrinsics act on a single lane.  The hardware approximations used by the
 * mixed path (v_log_f32, v_exp_f32, v_rcp_f32, v_rsq_f32) are replaced by the
 * correctly rounded C library functions, so results match the GPU to within
 * those instructions' ~1 ulp, not bit for bit.
 */
#ifndef JSX_HOSTBUILD_HIP_RUNTIME_H
#define JSX_HOSTBUILD_HIP_RUNTIME_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define __host__
#define __device__
#define __global__ static inline
#define __shared__ static
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__ __restrict

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
static const dim3 threadIdx(0, 0, 0), blockIdx(0, 0, 0), blockDim(1, 1, 1), gridDim(1, 1, 1);

static inline float jsx_hb_logf(float x) { return log2f(x); }
static inline float jsx_hb_exp2f(float x) { return exp2f(x); }
#define __builtin_amdgcn_logf(x) jsx_hb_logf(x)
#define __builtin_amdgcn_exp2f(x) jsx_hb_exp2f(x)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_rsqf(x) (1.0f / sqrtf(x))
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __builtin_amdgcn_s_sleep(n) ((void)0)
#define __builtin_amdgcn_s_waitcnt(n) ((void)0)

static inline int __lane_id() { return 0; }
static inline unsigned long long __ballot(int p) { return p ? 1ull : 0ull; }
static inline int __any(int p) { return p != 0; }
static inline int __all(int p) { return p != 0; }
static inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
static inline int __ffsll(long long v) { return __builtin_ffsll(v); }
template <class T> static inline T __shfl(T v, int, int = 64) { return v; }
template <class T> static inline T __shfl_up(T v, unsigned, int = 64) { return v; }
template <class T> static inline T __shfl_xor(T, int, int = 64) { return T(0); }
static inline void __syncthreads() {}
template <class T, class U> static inline T atomicAdd(T *p, U v) { T o = *p; *p = o + (T)v; return o; }
template <class T, class U> static inline T atomicSub(T *p, U v) { T o = *p; *p = o - (T)v; return o; }
template <class T, class U> static inline T atomicOr(T *p, U v) { T o = *p; *p = o | (T)v; return o; }
template <class T, class U> static inline T atomicCAS(T *p, U cmp, U v) { T o = *p; if (o == (T)cmp) *p = (T)v; return o; }
#define __HIP_MEMORY_SCOPE_AGENT 3
template <class T, class U> static inline void __hip_atomic_store(T *p, U v, int, int) { *p = (T)v; }
template <class T> static inline T __hip_atomic_load(T *p, int, int) { return *p; }

typedef int hipError_t;
#define hipSuccess 0
#define hipErrorNoDevice 100
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
#define hipEventDisableTiming 2
#define hipStreamCaptureModeThreadLocal 1
struct hipDeviceProp_t { int multiProcessorCount; };
struct double2 { double x, y; };
static inline double2 make_double2(double a, double b) { return double2{a, b}; }

static inline hipError_t hipMalloc(void *p, size_t n) { *(void **)p = calloc(1, n ? n : 1); return hipSuccess; }
static inline hipError_t hipMemsetAsync(void *p, int v, size_t n, hipStream_t) { memset(p, v, n); return hipSuccess; }
/* file-scope device arrays: ordinary host arrays here, their own address */
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
static inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = (void *)sym; return hipSuccess; }
static inline hipError_t hipHostMalloc(void *p, size_t n, int) { return hipMalloc(p, n); }
static inline hipError_t hipFree(void *p) { free(p); return hipSuccess; }
static inline hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
static inline hipError_t hipMemcpy(void *d, const void *s, size_t n, int) { memmove(d, s, n); return hipSuccess; }
static inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int, hipStream_t) { memmove(d, s, n); return hipSuccess; }
static inline hipError_t hipSetDevice(int) { return hipSuccess; }
static inline hipError_t hipGetDeviceCount(int *n) { *n = 1; return hipSuccess; }
static inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) { p->multiProcessorCount = 1; return hipSuccess; }
static inline hipError_t hipMemGetInfo(size_t *f, size_t *t) { *f = *t = (size_t)1 << 34; return hipSuccess; }
static inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, int) { *s = nullptr; return hipSuccess; }
static inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
static inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
static inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, int) { *e = nullptr; return hipSuccess; }
static inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
static inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
static inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
/* what hipEventQuery reports (set by the host tests to force error paths) */
#define hipErrorNotReady 600
#define hipErrorUnknown 999
static int jsx_hb_event_status = 0;
static inline hipError_t hipEventQuery(hipEvent_t) { return jsx_hb_event_status; }
static inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
static inline hipError_t hipEventCreate(hipEvent_t *e) { *e = nullptr; return hipSuccess; }
static inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
static inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
static inline hipError_t hipGetLastError() { return hipSuccess; }
static inline const char *hipGetErrorString(hipError_t) { return "host build"; }
typedef enum { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 } hipStreamCaptureStatus;
/* no graphs here: the capture query fails, so the library launches eagerly */
static inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *) { return hipErrorNoDevice; }
static inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorNoDevice; }
static inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNoDevice; }
static inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, void *, size_t) { return hipErrorNoDevice; }
static inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
static inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
static inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNoDevice; }

/* kernels are not run by the host build */
#define hipLaunchKernelGGL(...) ((void)0)

#endif
