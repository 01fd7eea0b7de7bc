// Host-only stand-in for the HIP runtime, used to compile the engine's host
// code (bs_api.cpp) for the CPU-side buffer checker.  Device memory is plain
// host memory; every allocation is recorded so that launches can be checked
// against the extents their kernels touch.  Test infrastructure only.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3 };
typedef struct ihipStream_t *hipStream_t;
typedef struct ihipEvent_t *hipEvent_t;
#define hipStreamNonBlocking 1

struct float4
{
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
struct double4
{
    double x, y, z, w;
};
inline double4 make_double4(double x, double y, double z, double w) { return double4{x, y, z, w}; }

struct hipDeviceProp_t
{
    char gcnArchName[256];
};

// allocation tracking (hostcheck_runtime.cpp)
bool hc_valid(const void *p, size_t bytes);
void hc_fail(const char *what);

hipError_t hipMalloc(void **p, size_t bytes);
hipError_t hipFree(void *p);
#define hipHostMallocDefault 0u
hipError_t hipHostMalloc(void **p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipMemGetInfo(size_t *free, size_t *total);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
const char *hipGetErrorString(hipError_t e);
// Device-side file-scope arrays are host arrays here; HIP_SYMBOL registers
// one as a live device allocation so launches into it pass the extent checks.
#define __device__
#define __align__(x) __attribute__((aligned(x)))
const void *hc_symbol(const void *p, size_t bytes);
#define HIP_SYMBOL(x) hc_symbol(&(x), sizeof(x))
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
// Graphs: capture is not emulated (the launchers here run at once); the
// calls report "not supported", which the engine turns into an error.
typedef struct ihipGraph_t *hipGraph_t;
typedef struct ihipGraphExec_t *hipGraphExec_t;
typedef struct ihipGraphNode_t *hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
enum { hipErrorNotSupported = 801 };
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
