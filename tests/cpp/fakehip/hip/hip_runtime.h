// A host-only stand-in for the HIP runtime API, enough of it to compile the
// serving loop's host code (tpr_amd/csrc/server.hip) with g++ and run it on
// the CPU under ThreadSanitizer.  "Device" memory is host memory; each stream
// is a thread that runs its queued operations in order; events complete when
// their stream reaches them.  Kernel launchers are replaced by CPU stubs in
// server_stress.cpp.  Nothing here is used by the library itself.
#pragma once
#include <stddef.h>
#include <stdint.h>

#define __host__
#define __device__
#define __global__
#define __forceinline__ inline

typedef enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNotReady = 600,
    hipErrorUnknown = 999
} hipError_t;

typedef enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4
} hipMemcpyKind;

typedef enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount = 1 } hipDeviceAttribute_t;
typedef enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize = 8 } hipFuncAttribute;

struct FakeStream;
struct FakeEvent;
typedef FakeStream* hipStream_t;
typedef FakeEvent* hipEvent_t;

#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u

hipError_t hipGetDevice(int* dev);
hipError_t hipSetDevice(int dev);
hipError_t hipDeviceGetAttribute(int* value, hipDeviceAttribute_t attr, int dev);
const char* hipGetErrorString(hipError_t e);
hipError_t hipFuncSetAttribute(const void* fn, hipFuncAttribute attr, int value);

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void* p);

hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);

hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);

hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
