// TEST INFRASTRUCTURE: a host-memory stand-in for the few HIP runtime calls libtwins_core's host
// side makes, so that the unchanged twins_core.cpp / twins_cli.cpp can be linked into a CPU
// program (with AddressSanitizer) that runs the device functions on the CPU for bounds checks.
// Never linked into the product.
#include <cstdlib>
#include <cstring>
#include <hip/hip_runtime_api.h>

extern "C" {
hipError_t hipMalloc(void **p, size_t n)
{
    *p = calloc(1, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p)
{
    free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind)
{
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n)
{
    memset(d, v, n);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t)
{
    memset(d, v, n);
    return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n)
{
    *n = 1;
    return hipSuccess;
}
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int)
{
    *v = 256;
    return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t *s)
{
    *s = (hipStream_t)malloc(8);
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s)
{
    free(s);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e)
{
    *e = (hipEvent_t)malloc(8);
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e)
{
    free(e);
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t)
{
    *ms = 0;
    return hipSuccess;
}
hipError_t hipGetLastError(void) { return hipSuccess; }
// no graphs on the CPU: the capture is refused and the batch launches directly
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g)
{
    *g = nullptr;
    return hipErrorNotSupported;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, hipGraphNode_t *, char *, size_t) { return hipErrorNotSupported; }
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
// the serving loop needs a GPU: its calls are refused here
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)malloc(8); return hipSuccess; }
hipError_t hipStreamQuery(hipStream_t) { return hipSuccess; }
hipError_t hipHostMalloc(void **, size_t, unsigned) { return hipErrorNotSupported; }
hipError_t hipHostFree(void *) { return hipSuccess; }
hipError_t hipHostGetDevicePointer(void **, void *, unsigned) { return hipErrorNotSupported; }
hipError_t hipMemGetInfo(size_t *f, size_t *t) { *f = *t = 0; return hipSuccess; }
hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "host stand-in"; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned)
{
    *e = (hipEvent_t)malloc(8);
    return hipSuccess;
}
// (launches run to their end on the CPU: an event is always complete)
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
}
