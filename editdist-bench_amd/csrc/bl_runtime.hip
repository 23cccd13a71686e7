/*
 * bl_runtime.hip -- device, memory, stream and event entry points of the C ABI
 * (declared in include/bitlev_hip.h). The reference leaves all of this to
 * CuPy (bitlev/gpu.py:44-51, :81, :87-88, :137); here the library owns it, so
 * callers need nothing but this shared object and a HIP driver.
 */
#include <hip/hip_runtime.h>
#include <string.h>

#include "bl_internal.h"

namespace bl {

static thread_local char g_err[512];

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace bl

extern "C" {

int bl_version(void) { return 1; }

const char *bl_last_error(void) { return bl::g_err; }

int bl_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int bl_set_device(int device)
{
    BL_HIP(hipSetDevice(device));
    return BL_OK;
}

int bl_device_info(char *name, size_t name_len, int *cu_count, int *arch_major, int *arch_minor)
{
    int dev = 0;
    BL_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    BL_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (arch_major) *arch_major = p.major;
    if (arch_minor) *arch_minor = p.minor;
    return BL_OK;
}

int bl_malloc(void **d_ptr, size_t bytes)
{
    if (!d_ptr) return bl::fail(BL_EINVAL, "bl_malloc: null output pointer");
    *d_ptr = nullptr;
    if (bytes == 0) return BL_OK;
    BL_HIP(hipMalloc(d_ptr, bytes));
    return BL_OK;
}

int bl_free(void *d_ptr)
{
    if (d_ptr) BL_HIP(hipFree(d_ptr));
    return BL_OK;
}

int bl_host_malloc(void **h_ptr, size_t bytes)
{
    if (!h_ptr) return bl::fail(BL_EINVAL, "bl_host_malloc: null output pointer");
    *h_ptr = nullptr;
    if (bytes == 0) return BL_OK;
    BL_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault));
    return BL_OK;
}

int bl_host_free(void *h_ptr)
{
    if (h_ptr) BL_HIP(hipHostFree(h_ptr));
    return BL_OK;
}

int bl_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream)
{
    if (!bytes) return BL_OK;
    BL_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return BL_OK;
}

int bl_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream)
{
    if (!bytes) return BL_OK;
    BL_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return BL_OK;
}

int bl_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
    if (!bytes) return BL_OK;
    BL_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return BL_OK;
}

int bl_memset(void *d_dst, int value, size_t bytes, void *stream)
{
    if (!bytes) return BL_OK;
    BL_HIP(hipMemsetAsync(d_dst, value, bytes, (hipStream_t)stream));
    return BL_OK;
}

int bl_stream_create(void **stream)
{
    if (!stream) return bl::fail(BL_EINVAL, "bl_stream_create: null output pointer");
    hipStream_t s;
    BL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return BL_OK;
}

int bl_stream_destroy(void *stream)
{
    if (stream) BL_HIP(hipStreamDestroy((hipStream_t)stream));
    return BL_OK;
}

int bl_stream_sync(void *stream)
{
    BL_HIP(hipStreamSynchronize((hipStream_t)stream));
    return BL_OK;
}

int bl_device_sync(void)
{
    BL_HIP(hipDeviceSynchronize());
    return BL_OK;
}

int bl_event_create(void **event)
{
    if (!event) return bl::fail(BL_EINVAL, "bl_event_create: null output pointer");
    hipEvent_t e;
    BL_HIP(hipEventCreate(&e));
    *event = (void *)e;
    return BL_OK;
}

int bl_event_destroy(void *event)
{
    if (event) BL_HIP(hipEventDestroy((hipEvent_t)event));
    return BL_OK;
}

int bl_event_record(void *event, void *stream)
{
    BL_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return BL_OK;
}

int bl_stream_wait_event(void *stream, void *event)
{
    if (!event) return bl::fail(BL_EINVAL, "bl_stream_wait_event: null event");
    BL_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return BL_OK;
}

int bl_event_sync(void *event)
{
    BL_HIP(hipEventSynchronize((hipEvent_t)event));
    return BL_OK;
}

int bl_event_query(void *event)
{
    hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return bl::fail(BL_EHIP, "hipEventQuery: %s", hipGetErrorString(e));
}

int bl_graph_begin(void *stream)
{
    if (!stream) return bl::fail(BL_EINVAL, "bl_graph_begin: capture needs a non-default stream");
    BL_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return BL_OK;
}

int bl_graph_end(void *stream, void **graph_exec)
{
    if (!graph_exec) return bl::fail(BL_EINVAL, "bl_graph_end: null output pointer");
    hipGraph_t g = nullptr;
    BL_HIP(hipStreamEndCapture((hipStream_t)stream, &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return bl::fail(BL_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    *graph_exec = (void *)ge;
    return BL_OK;
}

int bl_graph_launch(void *graph_exec, void *stream)
{
    BL_HIP(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return BL_OK;
}

int bl_graph_destroy(void *graph_exec)
{
    if (graph_exec) BL_HIP(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return BL_OK;
}

int bl_event_elapsed_ms(void *start, void *stop, float *ms)
{
    BL_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return BL_OK;
}

}  /* extern "C" */
