// runtime.hip -- allocation, copies, streams and error reporting of the
// C ABI (the HIP-native replacement of GPUtils' Session plumbing,
// include/tensor.cuh:133-247 and the cudaMalloc/cudaMemcpy call sites).
#include <cstdarg>
#include <cstring>

#include "common.hpp"

namespace gputils {

static thread_local char g_last_error[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what) {
    set_error("%s failed: %s (%d)", what, hipGetErrorString(e), (int)e);
    return (int)e > 0 ? (int)e : 1;
}

// Fill `bytes` bytes with `value`: 16-byte vector stores for the aligned
// body, single bytes for the ragged head and tail.
__global__ void fill_bytes_kernel(unsigned char *p, unsigned char value, size_t bytes) {
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const size_t nthreads = (size_t)gridDim.x * blockDim.x;
    const uintptr_t addr = reinterpret_cast<uintptr_t>(p);
    size_t head = (16 - (addr & 15)) & 15;
    if (head > bytes) head = bytes;
    const size_t body = (bytes - head) / 16;
    const unsigned v = value * 0x01010101u;
    const uint4 v4 = make_uint4(v, v, v, v);
    uint4 *b = reinterpret_cast<uint4 *>(p + head);
    for (size_t i = tid; i < body; i += nthreads) b[i] = v4;
    const size_t tail_start = head + body * 16;
    if (tid < head) p[tid] = value;
    for (size_t i = tail_start + tid; i < bytes; i += nthreads) p[i] = value;
}

__global__ void fill_pointers_kernel(void **dst, const char *base, int64_t stride_bytes, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (void *)(base + i * stride_bytes);
}

static hipMemcpyKind to_kind(int kind) {
    switch (kind) {
        case 0: return hipMemcpyHostToDevice;
        case 1: return hipMemcpyDeviceToHost;
        case 2: return hipMemcpyDeviceToDevice;
        default: return hipMemcpyDefault;
    }
}

}  // namespace gputils

using namespace gputils;

extern "C" {

const char *gputils_last_error(void) { return g_last_error; }

const char *gputils_version(void) { return "gputils-mi355x 0.1.0 (gfx950)"; }

int gputils_device_count(int *count) {
    GPUTILS_ARG(count != nullptr, 1, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) {
        (void)hipGetLastError();
        *count = 0;
        return GPUTILS_OK;
    }
    GPUTILS_HIP(e);
    *count = c;
    return GPUTILS_OK;
}

int gputils_malloc(void **ptr, size_t bytes) {
    GPUTILS_ARG(ptr != nullptr, 1, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return GPUTILS_OK;
    GPUTILS_HIP(hipMalloc(ptr, bytes));
    return GPUTILS_OK;
}

int gputils_free(void *ptr) {
    if (ptr == nullptr) return GPUTILS_OK;
    GPUTILS_HIP(hipFree(ptr));
    return GPUTILS_OK;
}

int gputils_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    if (bytes == 0) return GPUTILS_OK;
    GPUTILS_ARG(dst != nullptr, 1, "dst is NULL");
    GPUTILS_ARG(src != nullptr, 2, "src is NULL");
    GPUTILS_ARG(kind >= 0 && kind <= 3, 4, "kind must be 0..3");
    GPUTILS_HIP(hipMemcpy(dst, src, bytes, to_kind(kind)));
    return GPUTILS_OK;
}

int gputils_memcpy_async(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    if (bytes == 0) return GPUTILS_OK;
    GPUTILS_ARG(dst != nullptr, 1, "dst is NULL");
    GPUTILS_ARG(src != nullptr, 2, "src is NULL");
    GPUTILS_ARG(kind >= 0 && kind <= 3, 4, "kind must be 0..3");
    GPUTILS_HIP(hipMemcpyAsync(dst, src, bytes, to_kind(kind), as_stream(stream)));
    return GPUTILS_OK;
}

int gputils_memset(void *ptr, int value, size_t bytes) {
    if (bytes == 0) return GPUTILS_OK;
    GPUTILS_ARG(ptr != nullptr, 1, "ptr is NULL");
    GPUTILS_HIP(hipMemset(ptr, value, bytes));
    return GPUTILS_OK;
}

int gputils_fill_bytes_async(void *ptr, unsigned char value, size_t bytes, void *stream) {
    if (bytes == 0) return GPUTILS_OK;
    GPUTILS_ARG(ptr != nullptr, 1, "ptr is NULL");
    const unsigned grid = grid_1d((int64_t)(bytes / 16 + 1), 256, 4096);
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(grid), dim3(256), 0, as_stream(stream),
                       (unsigned char *)ptr, value, bytes);
    return launch_check("fill_bytes_kernel");
}

int gputils_stream_create(void **stream) {
    GPUTILS_ARG(stream != nullptr, 1, "stream is NULL");
    hipStream_t s;
    GPUTILS_HIP(hipStreamCreate(&s));
    *stream = (void *)s;
    return GPUTILS_OK;
}

int gputils_stream_destroy(void *stream) {
    if (stream == nullptr) return GPUTILS_OK;
    GPUTILS_HIP(hipStreamDestroy(as_stream(stream)));
    return GPUTILS_OK;
}

int gputils_stream_synchronize(void *stream) {
    GPUTILS_HIP(hipStreamSynchronize(as_stream(stream)));
    return GPUTILS_OK;
}

int gputils_device_synchronize(void) {
    GPUTILS_HIP(hipDeviceSynchronize());
    return GPUTILS_OK;
}

int gputils_fill_pointer_array(void **dst, const void *base, int64_t stride_bytes, int64_t count,
                               void *stream) {
    GPUTILS_ARG(count >= 0, 4, "count < 0");
    if (count == 0) return GPUTILS_OK;
    GPUTILS_ARG(dst != nullptr, 1, "dst is NULL");
    hipLaunchKernelGGL(fill_pointers_kernel, dim3(grid_1d(count, 256, 1024)), dim3(256), 0,
                       as_stream(stream), dst, (const char *)base, stride_bytes, count);
    return launch_check("fill_pointers_kernel");
}

struct gputils_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int gputils_graph_capture_begin(void *stream) {
    GPUTILS_ARG(stream != nullptr, 1, "the null stream cannot be captured");
    GPUTILS_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
    return GPUTILS_OK;
}

int gputils_graph_capture_end(void *stream, gputils_graph_t *graph) {
    GPUTILS_ARG(stream != nullptr, 1, "the null stream cannot be captured");
    GPUTILS_ARG(graph != nullptr, 2, "graph is NULL");
    hipGraph_t g = nullptr;
    GPUTILS_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    if (rc != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(rc, "hipGraphInstantiate");
    }
    *graph = new gputils_graph{g, e};
    return GPUTILS_OK;
}

int gputils_graph_launch(gputils_graph_t graph, void *stream) {
    GPUTILS_ARG(graph != nullptr, 1, "graph is NULL");
    GPUTILS_HIP(hipGraphLaunch(graph->exec, as_stream(stream)));
    return GPUTILS_OK;
}

int gputils_graph_destroy(gputils_graph_t graph) {
    if (graph == nullptr) return GPUTILS_OK;
    const hipError_t a = hipGraphExecDestroy(graph->exec);
    const hipError_t b = hipGraphDestroy(graph->graph);
    delete graph;
    if (a != hipSuccess) return hip_fail(a, "hipGraphExecDestroy");
    if (b != hipSuccess) return hip_fail(b, "hipGraphDestroy");
    return GPUTILS_OK;
}

}  // extern "C"
