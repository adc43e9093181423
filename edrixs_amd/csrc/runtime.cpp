// Error state, device-memory accounting and device selection for libedx.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "common.h"

namespace edx {

static thread_local std::string g_last_error;
static thread_local int g_device = 0;
static thread_local int g_corehole_layout = EDX_COREHOLE_REFERENCE;
thread_local hipStream_t g_launch_stream = 0;

int corehole_layout() { return g_corehole_layout; }

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    return set_error(EDX_ERR_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
                     line);
}

static std::mutex g_mem_mu;
static std::unordered_map<void *, size_t> g_allocs;
static int64_t g_in_use = 0;

void *dmalloc(size_t bytes, const char *what) {
    void *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error(EDX_ERR_NOMEM, "device allocation of %zu bytes for %s failed: %s", bytes, what,
                  hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mem_mu);
    g_allocs[p] = bytes;
    g_in_use += (int64_t)bytes;
    return p;
}

void dfree(void *p) {
    if (!p) return;
    {
        std::lock_guard<std::mutex> lk(g_mem_mu);
        auto it = g_allocs.find(p);
        if (it != g_allocs.end()) {
            g_in_use -= (int64_t)it->second;
            g_allocs.erase(it);
        }
    }
    (void)hipFree(p);
}

int64_t mem_in_use() {
    std::lock_guard<std::mutex> lk(g_mem_mu);
    return g_in_use;
}

int current_device() { return g_device; }

int num_cus() {
    static int cached = -1;
    if (cached < 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, g_device) == hipSuccess)
            cached = prop.multiProcessorCount;
        else
            cached = 256;
    }
    return cached;
}

}  // namespace edx

extern "C" {

const char *edx_last_error(void) { return edx::g_last_error.c_str(); }

int edx_set_corehole_layout(int layout, int *previous) {
    EDX_CHECK(layout == EDX_COREHOLE_REFERENCE || layout == EDX_COREHOLE_INTERLEAVED, EDX_ERR_ARG,
              "unknown core-hole layout %d", layout);
    if (previous) *previous = edx::g_corehole_layout;
    edx::g_corehole_layout = layout;
    return EDX_OK;
}

int edx_init(int device) {
    int n = 0;
    EDX_HIP(hipGetDeviceCount(&n));
    EDX_CHECK(device >= 0 && device < n, EDX_ERR_ARG, "device %d out of range (%d devices)",
              device, n);
    EDX_HIP(hipSetDevice(device));
    edx::g_device = device;
    return edx::ensure_binomials();
}

int edx_mem_in_use(int64_t *bytes) {
    EDX_CHECK(bytes, EDX_ERR_ARG, "null output");
    *bytes = edx::mem_in_use();
    return EDX_OK;
}

int edx_set_stream(void *stream) {
    edx::g_launch_stream = (hipStream_t)stream;
    return EDX_OK;
}

}  // extern "C"
