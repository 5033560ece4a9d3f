// runtime.hip - error state, device/memory/stream plumbing of the C ABI.
// Replaces Utils::* (include/utils/utils.hpp:13-95) and the error reporting of
// ErrorChecker (include/utils/exceptions.hpp:13-74), which synchronised the
// device after every call; here errors are returned as status codes.
#include <cstdio>
#include <cstring>

#include "common.hpp"

namespace psx {

static thread_local char g_err[1024] = "";

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int ok() {
  g_err[0] = 0;
  return PSX_OK;
}

}  // namespace psx

using psx::as_stream;

extern "C" const char* psx_last_error_string(void) { return psx::g_err; }
extern "C" int psx_version(void) { return 10000; }

extern "C" int psx_device_count(int* count) {
  PSX_REQUIRE(count, "psx_device_count: null count");
  PSX_CHECK_HIP(hipGetDeviceCount(count));
  return psx::ok();
}

extern "C" int psx_set_device(int device) {
  PSX_CHECK_HIP(hipSetDevice(device));
  return psx::ok();
}

extern "C" int psx_get_device(int* device) {
  PSX_REQUIRE(device, "psx_get_device: null device");
  PSX_CHECK_HIP(hipGetDevice(device));
  return psx::ok();
}

extern "C" int psx_device_info(int device, char* name, size_t cap, uint64_t* hbm_bytes, int* cu_count,
                               int* arch_gfx) {
  hipDeviceProp_t p;
  PSX_CHECK_HIP(hipGetDeviceProperties(&p, device));
  if (name && cap) {
    strncpy(name, p.name, cap - 1);
    name[cap - 1] = 0;
  }
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (arch_gfx) {
    int v = 0;
    sscanf(p.gcnArchName, "gfx%x", &v);
    *arch_gfx = v;
  }
  return psx::ok();
}

extern "C" int psx_malloc(void** ptr, uint64_t bytes) {
  PSX_REQUIRE(ptr, "psx_malloc: null ptr");
  *ptr = nullptr;
  if (bytes == 0) return psx::ok();
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return psx::fail(PSX_ERR_ALLOC, "hipMalloc(%llu) failed: %s", (unsigned long long)bytes, hipGetErrorString(e));
  }
  return psx::ok();
}

extern "C" int psx_free(void* ptr) {
  if (ptr) PSX_CHECK_HIP(hipFree(ptr));
  return psx::ok();
}

extern "C" int psx_host_malloc(void** ptr, uint64_t bytes) {
  PSX_REQUIRE(ptr, "psx_host_malloc: null ptr");
  *ptr = nullptr;
  if (bytes == 0) return psx::ok();
  hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *ptr = nullptr;
    (void)hipGetLastError();
    return psx::fail(PSX_ERR_ALLOC, "hipHostMalloc(%llu) failed: %s", (unsigned long long)bytes,
                     hipGetErrorString(e));
  }
  return psx::ok();
}

extern "C" int psx_host_free(void* ptr) {
  if (ptr) PSX_CHECK_HIP(hipHostFree(ptr));
  return psx::ok();
}

extern "C" int psx_memcpy_h2d(void* dst, const void* src, uint64_t bytes, psx_stream s) {
  if (!bytes) return psx::ok();
  PSX_REQUIRE(dst && src, "psx_memcpy_h2d: null pointer");
  PSX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(s)));
  PSX_CHECK_HIP(hipStreamSynchronize(as_stream(s)));
  return psx::ok();
}

extern "C" int psx_memcpy_d2h(void* dst, const void* src, uint64_t bytes, psx_stream s) {
  if (!bytes) return psx::ok();
  PSX_REQUIRE(dst && src, "psx_memcpy_d2h: null pointer");
  PSX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(s)));
  PSX_CHECK_HIP(hipStreamSynchronize(as_stream(s)));
  return psx::ok();
}

extern "C" int psx_memcpy_d2d(void* dst, const void* src, uint64_t bytes, psx_stream s) {
  if (!bytes) return psx::ok();
  PSX_REQUIRE(dst && src, "psx_memcpy_d2d: null pointer");
  PSX_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(s)));
  return psx::ok();
}

extern "C" int psx_stream_create(psx_stream* s) {
  PSX_REQUIRE(s, "psx_stream_create: null stream");
  hipStream_t h;
  PSX_CHECK_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = (psx_stream)h;
  return psx::ok();
}

extern "C" int psx_stream_destroy(psx_stream s) {
  if (s) PSX_CHECK_HIP(hipStreamDestroy(as_stream(s)));
  return psx::ok();
}

extern "C" int psx_stream_synchronize(psx_stream s) {
  PSX_CHECK_HIP(hipStreamSynchronize(as_stream(s)));
  return psx::ok();
}

extern "C" int psx_device_synchronize(void) {
  PSX_CHECK_HIP(hipDeviceSynchronize());
  return psx::ok();
}

extern "C" int psx_event_create(psx_event* e) {
  PSX_REQUIRE(e, "psx_event_create: null event");
  hipEvent_t h;
  PSX_CHECK_HIP(hipEventCreate(&h));
  *e = (psx_event)h;
  return psx::ok();
}

extern "C" int psx_event_destroy(psx_event e) {
  if (e) PSX_CHECK_HIP(hipEventDestroy((hipEvent_t)e));
  return psx::ok();
}

extern "C" int psx_event_record(psx_event e, psx_stream s) {
  PSX_CHECK_HIP(hipEventRecord((hipEvent_t)e, as_stream(s)));
  return psx::ok();
}

extern "C" int psx_event_synchronize(psx_event e) {
  PSX_CHECK_HIP(hipEventSynchronize((hipEvent_t)e));
  return psx::ok();
}

extern "C" int psx_event_elapsed_ms(psx_event start, psx_event stop, float* ms) {
  PSX_REQUIRE(ms, "psx_event_elapsed_ms: null ms");
  PSX_CHECK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return psx::ok();
}

extern "C" int psx_mem_info(uint64_t* free_bytes, uint64_t* total_bytes) {
  size_t f = 0, t = 0;
  PSX_CHECK_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return psx::ok();
}
