// Runtime part of the C ABI: errors, device memory, streams, events.
#include <cstdio>
#include <cstring>

#include "fvm_internal.hpp"

namespace fvm {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? FVM_ERR_ALLOC : FVM_ERR_HIP;
}
const std::string& last_error() { return g_last_error; }
}  // namespace fvm

extern "C" {

int fvm_last_error(char* buf, size_t n) {
  if (!buf || n == 0) return FVM_ERR_ARG;
  const std::string& s = fvm::last_error();
  size_t k = s.size() < n - 1 ? s.size() : n - 1;
  std::memcpy(buf, s.data(), k);
  buf[k] = 0;
  return FVM_OK;
}

int fvm_build_info(char* buf, size_t n) {
  FVM_CHECK(buf && n > 0, "fvm_build_info: null buffer");
  std::snprintf(buf, n, "gfx950 rhs_rev=%d geo_level=%d solver_threads=%d spmv_rows=%d", FVM_RHS_REV,
                FVM_GEO_LEVEL, FVM_SOLVER_THREADS, FVM_SPMV_ROWS);
  return FVM_OK;
}

int fvm_device_count(int* n) {
  FVM_CHECK(n, "fvm_device_count: null output");
  hipError_t e = hipGetDeviceCount(n);
  if (e != hipSuccess) {
    *n = 0;
    return fvm::hip_fail(e, "hipGetDeviceCount");
  }
  return FVM_OK;
}

int fvm_device_synchronize(void) {
  FVM_HIP(hipDeviceSynchronize());
  return FVM_OK;
}

int fvm_alloc(void** d_ptr, size_t bytes) {
  FVM_CHECK(d_ptr, "fvm_alloc: null output");
  *d_ptr = nullptr;
  if (bytes == 0) bytes = 1;
  FVM_HIP(hipMalloc(d_ptr, bytes));
  return FVM_OK;
}

int fvm_free(void* d_ptr) {
  if (!d_ptr) return FVM_OK;
  FVM_HIP(hipFree(d_ptr));
  return FVM_OK;
}

int fvm_memcpy_h2d(void* d_dst, const void* src, size_t bytes, fvm_stream s) {
  if (bytes == 0) return FVM_OK;
  FVM_CHECK(d_dst && src, "fvm_memcpy_h2d: null pointer");
  FVM_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
  FVM_HIP(hipStreamSynchronize((hipStream_t)s));
  return FVM_OK;
}

int fvm_memcpy_d2h(void* dst, const void* d_src, size_t bytes, fvm_stream s) {
  if (bytes == 0) return FVM_OK;
  FVM_CHECK(dst && d_src, "fvm_memcpy_d2h: null pointer");
  FVM_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
  FVM_HIP(hipStreamSynchronize((hipStream_t)s));
  return FVM_OK;
}

int fvm_memcpy_d2d(void* d_dst, const void* d_src, size_t bytes, fvm_stream s) {
  if (bytes == 0) return FVM_OK;
  FVM_CHECK(d_dst && d_src, "fvm_memcpy_d2d: null pointer");
  FVM_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
  return FVM_OK;
}

int fvm_memset(void* d_dst, int value, size_t bytes, fvm_stream s) {
  if (bytes == 0) return FVM_OK;
  FVM_CHECK(d_dst, "fvm_memset: null pointer");
  FVM_HIP(hipMemsetAsync(d_dst, value, bytes, (hipStream_t)s));
  return FVM_OK;
}

int fvm_stream_create(fvm_stream* s) {
  FVM_CHECK(s, "fvm_stream_create: null output");
  hipStream_t h;
  FVM_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
  *s = h;
  return FVM_OK;
}

int fvm_stream_destroy(fvm_stream s) {
  if (!s) return FVM_OK;
  FVM_HIP(hipStreamDestroy((hipStream_t)s));
  return FVM_OK;
}

int fvm_stream_synchronize(fvm_stream s) {
  FVM_HIP(hipStreamSynchronize((hipStream_t)s));
  return FVM_OK;
}

int fvm_event_create(fvm_event* e) {
  FVM_CHECK(e, "fvm_event_create: null output");
  hipEvent_t h;
  FVM_HIP(hipEventCreate(&h));
  *e = h;
  return FVM_OK;
}

int fvm_event_destroy(fvm_event e) {
  if (!e) return FVM_OK;
  FVM_HIP(hipEventDestroy((hipEvent_t)e));
  return FVM_OK;
}

int fvm_event_record(fvm_event e, fvm_stream s) {
  FVM_CHECK(e, "fvm_event_record: null event");
  FVM_HIP(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
  return FVM_OK;
}

int fvm_event_elapsed_ms(float* ms, fvm_event start, fvm_event stop) {
  FVM_CHECK(ms && start && stop, "fvm_event_elapsed_ms: null argument");
  FVM_HIP(hipEventSynchronize((hipEvent_t)stop));
  FVM_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
  return FVM_OK;
}

}  // extern "C"
