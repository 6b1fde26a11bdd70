// Context, error reporting and plain memory helpers of the C ABI.
#include <cstring>
#include <string>

#include "internal.hpp"

namespace dda {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace dda

using namespace dda;

extern "C" const char* dda_last_error(void) { return g_last_error.c_str(); }

extern "C" dda_status dda_ctx_create(int device, dda_ctx_t* out) {
  if (!out) {
    set_error("dda_ctx_create: null output");
    return DDA_INVALID_ARG;
  }
  *out = nullptr;
  int n = 0;
  DDA_TRY_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("dda_ctx_create: no such device");
    return DDA_INVALID_ARG;
  }
  DDA_TRY_HIP(hipSetDevice(device));
  dda_ctx_s* c = new dda_ctx_s();
  c->device = device;
  hipError_t e = hipGetDeviceProperties(&c->prop, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    set_error(std::string("dda_ctx_create: ") + hipGetErrorString(e));
    delete c;
    return DDA_HIP_ERROR;
  }
  const dda_status rs = noalloc_scratch_setup(device);
  if (rs != DDA_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rs;
  }
  *out = c;
  return DDA_OK;
}

extern "C" dda_status dda_ctx_destroy(dda_ctx_t ctx) {
  if (!ctx) return DDA_OK;
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return DDA_OK;
}

extern "C" dda_status dda_ctx_synchronize(dda_ctx_t ctx) {
  if (!ctx) {
    set_error("dda_ctx_synchronize: null context");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  DDA_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return DDA_OK;
}

extern "C" dda_status dda_ctx_info(dda_ctx_t ctx, char* name, size_t name_len, size_t* hbm_bytes, int* cu_count) {
  if (!ctx) {
    set_error("dda_ctx_info: null context");
    return DDA_INVALID_ARG;
  }
  if (name && name_len) {
    std::strncpy(name, ctx->prop.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  if (hbm_bytes) *hbm_bytes = ctx->prop.totalGlobalMem;
  if (cu_count) *cu_count = ctx->prop.multiProcessorCount;
  return DDA_OK;
}

extern "C" dda_status dda_malloc(dda_ctx_t ctx, size_t bytes, void** out) {
  if (!ctx || !out) {
    set_error("dda_malloc: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY_HIP(hipSetDevice(ctx->device));
  DDA_TRY_HIP(hipMalloc(out, bytes ? bytes : 1));
  return DDA_OK;
}

extern "C" dda_status dda_free(dda_ctx_t ctx, void* ptr) {
  if (!ctx) {
    set_error("dda_free: null context");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  if (ptr) {
    DDA_TRY_HIP(hipStreamSynchronize(ctx->stream));
    DDA_TRY_HIP(hipFree(ptr));
  }
  return DDA_OK;
}

extern "C" dda_status dda_memcpy_h2d(dda_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) {
    set_error("dda_memcpy_h2d: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  DDA_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  DDA_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return DDA_OK;
}

extern "C" dda_status dda_memcpy_d2h(dda_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) {
    set_error("dda_memcpy_d2h: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  DDA_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  DDA_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return DDA_OK;
}

extern "C" dda_status dda_memcpy_d2d(dda_ctx_t ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx || (bytes && (!dst || !src))) {
    set_error("dda_memcpy_d2d: null argument");
    return DDA_INVALID_ARG;
  }
  DDA_TRY(bind_device(ctx));
  DDA_TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  DDA_TRY_HIP(hipStreamSynchronize(ctx->stream));
  return DDA_OK;
}
