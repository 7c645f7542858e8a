// Helpers shared by the C-ABI translation units: status/error recording,
// HIP error mapping and the context's staging buffer for host pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <new>
#include <string>

#include "host/errors.hpp"
#include "internal.hpp"

namespace re {

inline re_status set_err(re_ctx* ctx, re_status st, const char* msg) {
  fail(st, "%s", msg);
  if (ctx) ctx->err = msg;
  return st;
}

inline re_status set_err(re_ctx* ctx, re_status st, const std::string& msg) { return set_err(ctx, st, msg.c_str()); }

inline re_status hip_err(re_ctx* ctx, hipError_t e, const char* where) {
  char buf[400];
  snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  return set_err(ctx, e == hipErrorOutOfMemory ? RE_ERR_OOM : RE_ERR_HIP, buf);
}

#define HIP_TRY(ctx, expr, where)                          \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return re::hip_err(ctx, _e, where); \
  } while (0)

// Runs an entry point's body; a host allocation that fails inside it (a
// std::vector sized from the caller's arguments or a file header) becomes
// RE_ERR_OOM instead of an exception crossing the C ABI.
template <class F>
re_status guarded(re_ctx* ctx, const char* what, F&& body) {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return set_err(ctx, RE_ERR_OOM, std::string(what) + ": host allocation failed");
  } catch (const std::length_error&) {
    return set_err(ctx, RE_ERR_OOM, std::string(what) + ": host allocation failed");
  }
}

// Grows the context's staging buffer to at least `bytes`.
inline re_status ensure_stage(re_ctx* ctx, size_t bytes) {
  if (ctx->stage_bytes >= bytes) return RE_OK;
  HIP_TRY(ctx, hipStreamSynchronize(ctx->stream), "stage sync");
  if (ctx->stage) (void)hipFree(ctx->stage);
  ctx->stage = nullptr;
  ctx->stage_bytes = 0;
  size_t want = bytes + bytes / 4;
  HIP_TRY(ctx, hipMalloc(&ctx->stage, want), "stage hipMalloc");
  ctx->stage_bytes = want;
  return RE_OK;
}

// Carves aligned sub-buffers out of the staging buffer.
struct StageCursor {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~(size_t)255;
    T* p = (T*)(base + off);
    off += count * sizeof(T);
    return p;
  }
};

inline size_t stage_size(std::initializer_list<size_t> parts) {
  size_t s = 0;
  for (size_t p : parts) s += ((p + 255) & ~(size_t)255);
  return s + 256;
}

}  // namespace re
