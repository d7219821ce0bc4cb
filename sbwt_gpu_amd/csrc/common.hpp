// Shared plumbing of libsbwt_gpu_amd: status codes, per-thread error text,
// HIP call checking.
#pragma once
#include <cstring>  // before rocPRIM headers, which call memset in host code
#include <cstdint>
#include <cstdio>
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/sbwt_gpu_amd.h"

namespace sbwt {

using u64 = uint64_t;
using u32 = uint32_t;
using u8 = uint8_t;
using u16 = uint16_t;

constexpr u64 kNone = ~0ull;

// Records the message for sbwt_last_error() and returns the code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace sbwt

#define SBWT_HIP(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::sbwt::fail(SBWT_ERR_HIP, "%s failed: %s (%s:%d)", #call,              \
                          hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

#define SBWT_TRY(expr)         \
  do {                         \
    int s_ = (expr);           \
    if (s_ != SBWT_OK) return s_; \
  } while (0)

#define SBWT_REQUIRE(cond, ...)                            \
  do {                                                     \
    if (!(cond)) return ::sbwt::fail(SBWT_ERR_ARG, __VA_ARGS__); \
  } while (0)
