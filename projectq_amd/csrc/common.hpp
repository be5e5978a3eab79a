// Shared types and error plumbing of the MI355X simulator library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>

#include "layout.hpp"

namespace qs {

// Error classes surfaced through the C ABI (QS_ERR_* in include/projectq_amd.h).
struct ValueError : std::invalid_argument {
  using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct CallbackError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}
#define QS_HIP(expr) ::qs::hip_check((expr), #expr)

// Amplitude types: interleaved (re, im).
template <typename R>
struct Amp;
template <>
struct Amp<double> {
  using type = double2;
};
template <>
struct Amp<float> {
  using type = float2;
};
template <typename R>
using amp_t = typename Amp<R>::type;

}  // namespace qs
