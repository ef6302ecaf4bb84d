// cospanmdao_amd/csrc/internal.hpp — shared by the translation units of libcospan_hip (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include "cospan_hip.h"

namespace cs_hip {

/// Records a thread-local message and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
/// CS_HIP_OK, or CS_HIP_ERR_RUNTIME with "<what>: <hip error string>".
int check(hipError_t e, const char* what);
/// The HIP stream behind a handle (null: the null stream).
hipStream_t raw(cs_hip_stream s);
/// The device address of the current device's fault word (looked up once per device).
int fault_word(uint32_t** out);

}  // namespace cs_hip
