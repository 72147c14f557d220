// SafeCall.h — error behaviour of the reference (Utility/SafeCall.h:13-21):
// a failed call prints its message and ends the process with exit(0).
// `SafeCallMode::Throw` turns the same failures into `SafeCallError`
// exceptions instead (for tests and for callers that recover).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "sicp.h"

enum class SafeCallMode { Exit, Throw };

inline SafeCallMode& SafeCallModeRef() {
    static SafeCallMode mode = SafeCallMode::Exit;
    return mode;
}

struct SafeCallError : std::runtime_error {
    int status;
    SafeCallError(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void SafeCallFail(int status, const std::string& msg) {
    if (SafeCallModeRef() == SafeCallMode::Throw) throw SafeCallError(status, msg);
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(0);
}

// HIP runtime status.
inline void SafeCall(hipError_t err) {
    if (err != hipSuccess)
        SafeCallFail(SICP_ERR_HIP, std::string("HIP error: ") + hipGetErrorString(err));
}

// Status of a C ABI entry point.
inline void SicpCall(int status, const char* what) {
    if (status != SICP_OK)
        SafeCallFail(status, std::string(what) + ": " + sicp_last_error());
}
