// status.h -- status codes shared by the C-ABI and the internal C++ code.
#pragma once

#include "../../include/ldpc_api.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

namespace nrldpc {

// Records the last error message of the calling thread (read back by ldpc_last_error()).
void set_last_error(const std::string& msg);

}  // namespace nrldpc

// Evaluate a HIP runtime call; on failure record the message and return LDPC_EHIP.
#define NRLDPC_HIP(call)                                                                   \
    do {                                                                                   \
        hipError_t nrldpc_e_ = (call);                                                     \
        if (nrldpc_e_ != hipSuccess) {                                                     \
            ::nrldpc::set_last_error(std::string(#call) + ": " + hipGetErrorString(nrldpc_e_)); \
            return LDPC_EHIP;                                                              \
        }                                                                                  \
    } while (0)
