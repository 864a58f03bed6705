// Engine context shared by the single-LP sessions and the batched solver:
// one device, one non-blocking HIP stream (default priority), last error.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct nir_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string error;
};
