// Internal helpers shared by the HIP translation units of libozl_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "ozl_errors.h"

#define OZL_HIP_CHECK(expr)                                   \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return OZL_ERR_HIP;             \
    } while (0)

static inline unsigned ozl_blocks(size_t work, unsigned perBlock, unsigned cap = 1u << 20)
{
    size_t b = (work + perBlock - 1) / perBlock;
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}
