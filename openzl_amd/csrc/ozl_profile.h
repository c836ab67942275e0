// Optional per-kernel timing with HIP events (used by the benchmark to time
// the dominant kernel live; off by default, no cost when off).
#pragma once
#include <hip/hip_runtime.h>

struct ozl_prof_token {
    int slot;
};
ozl_prof_token ozl_prof_start(const char* name, hipStream_t st);
void ozl_prof_stop(ozl_prof_token t, hipStream_t st);

#define OZL_LAUNCH(name, kernel, grid, block, shmem, stream, ...)                   \
    do {                                                                           \
        ozl_prof_token _t = ozl_prof_start(name, stream);                          \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);       \
        ozl_prof_stop(_t, stream);                                                 \
    } while (0)
