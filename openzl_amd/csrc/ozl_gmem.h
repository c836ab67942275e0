// ozl_gmem.h -- loads and stores through global-address-space pointers.
//
// A generic pointer compiles to flat_load / flat_store, which count in both
// vmcnt and lgkmcnt: every wait for an LDS or shuffle result then also waits
// for the memory operations in flight. These helpers cast to address space 1
// (global), whose instructions count in vmcnt only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef unsigned int ozl_u32x4 __attribute__((ext_vector_type(4)));
#define OZL_GLOBAL(T) __attribute__((address_space(1))) T

__device__ inline ozl_u32x4 ozl_gld16(const void* p) { return *(const OZL_GLOBAL(ozl_u32x4)*)(uintptr_t)p; }
__device__ inline uint64_t ozl_gld8(const void* p) { return *(const OZL_GLOBAL(uint64_t)*)(uintptr_t)p; }
__device__ inline uint32_t ozl_gld4(const void* p) { return *(const OZL_GLOBAL(uint32_t)*)(uintptr_t)p; }
__device__ inline uint32_t ozl_gld2(const void* p) { return *(const OZL_GLOBAL(uint16_t)*)(uintptr_t)p; }
__device__ inline uint32_t ozl_gld1(const void* p) { return *(const OZL_GLOBAL(uint8_t)*)(uintptr_t)p; }
__device__ inline void ozl_gst8(void* p, uint64_t v) { *(OZL_GLOBAL(uint64_t)*)(uintptr_t)p = v; }
__device__ inline void ozl_gst4(void* p, uint32_t v) { *(OZL_GLOBAL(uint32_t)*)(uintptr_t)p = v; }
__device__ inline void ozl_gst2(void* p, uint32_t v) { *(OZL_GLOBAL(uint16_t)*)(uintptr_t)p = (uint16_t)v; }
__device__ inline void ozl_gst1(void* p, uint32_t v) { *(OZL_GLOBAL(uint8_t)*)(uintptr_t)p = (uint8_t)v; }
// element of width w (1, 2, 4, 8) to p
__device__ inline void ozl_gst_w(void* p, uint64_t v, int w)
{
    switch (w) {
    case 1: ozl_gst1(p, (uint32_t)v); break;
    case 2: ozl_gst2(p, (uint32_t)v); break;
    case 4: ozl_gst4(p, (uint32_t)v); break;
    default: ozl_gst8(p, v); break;
    }
}
// element of width w (1, 2, 4, 8) at p
__device__ inline uint64_t ozl_gld_w(const void* p, int w)
{
    switch (w) {
    case 1: return ozl_gld1(p);
    case 2: return ozl_gld2(p);
    case 4: return ozl_gld4(p);
    default: return ozl_gld8(p);
    }
}
