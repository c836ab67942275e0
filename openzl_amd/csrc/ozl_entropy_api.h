// Internal interface between the batched entropy kernels and the graph engine.
#pragma once
#include <hip/hip_runtime.h>
#include "ozl_frame.h"

#define OZL_TAB_BYTES 10240  // per block: FSE stateTable u16[4096] + tt u32[256]; HUF len[256] + val u16[256]

struct ozl_ews {
    ozl_eplan* plans;
    uint8_t* tables;
    uint8_t* aux;
    uint8_t* payload;
    uint64_t* frameOff;  // B+1 (frame path only)
    size_t payloadCap;
    int minSS;           // ZL_CParam_minStreamSize in force (<= 0: no automatic store)
};

// A stream of this many bytes is stored by the minStreamSize rule
// (src/openzl/compress/cctx.c:1127-1140).
__host__ __device__ inline bool ozl_small(uint32_t bytes, int minSS) { return minSS > 0 && bytes < (uint32_t)minSS; }

// Payload capacity per block of at most maxBlock bytes.
static inline size_t ozl_payload_cap(size_t maxBlock) { return ((maxBlock + maxBlock / 128 + 256) + 15) & ~(size_t)15; }

// Carves the workspace into the per-block areas.
static inline ozl_ews ozl_ews_carve(void* workspace, uint32_t B, size_t maxBlock)
{
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    ozl_ews ws;
    ws.payloadCap = ozl_payload_cap(maxBlock);
    ws.plans = (ozl_eplan*)w; w += (size_t)B * sizeof(ozl_eplan);
    ws.tables = w; w += (size_t)B * OZL_TAB_BYTES;
    ws.aux = w; w += (size_t)B * OZL_AUX_BYTES;
    ws.payload = w;
    ws.frameOff = nullptr;
    ws.minSS = OZL_MIN_STREAM_SIZE;
    return ws;
}

// Plans and encodes B chunks (chunk b = src[off[b], offEnd[b]), each < 100000
// bytes; offEnd = off + 1 for back-to-back chunks) for the FSE (hufMode 0) or
// Huffman (1) graph: plans, tables, NCount / weights (with their nested graph)
// and payloads land in the workspace.
int ozl_entropy_encode_chunks(int hufMode, const uint8_t* src, const uint64_t* off, const uint64_t* offEnd, uint32_t B,
                              size_t maxBlock, const ozl_ews& ws, hipStream_t st);
