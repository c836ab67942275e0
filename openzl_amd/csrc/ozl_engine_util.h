// Device utilities shared by the graph engine and the universal decoder
// (ozl_util.hip): stream gathers, XXH3-64 of long device buffers, frame tails.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

// one XXH3 job of a batch: ptr/len of the buffer, its block sums' offset (words)
struct ozl_xjob {
    uint64_t ptr, len, sumOff;
};

int ozl_xxh3_device(const uint8_t* data, size_t len, uint64_t* dOut, void* scratch, size_t scratchSize, hipStream_t st);
size_t ozl_xxh3_scratch_size(size_t len);
size_t ozl_xxh3_batch_sums(size_t len);
int ozl_xxh3_device_batch(const ozl_xjob* jobs, uint32_t count, size_t maxLen, uint64_t* sums, uint64_t* dOut,
                          hipStream_t st);
int ozl_put_tails(uint8_t* dst, const uint64_t* at, const uint64_t* hashes, uint32_t count, int endByte, hipStream_t st);
int ozl_gather_sized(const uint8_t* const* srcs, const uint64_t* sizes, const uint64_t* dstOffs, int count,
                     size_t maxSize, uint8_t* dst, hipStream_t st);
bool ozl_is_device_ptr(const void* p);
