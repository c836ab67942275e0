// Device utilities shared by the graph engine and the universal decoder
// (ozl_util.hip): stream gathers, XXH3-64 of long device buffers, frame tails.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

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

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

// f(0) .. f(n-1) on the library's host worker threads and the caller (pure
// host work: no HIP calls). The workers (up to 15 besides the caller) are
// started once and kept; calls from several threads take turns, and a call
// made from inside f runs inline.
void ozl_pool_run(size_t n, void (*fn)(void*, size_t), void* ctx);
template <class F>
void ozl_parallel_for(size_t n, F f)
{
    ozl_pool_run(n, [](void* c, size_t i) { (*static_cast<F*>(c))(i); }, &f);
}

// Waits for up to three streams when it goes out of scope.
struct OzlStreamsDone {
    hipStream_t a, b, c = nullptr;
    ~OzlStreamsDone()
    {
        if (a) (void)hipStreamSynchronize(a);
        if (b) (void)hipStreamSynchronize(b);
        if (c) (void)hipStreamSynchronize(c);
    }
};

// Pinned host staging for the small host-to-device uploads of one call
// (metadata, headers): hipMemcpyAsync from pageable memory waits for the
// stream, from pinned memory it does not. Blocks are kept across calls;
// reset() once the call's streams have drained.
struct OzlHostStage {
    std::vector<std::pair<uint8_t*, size_t>> blocks;
    size_t cur = 0, used = 0;
    ~OzlHostStage()
    {
        for (auto& b : blocks) (void)hipHostFree(b.first);
    }
    void reset()
    {
        cur = 0;
        used = 0;
    }
    // a pinned copy of p[0, n) (p null: n bytes, not written) that stays
    // valid until reset(); null on failure
    uint8_t* put(const void* p, size_t n)
    {
        const size_t need = (n + 63) & ~(size_t)63;
        while (cur < blocks.size() && used + need > blocks[cur].second) {
            cur++;
            used = 0;
        }
        if (cur == blocks.size()) {
            size_t sz = blocks.empty() ? ((size_t)4 << 20) : blocks.back().second * 2;
            if (sz < need) sz = need;
            void* b = nullptr;
            if (hipHostMalloc(&b, sz, hipHostMallocDefault) != hipSuccess) return nullptr;
            blocks.push_back({ (uint8_t*)b, sz });
            used = 0;
        }
        uint8_t* d = blocks[cur].first + used;
        used += need;
        if (n && p) memcpy(d, p, n);
        return d;
    }
};
