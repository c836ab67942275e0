// Device utilities of the graph engine: gathering stored streams into a frame
// and XXH3-64 of long device buffers.
#include "ozl_internal.h"
#include "ozl_xxh3.h"
#include "ozl_gmem.h"
#include "ozl_profile.h"
#include <condition_variable>
#include <mutex>

#include "ozl_engine_util.h"

namespace {

constexpr int kT = 256;
constexpr size_t kGatherPiece = 1 << 16;

// grid.y = stream, grid.x = 64 KiB piece of the stream. Destination dwords
// are written whole; a source that is not dword-aligned with its destination
// is read as two aligned dwords and a funnel shift (an aligned dword holding
// a byte of the stream lies in the same page as that byte, so no read leaves
// the stream's pages). The few bytes before the first and after the last
// whole destination dword of a piece are copied one by one.
__global__ void __launch_bounds__(kT) k_gather(const uint8_t* const* __restrict__ srcs, const uint64_t* __restrict__ sizes,
                                               const uint64_t* __restrict__ dstOffs, uint8_t* __restrict__ dst)
{
    const int k = blockIdx.y;
    const size_t size = sizes[k];
    const size_t a = (size_t)blockIdx.x * kGatherPiece;
    if (a >= size) return;
    const size_t b = a + kGatherPiece < size ? a + kGatherPiece : size;
    const uint8_t* s = srcs[k];
    uint8_t* d = dst + dstOffs[k];
    const size_t h = (4 - ((uintptr_t)d & 3)) & 3;     // first position whose destination is dword-aligned
    const size_t lo = a + h < b ? a + h : b;
    const size_t hi = lo + ((b - lo) & ~(size_t)3);
    const size_t nw = (hi - lo) >> 2;
    const uintptr_t sa = (uintptr_t)(s + lo);
    const uint32_t sh = (uint32_t)(sa & 3) * 8;
    const uint8_t* s4 = (const uint8_t*)(sa & ~(uintptr_t)3);
    uint8_t* d4 = d + lo;
    if (sh == 0) {
        for (size_t j = threadIdx.x; j < nw; j += kT) ozl_gst4(d4 + 4 * j, ozl_gld4(s4 + 4 * j));
    } else {
        for (size_t j = threadIdx.x; j < nw; j += kT) {
            const uint32_t w0 = ozl_gld4(s4 + 4 * j), w1 = ozl_gld4(s4 + 4 * j + 4);
            ozl_gst4(d4 + 4 * j, __builtin_amdgcn_alignbit(w1, w0, sh));
        }
    }
    if (threadIdx.x < lo - a) d[a + threadIdx.x] = s[a + threadIdx.x];
    if (threadIdx.x < b - hi) d[hi + threadIdx.x] = s[hi + threadIdx.x];
}

// XXH3 of a long buffer: per 1 KiB block, the 8 accumulator sums of its 16
// stripes (independent across blocks), by one wave per 8 blocks: lane l
// reads bytes 16 l .. 16 l + 15 of a block (stripe l / 4, words 2 (l % 4)
// and 2 (l % 4) + 1) with 16-byte loads, all 8 blocks' loads in flight
// together, and the 16 stripes' terms are summed across lanes
// (ozl_xxh3_sums_of) ...
static_assert(kT == 256, "a block-sum workgroup covers 32 blocks: 4 waves of 8");
__device__ inline void xxh3_block_sums(const uint8_t* in, size_t len, uint64_t* sums, size_t b0)
{
    const size_t nbBlocks = len > 240 ? (len - 1) / 1024 : 0;
    if (b0 >= nbBlocks) return;
    const ozl_xxh3_lane L = ozl_xxh3_lane_init();
    const uint32_t m = nbBlocks - b0 < 8 ? (uint32_t)(nbBlocks - b0) : 8u;
    uint64_t v0[8], v1[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++)
        if (k < m) ozl_xxh3_block_load(in + (b0 + k) * 1024, ozl_xxh3_mode(in, len, b0 + k), v0[k], v1[k]);
#pragma unroll
    for (uint32_t k = 0; k < 8; k++) {
        if (k < m) {
            uint64_t sa, sb;
            ozl_xxh3_sums_of(v0[k], v1[k], L, sa, sb);
            if ((threadIdx.x & 63) < 4) {
                sums[(b0 + k) * 8 + 2 * L.j] = sa;
                sums[(b0 + k) * 8 + 2 * L.j + 1] = sb;
            }
        }
    }
}

// ... then the scramble chain, the tail and the merge by one wave: lane i
// (mod 8) runs accumulator word i. The chain's vector steps are a 64-bit add
// (one instruction), a shift and a three-way xor, and one 32 x 32 -> 64-bit
// multiply-add for the product with the 32-bit prime; the block sums come
// through per-lane loads issued a batch of 32 blocks ahead, which covers the
// memory latency (loads complete in order).
__constant__ uint64_t kXxh3Init[8] = { OZL_P32_3, OZL_P64_1, OZL_P64_2, OZL_P64_3,
                                       OZL_P64_4, OZL_P32_2, OZL_P64_5, OZL_P32_1 };

__device__ __attribute__((always_inline)) inline void xxh3_finish_wave(const uint8_t* __restrict__ in, size_t len, const uint64_t* __restrict__ sums,
                                        uint64_t* __restrict__ out, uint64_t* accs)
{
    if (len <= 240) {
        if (threadIdx.x == 0) *out = ozl_xxh3_short(in, len);
        return;
    }
    const int w = (int)(threadIdx.x & 7);
    uint64_t acc = kXxh3Init[w];
    const size_t nbBlocks = (len - 1) / 1024;
    const uint64_t sec = ozl_sec64(192 - 64 + 8 * w);
    const uint64_t* sp = sums + w;
    const uint32_t secLo = (uint32_t)sec, secHi = (uint32_t)(sec >> 32);
    // acc = ((acc + v) ^ ((acc + v) >> 47) ^ sec) * P32_1 in 32-bit halves:
    // the high half's product runs beside the low half's shift and xor
    auto scr = [&](uint64_t v) {
        const uint64_t a = acc + v;
        const uint32_t lo = (uint32_t)a, hi = (uint32_t)(a >> 32);
        const uint32_t l2 = lo ^ (hi >> 15) ^ secLo;
        const uint32_t h2 = (hi ^ secHi) * (uint32_t)OZL_P32_1;
        acc = (uint64_t)l2 * (uint32_t)OZL_P32_1 + ((uint64_t)h2 << 32);
    };
    constexpr int kB = 32;
    // the block count is uniform: kept in a scalar register, so that the
    // batch loop's branches are scalar ones
    const uint32_t nb = (uint32_t)__builtin_amdgcn_readfirstlane((int)nbBlocks);
    if (nb > 0) {
        // two batches of sums in flight at a time: a batch is scrambled while
        // the next one loads. Loads of blocks past the end read the last one
        // again (in bounds, never scrambled): no branch around a load.
        const uint32_t last = nb - 1;
        uint64_t A[kB], B[kB];
        auto load = [&](uint64_t* d, uint32_t b0) {
#pragma unroll
            for (int k = 0; k < kB; k++) {
                const uint32_t x = b0 + (uint32_t)k;
                d[k] = sp[8 * (size_t)(x < last ? x : last)];
            }
        };
        auto run = [&](const uint64_t* v) {
#pragma unroll
            for (int k = 0; k < kB; k++) scr(v[k]);
        };
        load(A, 0);
        uint32_t b = 0;
        for (;;) {
            if (b + kB > nb) break;
            load(B, b + kB);
            run(A);
            b += kB;
            if (b + kB > nb) {
#pragma unroll
                for (int k = 0; k < kB; k++) A[k] = B[k];
                break;
            }
            load(A, b + kB);
            run(B);
            b += kB;
        }
        // the last nb mod 32 blocks (in A), in pieces of 16, 8, 4, 2, 1
        const uint32_t r = nb - b;
#pragma unroll
        for (int piece = 16; piece >= 1; piece >>= 1) {
            if (r & piece) {
#pragma unroll
                for (int k = 0; k < piece; k++) scr(A[k]);
#pragma unroll
                for (int k = 0; k + piece < kB; k++) A[k] = A[k + piece];
            }
        }
    }
    const size_t nbStripes = ((len - 1) - 1024 * nbBlocks) / 64;
    for (size_t s = 0; s < nbStripes; s++) acc += ozl_stripe_term(in + nbBlocks * 1024 + s * 64, (int)s * 8, w);
    acc += ozl_stripe_term(in + len - 64, 192 - 64 - 7, w);
    if (threadIdx.x < 8) accs[w] = acc;
    __syncthreads();
    if (threadIdx.x == 0) *out = ozl_xxh3_merge(accs, len);
}

__global__ void __launch_bounds__(kT) k_xxh3_blocks(const uint8_t* __restrict__ in, size_t len,
                                                    uint64_t* __restrict__ sums)
{
    xxh3_block_sums(in, len, sums, (size_t)blockIdx.x * 32 + 8 * (threadIdx.x >> 6));
}

__global__ void __launch_bounds__(64) k_xxh3_finish(const uint8_t* __restrict__ in, size_t len,
                                                    const uint64_t* __restrict__ sums, uint64_t* __restrict__ out)
{
    __shared__ uint64_t accs[8];
    xxh3_finish_wave(in, len, sums, out, accs);
}

// many buffers: grid.y = buffer (block sums), grid.x = buffer (chains)
__global__ void __launch_bounds__(kT) k_xxh3_blocks_n(const ozl_xjob* __restrict__ jobs, uint64_t* __restrict__ sums)
{
    const ozl_xjob j = jobs[blockIdx.y];
    xxh3_block_sums((const uint8_t*)(uintptr_t)j.ptr, j.len, sums + j.sumOff, (size_t)blockIdx.x * 32 + 8 * (threadIdx.x >> 6));
}

__global__ void __launch_bounds__(64) k_xxh3_finish_n(const ozl_xjob* __restrict__ jobs,
                                                      const uint64_t* __restrict__ sums, uint64_t* __restrict__ out)
{
    __shared__ uint64_t accs[8];
    const ozl_xjob j = jobs[blockIdx.x];
    xxh3_finish_wave((const uint8_t*)(uintptr_t)j.ptr, j.len, sums + j.sumOff, out + blockIdx.x, accs);
}

// frame tails: the low 4 bytes of hashes[k] (little endian) at dst + at[k]
// (when hashes is given), then a zero end byte when endByte is set
__global__ void __launch_bounds__(64) k_put_tails(uint8_t* __restrict__ dst, const uint64_t* __restrict__ at,
                                                  const uint64_t* __restrict__ hashes, uint32_t count, int endByte)
{
    const uint32_t k = blockIdx.x * 64 + threadIdx.x;
    if (k >= count) return;
    uint8_t* p = dst + at[k];
    if (hashes) {
        const uint64_t h = hashes[k];
        for (int q = 0; q < 4; q++) p[q] = (uint8_t)(h >> (8 * q));
        p += 4;
    }
    if (endByte) p[0] = 0;
}

}  // namespace

// ---------------------------------------------------------------- host worker threads
namespace {
struct HostPool {
    std::mutex callMu;   // one job at a time
    std::mutex mu;
    std::condition_variable cv, done;
    std::vector<std::thread> th;
    void (*fn)(void*, size_t) = nullptr;
    void* ctx = nullptr;
    size_t n = 0;
    std::atomic<size_t> next{ 0 };
    size_t active = 0;   // workers still on the job
    uint64_t gen = 0;
    bool quit = false;
    ~HostPool()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            quit = true;
        }
        cv.notify_all();
        for (auto& t : th) t.join();
    }
};
thread_local bool tInHostPool = false;

void host_worker(HostPool* P)
{
    tInHostPool = true;
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(P->mu);
    for (;;) {
        P->cv.wait(lk, [&]() { return P->quit || P->gen != seen; });
        if (P->quit) return;
        seen = P->gen;
        void (*fn)(void*, size_t) = P->fn;
        void* ctx = P->ctx;
        const size_t n = P->n;
        lk.unlock();
        for (size_t i; (i = P->next.fetch_add(1)) < n;) fn(ctx, i);
        lk.lock();
        if (--P->active == 0) P->done.notify_all();
    }
}
}  // namespace

void ozl_pool_run(size_t n, void (*fn)(void*, size_t), void* ctx)
{
    if (tInHostPool || n < 4) {
        for (size_t i = 0; i < n; i++) fn(ctx, i);
        return;
    }
    static HostPool P;
    std::lock_guard<std::mutex> call(P.callMu);
    {
        std::lock_guard<std::mutex> lk(P.mu);
        if (P.th.empty()) {
            const unsigned hc = std::thread::hardware_concurrency();
            const unsigned W = (hc ? (hc < 16 ? hc : 16) : 1) - 1;
            for (unsigned t = 0; t < W; t++) P.th.emplace_back(host_worker, &P);
        }
        P.fn = fn;
        P.ctx = ctx;
        P.n = n;
        P.next = 0;
        P.active = P.th.size();
        P.gen++;
    }
    P.cv.notify_all();
    tInHostPool = true;
    for (size_t i; (i = P.next.fetch_add(1)) < n;) fn(ctx, i);
    tInHostPool = false;
    std::unique_lock<std::mutex> lk(P.mu);
    P.done.wait(lk, [&]() { return P.active == 0; });
}

size_t ozl_xxh3_scratch_size(size_t len) { return ((len / 1024) + 1) * 64 + 256; }

// XXH3-64 of `len` device bytes into *dOut (device), asynchronous.
int ozl_xxh3_device(const uint8_t* data, size_t len, uint64_t* dOut, void* scratch, size_t scratchSize, hipStream_t st)
{
    size_t nbBlocks = len > 240 ? (len - 1) / 1024 : 0;
    if (scratchSize < ozl_xxh3_scratch_size(len)) return OZL_ERR_INTERNAL_BUFFER;
    uint64_t* sums = (uint64_t*)scratch;
    if (nbBlocks) OZL_LAUNCH("k_xxh3_blocks", k_xxh3_blocks, dim3((unsigned)((nbBlocks + 31) / 32)), dim3(kT), 0, st, data, len, sums);
    OZL_LAUNCH("k_xxh3_finish", k_xxh3_finish, dim3(1), dim3(64), 0, st, data, len, sums, dOut);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

size_t ozl_xxh3_batch_sums(size_t len) { return len > 240 ? ((len - 1) / 1024) * 8 : 0; }

// XXH3-64 of `count` device buffers (jobs: device array; the block sums of
// job k at sums + jobs[k].sumOff, ozl_xxh3_batch_sums(len) words) into
// dOut[k], asynchronous.
int ozl_xxh3_device_batch(const ozl_xjob* jobs, uint32_t count, size_t maxLen, uint64_t* sums, uint64_t* dOut,
                          hipStream_t st)
{
    if (count == 0) return OZL_OK;
    const size_t maxBlocks = maxLen > 240 ? (maxLen - 1) / 1024 : 0;
    if (maxBlocks)
        for (uint32_t y0 = 0; y0 < count; y0 += 65535) {
            const uint32_t cnt = count - y0 < 65535 ? count - y0 : 65535;
            OZL_LAUNCH("k_xxh3_blocks_n", k_xxh3_blocks_n, dim3((unsigned)((maxBlocks + 31) / 32), cnt), dim3(kT), 0, st,
                       jobs + y0, sums);
        }
    OZL_LAUNCH("k_xxh3_finish_n", k_xxh3_finish_n, dim3(count), dim3(64), 0, st, jobs, sums, dOut);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Writes frame tails (see k_put_tails); at and hashes are device arrays.
int ozl_put_tails(uint8_t* dst, const uint64_t* at, const uint64_t* hashes, uint32_t count, int endByte, hipStream_t st)
{
    if (count == 0) return OZL_OK;
    OZL_LAUNCH("k_put_tails", k_put_tails, dim3((count + 63) / 64), dim3(64), 0, st, dst, at, hashes, count, endByte);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Whether p points into device memory (host memory, pinned or pageable, is not).
bool ozl_is_device_ptr(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();   // plain host memory: clear the error so no later check sees it
        return false;
    }
    return a.type == hipMemoryTypeDevice;
}

// Copies `count` device buffers srcs[k] (sizes[k] bytes, at most maxSize) to
// dst + dstOffs[k]; the three arrays are device arrays.
int ozl_gather_sized(const uint8_t* const* srcs, const uint64_t* sizes, const uint64_t* dstOffs, int count,
                     size_t maxSize, uint8_t* dst, hipStream_t st)
{
    if (count <= 0) return OZL_OK;
    unsigned gx = (unsigned)((maxSize + kGatherPiece - 1) / kGatherPiece);
    if (gx == 0) gx = 1;
    for (int k0 = 0; k0 < count; k0 += 65535) {
        int cnt = count - k0 < 65535 ? count - k0 : 65535;
        OZL_LAUNCH("k_gather", k_gather, dim3(gx, cnt), dim3(kT), 0, st, srcs + k0, sizes + k0, dstOffs + k0, dst);
    }
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
