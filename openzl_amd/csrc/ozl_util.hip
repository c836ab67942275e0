// Device utilities of the graph engine: gathering stored streams into a frame
// and XXH3-64 of long device buffers.
#include "ozl_internal.h"
#include "ozl_xxh3.h"
#include "ozl_profile.h"

namespace {

constexpr int kT = 256;
constexpr size_t kGatherPiece = 1 << 16;

// grid.y = stream, grid.x = 64 KiB piece of the stream
__global__ void __launch_bounds__(kT) k_gather(const uint8_t* const* __restrict__ srcs, const uint64_t* __restrict__ sizes,
                                               const uint64_t* __restrict__ dstOffs, uint8_t* __restrict__ dst)
{
    int k = blockIdx.y;
    size_t size = sizes[k];
    size_t a = (size_t)blockIdx.x * kGatherPiece;
    if (a >= size) return;
    size_t b = a + kGatherPiece < size ? a + kGatherPiece : size;
    const uint8_t* s = srcs[k];
    uint8_t* d = dst + dstOffs[k];
    for (size_t i = a + threadIdx.x; i < b; i += kT) d[i] = s[i];
}

// XXH3 of a long buffer: per 1 KiB block, the 8 accumulator sums of its 16
// stripes (independent across blocks) ...
__global__ void __launch_bounds__(kT) k_xxh3_blocks(const uint8_t* __restrict__ in, size_t nbBlocks,
                                                    uint64_t* __restrict__ sums)
{
    size_t blk = (size_t)blockIdx.x * (kT / 8) + (threadIdx.x >> 3);
    int i = threadIdx.x & 7;
    if (blk >= nbBlocks) return;
    const uint8_t* p = in + blk * 1024;
    uint64_t s = 0;
    for (int st = 0; st < 16; st++) s += ozl_stripe_term(p + st * 64, st * 8, i);
    sums[blk * 8 + i] = s;
}

// ... then one wavefront runs the scramble chain, the tail and the merge.
__global__ void __launch_bounds__(64) k_xxh3_finish(const uint8_t* __restrict__ in, size_t len,
                                                    const uint64_t* __restrict__ sums, uint64_t* __restrict__ out)
{
    int lane = threadIdx.x;
    if (len <= 240) {
        if (lane == 0) *out = ozl_xxh3_short(in, len);
        return;
    }
    const uint64_t init[8] = { OZL_P32_3, OZL_P64_1, OZL_P64_2, OZL_P64_3, OZL_P64_4, OZL_P32_2, OZL_P64_5, OZL_P32_1 };
    uint64_t acc = lane < 8 ? init[lane] : 0;
    size_t nbBlocks = (len - 1) / 1024;
    if (lane < 8) {
        // the chain is serial: the block sums are read 32 blocks ahead of
        // their use, and the lane's secret word is read once
        const uint64_t sec = ozl_sec64(192 - 64 + 8 * lane);
        constexpr int kAhead = 32;
        uint64_t cur[kAhead], nxt[kAhead];
        size_t b = 0;
#pragma unroll
        for (int k = 0; k < kAhead; k++) cur[k] = (size_t)k < nbBlocks ? sums[(size_t)k * 8 + lane] : 0;
        for (; b + kAhead <= nbBlocks; b += kAhead) {
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                const size_t q = b + kAhead + (size_t)k;
                nxt[k] = q < nbBlocks ? sums[q * 8 + lane] : 0;
            }
#pragma unroll
            for (int k = 0; k < kAhead; k++) {
                uint64_t a2 = acc + cur[k];
                a2 ^= a2 >> 47;
                a2 ^= sec;
                acc = a2 * OZL_P32_1;
            }
#pragma unroll
            for (int k = 0; k < kAhead; k++) cur[k] = nxt[k];
        }
        for (; b < nbBlocks; b++) {
            uint64_t a2 = acc + sums[b * 8 + lane];
            a2 ^= a2 >> 47;
            a2 ^= sec;
            acc = a2 * OZL_P32_1;
        }
        size_t nbStripes = ((len - 1) - 1024 * nbBlocks) / 64;
        for (size_t s = 0; s < nbStripes; s++) acc += ozl_stripe_term(in + nbBlocks * 1024 + s * 64, (int)s * 8, lane);
        acc += ozl_stripe_term(in + len - 64, 192 - 64 - 7, lane);
    }
    uint64_t a[8];
    for (int i = 0; i < 8; i++) a[i] = __shfl(acc, i, 64);
    if (lane == 0) *out = ozl_xxh3_merge(a, len);
}

}  // namespace

size_t ozl_xxh3_scratch_size(size_t len) { return ((len / 1024) + 1) * 64 + 256; }

// XXH3-64 of `len` device bytes into *dOut (device), asynchronous.
int ozl_xxh3_device(const uint8_t* data, size_t len, uint64_t* dOut, void* scratch, size_t scratchSize, hipStream_t st)
{
    size_t nbBlocks = len > 240 ? (len - 1) / 1024 : 0;
    if (scratchSize < ozl_xxh3_scratch_size(len)) return OZL_ERR_INTERNAL_BUFFER;
    uint64_t* sums = (uint64_t*)scratch;
    if (nbBlocks) OZL_LAUNCH("k_xxh3_blocks", k_xxh3_blocks, dim3((unsigned)((nbBlocks + 31) / 32)), dim3(kT), 0, st, data, nbBlocks, sums);
    OZL_LAUNCH("k_xxh3_finish", k_xxh3_finish, dim3(1), dim3(64), 0, st, data, len, sums, dOut);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
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
