// Numeric transforms of the codec hot path on gfx950: delta_int, zigzag,
// transpose_split and bitpack, encode and decode.
//
// All kernels are HBM-bound elementwise or scan work. Layout rules used
// throughout: every lane moves 16 bytes per global access where the element
// width allows it (128-bit loads/stores), consecutive lanes touch consecutive
// 16-byte pieces, and grids are grid-stride so a launch of a few thousand
// workgroups covers any size. Streams inside frames may start at any byte, so
// every kernel that reads a compressed stream uses byte or unaligned-safe access.
#include "ozl_internal.h"

namespace {

template <int W> struct UIntOf;
template <> struct UIntOf<1> { using T = uint8_t; };
template <> struct UIntOf<2> { using T = uint16_t; };
template <> struct UIntOf<4> { using T = uint32_t; };
template <> struct UIntOf<8> { using T = uint64_t; };

constexpr int kThreads = 256;

// ---------------------------------------------------------------- delta
// Reference: src/openzl/codecs/delta/encode_delta_kernel.c:20-123.
// out[i] = src[i+1] - src[i] (mod 2^8W), i in [0, n-1). Each lane produces
// V = 16/W outputs from one 128-bit load plus the next element.
template <int W>
__global__ void __launch_bounds__(kThreads) k_delta_encode(const uint8_t* __restrict__ src, size_t n,
                                                           uint8_t* __restrict__ out)
{
    using T = typename UIntOf<W>::T;
    constexpr int V = 16 / W;
    const T* s = reinterpret_cast<const T*>(src);
    T* o = reinterpret_cast<T*>(out);
    size_t nOut = n - 1;
    size_t nVec = nOut / V;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nVec; v += stride) {
        size_t e = v * V;
        union { uint4 q; T t[V]; } a, r;
        a.q = *reinterpret_cast<const uint4*>(s + e);
        T next = s[e + V];
#pragma unroll
        for (int k = 0; k < V - 1; k++) r.t[k] = (T)(a.t[k + 1] - a.t[k]);
        r.t[V - 1] = (T)(next - a.t[V - 1]);
        *reinterpret_cast<uint4*>(o + e) = r.q;
    }
    // tail
    size_t t0 = nVec * V;
    for (size_t i = t0 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nOut; i += stride)
        o[i] = (T)(s[i + 1] - s[i]);
}

// Delta decode = inclusive prefix sum of the deltas seeded by the header value
// (src/openzl/codecs/delta/decode_delta_kernel.c:14-267). Three passes over
// tiles of kThreads*16 elements: tile sums, scan of tile sums, local scan.
constexpr int kScanPerThread = 16;
constexpr int kScanTile = kThreads * kScanPerThread;

template <typename T>
__device__ T block_exclusive_scan(T v, T* lds, T& total)
{
    // wave-level inclusive scan via shuffles, then across the 4 waves
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x = (T)(x + y);
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T waveOff = 0;
    T all = 0;
    for (int w = 0; w < kThreads / 64; w++) {
        T s = lds[w];
        if (w < wave) waveOff = (T)(waveOff + s);
        all = (T)(all + s);
    }
    __syncthreads();
    total = all;
    return (T)(waveOff + x - v);
}

template <int W>
__global__ void __launch_bounds__(kThreads) k_scan_tile_sums(const uint8_t* __restrict__ deltas, size_t m,
                                                             uint8_t* __restrict__ sums)
{
    using T = typename UIntOf<W>::T;
    __shared__ T lds[kThreads / 64];
    const T* d = reinterpret_cast<const T*>(deltas);
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPerThread;
    T acc = 0;
    for (int k = 0; k < kScanPerThread; k++) {
        size_t i = base + k;
        if (i < m) acc = (T)(acc + d[i]);
    }
    T total;
    block_exclusive_scan<T>(acc, lds, total);
    if (threadIdx.x == 0) reinterpret_cast<T*>(sums)[blockIdx.x] = total;
}

// one workgroup: exclusive scan of the tile sums, seeded with `first`
template <int W>
__global__ void __launch_bounds__(kThreads) k_scan_sums(uint8_t* __restrict__ sums, size_t nTiles,
                                                        const uint8_t* __restrict__ first)
{
    using T = typename UIntOf<W>::T;
    __shared__ T lds[kThreads / 64];
    T* s = reinterpret_cast<T*>(sums);
    T carry = *reinterpret_cast<const T*>(first);
    for (size_t b = 0; b < nTiles; b += kThreads) {
        size_t i = b + threadIdx.x;
        T v = i < nTiles ? s[i] : (T)0;
        T total;
        T ex = block_exclusive_scan<T>(v, lds, total);
        if (i < nTiles) s[i] = (T)(carry + ex);
        carry = (T)(carry + total);
    }
}

template <int W>
__global__ void __launch_bounds__(kThreads) k_scan_apply(const uint8_t* __restrict__ deltas, size_t m,
                                                         const uint8_t* __restrict__ sums,
                                                         uint8_t* __restrict__ dst)
{
    using T = typename UIntOf<W>::T;
    __shared__ T lds[kThreads / 64];
    const T* d = reinterpret_cast<const T*>(deltas);
    T* o = reinterpret_cast<T*>(dst);
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanPerThread;
    T vals[kScanPerThread];
    T acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        size_t i = base + k;
        vals[k] = i < m ? d[i] : (T)0;
        acc = (T)(acc + vals[k]);
    }
    T total;
    T ex = block_exclusive_scan<T>(acc, lds, total);
    T run = (T)(reinterpret_cast<const T*>(sums)[blockIdx.x] + ex);
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        size_t i = base + k;
        run = (T)(run + vals[k]);
        if (i < m) o[i + 1] = run;
    }
}

template <int W>
__global__ void k_store_first(const uint8_t* __restrict__ first, uint8_t* __restrict__ dst)
{
    using T = typename UIntOf<W>::T;
    if (threadIdx.x == 0 && blockIdx.x == 0)
        reinterpret_cast<T*>(dst)[0] = *reinterpret_cast<const T*>(first);
}

// ---------------------------------------------------------------- zigzag
// Reference: src/openzl/codecs/zigzag/{encode,decode}_zigzag_kernel.c.
template <int W, bool ENC>
__global__ void __launch_bounds__(kThreads) k_zigzag(const uint8_t* __restrict__ src, size_t n,
                                                     uint8_t* __restrict__ dst)
{
    using T = typename UIntOf<W>::T;
    constexpr int V = 16 / W;
    const T* s = reinterpret_cast<const T*>(src);
    T* o = reinterpret_cast<T*>(dst);
    size_t nVec = n / V;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    auto f = [](T x) -> T {
        if (ENC) {
            T sign = (T)(0 - (x >> (8 * W - 1)));
            return (T)((T)(x << 1) ^ sign);
        }
        return (T)((x >> 1) ^ (T)(0 - (x & 1)));
    };
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nVec; v += stride) {
        union { uint4 q; T t[V]; } a;
        a.q = *reinterpret_cast<const uint4*>(s + v * V);
#pragma unroll
        for (int k = 0; k < V; k++) a.t[k] = f(a.t[k]);
        *reinterpret_cast<uint4*>(o + v * V) = a.q;
    }
    for (size_t i = nVec * V + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) o[i] = f(s[i]);
}

// ---------------------------------------------------------------- transpose
// Reference: src/openzl/codecs/transpose/encode_transpose_kernel.c:472-515.
// Each lane owns 16 consecutive elements: it loads their 16*W bytes with
// W 128-bit loads and writes one 16-byte piece to each of the W planes, so
// every plane store is a fully coalesced 128-bit store. Byte gathering is
// done in registers (no LDS, no bank conflicts).
template <int W>
__global__ void __launch_bounds__(kThreads) k_transpose_split(const uint8_t* __restrict__ src, size_t n,
                                                              uint8_t* __restrict__ planes)
{
    size_t nVec = n / 16;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nVec; v += stride) {
        union { uint4 q[W]; uint8_t b[16 * W]; } in;
#pragma unroll
        for (int k = 0; k < W; k++) in.q[k] = reinterpret_cast<const uint4*>(src + v * 16 * W)[k];
#pragma unroll
        for (int p = 0; p < W; p++) {
            union { uint4 q; uint8_t b[16]; } out;
#pragma unroll
            for (int e = 0; e < 16; e++) out.b[e] = in.b[e * W + p];
            *reinterpret_cast<uint4*>(planes + (size_t)p * n + v * 16) = out.q;
        }
    }
    for (size_t i = nVec * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
#pragma unroll
        for (int p = 0; p < W; p++) planes[(size_t)p * n + i] = src[i * W + p];
}

// Reference: src/openzl/codecs/transpose/decode_transpose_kernel.c:290.
template <int W>
__global__ void __launch_bounds__(kThreads) k_transpose_join(const uint8_t* __restrict__ planes, size_t n,
                                                             uint8_t* __restrict__ dst)
{
    size_t nVec = n / 16;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nVec; v += stride) {
        union { uint4 q[W]; uint8_t b[16 * W]; } out;
#pragma unroll
        for (int p = 0; p < W; p++) {
            union { uint4 q; uint8_t b[16]; } in;
            in.q = *reinterpret_cast<const uint4*>(planes + (size_t)p * n + v * 16);
#pragma unroll
            for (int e = 0; e < 16; e++) out.b[e * W + p] = in.b[e];
        }
#pragma unroll
        for (int k = 0; k < W; k++) reinterpret_cast<uint4*>(dst + v * 16 * W)[k] = out.q[k];
    }
    for (size_t i = nVec * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
#pragma unroll
        for (int p = 0; p < W; p++) dst[i * W + p] = planes[(size_t)p * n + i];
}

// the same from W planes at their own addresses (16-byte aligned)
struct PlanePtrs {
    const uint8_t* p[8];
};
template <int W>
__global__ void __launch_bounds__(kThreads) k_transpose_join_ptrs(PlanePtrs pp, size_t n, uint8_t* __restrict__ dst)
{
    size_t nVec = n / 16;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < nVec; v += stride) {
        union { uint4 q[W]; uint8_t b[16 * W]; } out;
#pragma unroll
        for (int p = 0; p < W; p++) {
            union { uint4 q; uint8_t b[16]; } in;
            in.q = *reinterpret_cast<const uint4*>(pp.p[p] + v * 16);
#pragma unroll
            for (int e = 0; e < 16; e++) out.b[e * W + p] = in.b[e];
        }
#pragma unroll
        for (int k = 0; k < W; k++) reinterpret_cast<uint4*>(dst + v * 16 * W)[k] = out.q[k];
    }
    for (size_t i = nVec * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
#pragma unroll
        for (int p = 0; p < W; p++) dst[i * W + p] = pp.p[p][i];
}

// planes at any address, any W <= 8: byte per thread
__global__ void __launch_bounds__(kThreads) k_transpose_join_ptrs_any(PlanePtrs pp, size_t n, int w,
                                                                      uint8_t* __restrict__ dst)
{
    size_t total = n * (size_t)w;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
        size_t i = j / w, p = j - i * w;
        dst[j] = pp.p[p][i];
    }
}

// generic element width (any W >= 1): byte per thread
__global__ void __launch_bounds__(kThreads) k_transpose_split_any(const uint8_t* __restrict__ src, size_t n, int w,
                                                                  uint8_t* __restrict__ planes)
{
    size_t total = n * (size_t)w;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
        size_t p = j / n, i = j - p * n;
        planes[j] = src[i * w + p];
    }
}
__global__ void __launch_bounds__(kThreads) k_transpose_join_any(const uint8_t* __restrict__ planes, size_t n, int w,
                                                                 uint8_t* __restrict__ dst)
{
    size_t total = n * (size_t)w;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x; j < total; j += stride) {
        size_t i = j / w, p = j - i * w;
        dst[j] = planes[p * n + i];
    }
}

// ---------------------------------------------------------------- bitpack
// Reference: src/openzl/codecs/bitpack/common_bitpack_kernel.c:131-283 and
// encode_bitpack_binding.c:51-76. Max reduction -> nbBits on the host; the
// pack writes output 32-bit word k from the elements overlapping bits
// [32k, 32k+32), so writes are coalesced and no atomics are needed.
template <int W>
__global__ void __launch_bounds__(kThreads) k_max(const uint8_t* __restrict__ src, size_t n,
                                                  unsigned long long* __restrict__ mx)
{
    using T = typename UIntOf<W>::T;
    const T* s = reinterpret_cast<const T*>(src);
    uint64_t m = 0;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t v = s[i];
        m = v > m ? v : m;
    }
    for (int d = 32; d > 0; d >>= 1) {
        uint64_t o = __shfl_xor(m, d, 64);
        m = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(mx, (unsigned long long)m);
}

template <int W>
__global__ void __launch_bounds__(kThreads) k_bitpack_encode(const uint8_t* __restrict__ src, size_t n,
                                                             int nbBits, uint8_t* __restrict__ dst)
{
    using T = typename UIntOf<W>::T;
    const T* s = reinterpret_cast<const T*>(src);
    size_t totalBits = n * (size_t)nbBits;
    size_t nBytes = (totalBits + 7) / 8;
    size_t nWords = nBytes / 4;  // full 32-bit words
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k <= nWords; k += stride) {
        size_t bit0 = k * 32;
        if (bit0 >= totalBits) break;
        size_t bit1 = bit0 + 32 < totalBits ? bit0 + 32 : totalBits;
        size_t e0 = bit0 / nbBits, e1 = (bit1 - 1) / nbBits;
        uint32_t word = 0;
        for (size_t e = e0; e <= e1; e++) {
            uint64_t v = (uint64_t)s[e];
            long long off = (long long)(e * nbBits) - (long long)bit0;  // may be negative
            uint32_t part;
            if (off >= 0) part = (uint32_t)(v << off);
            else part = (uint32_t)(v >> (-off));
            word |= part;
        }
        if (k < nWords) {
            // dst is a scratch buffer aligned to 4 bytes
            reinterpret_cast<uint32_t*>(dst)[k] = word;
        } else {
            for (size_t b = 0; b < nBytes - nWords * 4; b++) dst[nWords * 4 + b] = (uint8_t)(word >> (8 * b));
        }
    }
}

// Unpack: element i occupies bits [i*nbBits, (i+1)*nbBits). Byte loads keep
// the kernel correct for streams at any byte offset inside a frame.
template <int W>
__global__ void __launch_bounds__(kThreads) k_bitpack_decode(const uint8_t* __restrict__ src, size_t n,
                                                             int nbBits, uint8_t* __restrict__ dst)
{
    using T = typename UIntOf<W>::T;
    T* o = reinterpret_cast<T*>(dst);
    size_t stride = (size_t)gridDim.x * blockDim.x;
    uint64_t mask = nbBits == 64 ? ~0ull : ((1ull << nbBits) - 1);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        size_t bit = i * (size_t)nbBits;
        size_t byte0 = bit >> 3;
        int sh = (int)(bit & 7);
        int nbytes = (sh + nbBits + 7) >> 3;  // <= 9
        uint64_t lo = 0;
        uint8_t hi = 0;
        for (int b = 0; b < nbytes; b++) {
            uint8_t x = src[byte0 + b];
            if (b < 8) lo |= (uint64_t)x << (8 * b);
            else hi = x;
        }
        uint64_t v = lo >> sh;
        if (sh && nbytes == 9) v |= (uint64_t)hi << (64 - sh);
        o[i] = (T)(v & mask);
    }
}

}  // namespace

// ====================================================================== C ABI
#define DISPATCH_W(w, FN, ...)                                     \
    switch (w) {                                                   \
    case 1: FN<1>(__VA_ARGS__); break;                             \
    case 2: FN<2>(__VA_ARGS__); break;                             \
    case 4: FN<4>(__VA_ARGS__); break;                             \
    case 8: FN<8>(__VA_ARGS__); break;                             \
    default: return OZL_ERR_PARAMETER_INVALID;                     \
    }

template <int W>
static void launch_delta_encode(const void* src, size_t n, void* deltas, void* first, hipStream_t st)
{
    (void)hipMemcpyAsync(first, src, W, hipMemcpyDeviceToDevice, st);
    if (n > 1)
        hipLaunchKernelGGL(k_delta_encode<W>, dim3(ozl_blocks((n - 1) / (16 / W) + 1, kThreads, 8192)), dim3(kThreads),
                           0, st, (const uint8_t*)src, n, (uint8_t*)deltas);
}

template <int W>
static void launch_delta_decode(const void* first, const void* deltas, size_t n, void* dst, void* scratch,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_store_first<W>, dim3(1), dim3(64), 0, st, (const uint8_t*)first, (uint8_t*)dst);
    size_t m = n - 1;
    if (m == 0) return;
    size_t tiles = (m + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile_sums<W>, dim3((unsigned)tiles), dim3(kThreads), 0, st, (const uint8_t*)deltas, m,
                       (uint8_t*)scratch);
    hipLaunchKernelGGL(k_scan_sums<W>, dim3(1), dim3(kThreads), 0, st, (uint8_t*)scratch, tiles,
                       (const uint8_t*)first);
    hipLaunchKernelGGL(k_scan_apply<W>, dim3((unsigned)tiles), dim3(kThreads), 0, st, (const uint8_t*)deltas, m,
                       (const uint8_t*)scratch, (uint8_t*)dst);
}

template <int W, bool ENC>
static void launch_zigzag(const void* src, size_t n, void* dst, hipStream_t st)
{
    hipLaunchKernelGGL((k_zigzag<W, ENC>), dim3(ozl_blocks(n / (16 / W) + 1, kThreads, 8192)), dim3(kThreads), 0, st,
                       (const uint8_t*)src, n, (uint8_t*)dst);
}
template <int W> static void launch_zigzag_enc(const void* s, size_t n, void* d, hipStream_t st) { launch_zigzag<W, true>(s, n, d, st); }
template <int W> static void launch_zigzag_dec(const void* s, size_t n, void* d, hipStream_t st) { launch_zigzag<W, false>(s, n, d, st); }

template <int W>
static void launch_max(const void* src, size_t n, unsigned long long* mx, hipStream_t st)
{
    hipLaunchKernelGGL(k_max<W>, dim3(ozl_blocks(n, kThreads * 8, 4096)), dim3(kThreads), 0, st, (const uint8_t*)src,
                       n, mx);
}
template <int W>
static void launch_bitpack_encode(const void* src, size_t n, int nbBits, void* dst, hipStream_t st)
{
    size_t words = (n * (size_t)nbBits + 31) / 32;
    hipLaunchKernelGGL(k_bitpack_encode<W>, dim3(ozl_blocks(words, kThreads, 8192)), dim3(kThreads), 0, st,
                       (const uint8_t*)src, n, nbBits, (uint8_t*)dst);
}
template <int W>
static void launch_bitpack_decode(const void* src, size_t n, int nbBits, void* dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_bitpack_decode<W>, dim3(ozl_blocks(n, kThreads, 8192)), dim3(kThreads), 0, st,
                       (const uint8_t*)src, n, nbBits, (uint8_t*)dst);
}

extern "C" {

// Reference: EI_delta_int, src/openzl/codecs/delta/encode_delta_binding.c:11-45.
int ozl_delta_encode(const void* src, size_t n, int width, void* deltas, void* first, void* stream)
{
    if (n == 0) return OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    DISPATCH_W(width, launch_delta_encode, src, n, deltas, first, st);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

size_t ozl_delta_decode_scratch_size(size_t n, int width)
{
    return ((n + kScanTile - 1) / kScanTile + 1) * (size_t)(width < 8 ? 8 : width);
}

// Reference: DI_delta_int, src/openzl/codecs/delta/decode_delta_binding.c:10-62.
int ozl_delta_decode(const void* first, const void* deltas, size_t n, int width, void* dst, void* scratch,
                     void* stream)
{
    if (n == 0) return OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    DISPATCH_W(width, launch_delta_decode, first, deltas, n, dst, scratch, st);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Reference: EI_zigzag_num / DI_zigzag_num (src/openzl/codecs/zigzag/).
int ozl_zigzag_encode(const void* src, size_t n, int width, void* dst, void* stream)
{
    if (n == 0) return OZL_OK;
    DISPATCH_W(width, launch_zigzag_enc, src, n, dst, (hipStream_t)stream);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
int ozl_zigzag_decode(const void* src, size_t n, int width, void* dst, void* stream)
{
    if (n == 0) return OZL_OK;
    DISPATCH_W(width, launch_zigzag_dec, src, n, dst, (hipStream_t)stream);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Reference: EI_transpose_split (encode_transpose_binding.c:42-79); planes are
// written back to back, plane p at planes + p*n.
int ozl_transpose_split(const void* src, size_t n, int width, void* planes, void* stream)
{
    if (n == 0 || width <= 0) return width <= 0 ? OZL_ERR_PARAMETER_INVALID : OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    unsigned g = ozl_blocks(n / 16 + 1, kThreads, 8192);
    const uint8_t* s = (const uint8_t*)src;
    uint8_t* d = (uint8_t*)planes;
    switch (width) {
    case 1: OZL_HIP_CHECK(hipMemcpyAsync(planes, src, n, hipMemcpyDeviceToDevice, st)); break;
    case 2: hipLaunchKernelGGL(k_transpose_split<2>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    case 4: hipLaunchKernelGGL(k_transpose_split<4>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    case 8: hipLaunchKernelGGL(k_transpose_split<8>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    default:
        hipLaunchKernelGGL(k_transpose_split_any, dim3(ozl_blocks(n * width, kThreads, 8192)), dim3(kThreads), 0, st,
                           s, n, width, d);
    }
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Reference: DI_transpose_split (decode_transpose_binding.c:37-74).
int ozl_transpose_join(const void* planes, size_t n, int width, void* dst, void* stream)
{
    if (n == 0 || width <= 0) return width <= 0 ? OZL_ERR_PARAMETER_INVALID : OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    unsigned g = ozl_blocks(n / 16 + 1, kThreads, 8192);
    const uint8_t* s = (const uint8_t*)planes;
    uint8_t* d = (uint8_t*)dst;
    switch (width) {
    case 1: OZL_HIP_CHECK(hipMemcpyAsync(dst, planes, n, hipMemcpyDeviceToDevice, st)); break;
    case 2: hipLaunchKernelGGL(k_transpose_join<2>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    case 4: hipLaunchKernelGGL(k_transpose_join<4>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    case 8: hipLaunchKernelGGL(k_transpose_join<8>, dim3(g), dim3(kThreads), 0, st, s, n, d); break;
    default:
        hipLaunchKernelGGL(k_transpose_join_any, dim3(ozl_blocks(n * width, kThreads, 8192)), dim3(kThreads), 0, st,
                           s, n, width, d);
    }
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// transpose_split decode from `width` (at most 8) planes of n bytes each at
// the device addresses planes[0..width) (a host array), into dst.
int ozl_transpose_join_planes(const void* const* planes, size_t n, int width, void* dst, void* stream)
{
    if (width <= 0 || width > 8) return OZL_ERR_PARAMETER_INVALID;
    if (n == 0) return OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    PlanePtrs pp = {};
    bool aligned = true;
    for (int p = 0; p < width; p++) {
        pp.p[p] = (const uint8_t*)planes[p];
        aligned = aligned && ((uintptr_t)planes[p] & 15) == 0;
    }
    aligned = aligned && ((uintptr_t)dst & 15) == 0;
    unsigned g = ozl_blocks(n / 16 + 1, kThreads, 8192);
    uint8_t* d = (uint8_t*)dst;
    if (width == 1) {
        OZL_HIP_CHECK(hipMemcpyAsync(dst, planes[0], n, hipMemcpyDeviceToDevice, st));
    } else if (aligned && width == 2) {
        hipLaunchKernelGGL(k_transpose_join_ptrs<2>, dim3(g), dim3(kThreads), 0, st, pp, n, d);
    } else if (aligned && width == 4) {
        hipLaunchKernelGGL(k_transpose_join_ptrs<4>, dim3(g), dim3(kThreads), 0, st, pp, n, d);
    } else if (aligned && width == 8) {
        hipLaunchKernelGGL(k_transpose_join_ptrs<8>, dim3(g), dim3(kThreads), 0, st, pp, n, d);
    } else {
        hipLaunchKernelGGL(k_transpose_join_ptrs_any, dim3(ozl_blocks(n * width, kThreads, 8192)), dim3(kThreads), 0,
                           st, pp, n, width, d);
    }
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Max element (for computeNbBits, encode_bitpack_binding.c:51-76) into a
// device u64 the caller zeroed. Asynchronous.
int ozl_max_async(const void* src, size_t n, int width, unsigned long long* dmax, void* stream)
{
    if (n == 0) return OZL_OK;
    DISPATCH_W(width, launch_max, src, n, dmax, (hipStream_t)stream);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Reference: ZS_bitpackEncode (common_bitpack_kernel.c:682-706). dst must be
// 4-byte aligned and hold ceil(n*nbBits/8) bytes.
int ozl_bitpack_encode(const void* src, size_t n, int width, int nbBits, void* dst, void* stream)
{
    if (nbBits < 1 || nbBits > 8 * width) return OZL_ERR_PARAMETER_INVALID;
    if (n == 0) return OZL_OK;
    DISPATCH_W(width, launch_bitpack_encode, src, n, nbBits, dst, (hipStream_t)stream);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Reference: ZS_bitpackDecode (common_bitpack_kernel.c:1371-1392).
int ozl_bitpack_decode(const void* src, size_t n, int width, int nbBits, void* dst, void* stream)
{
    if (nbBits < 1 || nbBits > 8 * width) return OZL_ERR_PARAMETER_INVALID;
    if (n == 0) return OZL_OK;
    DISPATCH_W(width, launch_bitpack_decode, src, n, nbBits, dst, (hipStream_t)stream);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"
