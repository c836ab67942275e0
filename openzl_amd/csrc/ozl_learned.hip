// Learned graph selector on gfx950: per-chunk features and a 40 -> 64 -> 64
// -> 8 MLP (bf16 weights and activations, fp32 accumulation on the MFMA
// units) that picks one of the eight candidate graphs of ozl_select.hip for
// each chunk. The reference's counterpart is its GBT selector
// (src/openzl/compress/selectors/ml/mlselector.c, features.c:158-242): a
// feature vector per input, a model, a label that names a graph. The
// features and the network here are this project's own.
//
// Features of a chunk of n elements of width w (element x_i, delta
// d_i = x_{i+1} - x_i mod 2^(8w), zigzag z_i; bits(v) = 0 for v = 0, else
// 1 + floor(log2 v)):
//   f[0..31]  32-bin histogram of the chunk's bytes (byte >> 3), as fractions
//   f[32]     byte entropy / 8 (256 bins, bits per byte)
//   f[33]     log2(w) / 3
//   f[34]     fraction of zero deltas
//   f[35]     mean bits(z_i) / 64
//   f[36]     max bits(z_i) / 64
//   f[37]     max bits(x_i) / 64
//   f[38]     mean over byte planes of the deltas' plane entropy / 8
//   f[39]     mean over byte planes of the elements' plane entropy / 8
// (delta features are 0 when n < 2).
//
// MLP: h1 = relu(W1 f + b1), h2 = relu(W2 h1 + b2), logits = W3 h2 + b3,
// choice = argmax over the 8 logits (first maximum). One wavefront computes
// 16 chunks with v_mfma_f32_16x16x32_bf16: rows are chunks, K the inputs,
// columns the neurons; activations pass between layers through LDS in bf16.
// Parameter block (bytes): W1 bf16[64][64] (inputs 40..63 zero), W2
// bf16[64][64], W3 bf16[16][64] (rows 8..15 zero), b1 f32[64], b2 f32[64],
// b3 f32[16].
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_profile.h"

namespace {

constexpr int kF = 40;
constexpr int kT = 256;

__device__ inline uint64_t ld_el(const uint8_t* p, int w)
{
    switch (w) {
    case 1: return *p;
    case 2: return *reinterpret_cast<const uint16_t*>(p);
    case 4: return *reinterpret_cast<const uint32_t*>(p);
    default: return *reinterpret_cast<const uint64_t*>(p);
    }
}
__device__ inline int bits64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

template <typename T>
__device__ T wg_sum(T v, T* lds)
{
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    T t = 0;
    for (int k = 0; k < kT / 64; k++) t += lds[k];
    return t;
}
template <typename T>
__device__ T wg_max(T v, T* lds)
{
    for (int d = 32; d >= 1; d >>= 1) {
        const T o = __shfl_xor(v, d, 64);
        v = o > v ? o : v;
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
    __syncthreads();
    T t = 0;
    for (int k = 0; k < kT / 64; k++) t = lds[k] > t ? lds[k] : t;
    return t;
}

// c * log2(c) for one histogram bin
__device__ inline float clog(uint32_t c) { return c ? (float)c * __log2f((float)c) : 0.f; }

// one workgroup per chunk
__global__ void __launch_bounds__(kT) k_features(const uint8_t* __restrict__ src, const uint64_t* __restrict__ spans,
                                                 const uint8_t* __restrict__ widths, uint32_t K,
                                                 float* __restrict__ feats)
{
    __shared__ uint32_t hx[8][256];
    __shared__ uint32_t hd[8][256];
    __shared__ float lf[kT / 64];
    __shared__ uint32_t lu[kT / 64];
    __shared__ unsigned long long lz[kT / 64];
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    const int w = widths[c];
    const uint64_t bytes = spans[2 * c + 1];
    const uint8_t* x = src + spans[2 * c];
    float* f = feats + (size_t)c * kF;
    const bool ok = (w == 1 || w == 2 || w == 4 || w == 8) && bytes >= (uint64_t)w && bytes % w == 0;
    if (!ok) {
        if (threadIdx.x < kF) f[threadIdx.x] = 0.f;
        return;
    }
    const uint32_t n = (uint32_t)(bytes / w), m = n - 1;
    for (int i = threadIdx.x; i < 8 * 256; i += kT) {
        (&hx[0][0])[i] = 0;
        (&hd[0][0])[i] = 0;
    }
    __syncthreads();
    uint32_t zeros = 0;
    unsigned long long bz = 0;
    uint32_t mz = 0, mx = 0;
    const uint64_t msk = w == 8 ? ~0ull : ((1ull << (8 * w)) - 1);
    for (uint32_t i = threadIdx.x; i < n; i += kT) {
        const uint64_t v = ld_el(x + (size_t)i * w, w);
        for (int k = 0; k < w; k++) atomicAdd(&hx[k][(v >> (8 * k)) & 255], 1u);
        const uint32_t bv = (uint32_t)bits64(v);
        mx = bv > mx ? bv : mx;
        if (i < m) {
            const uint64_t d = (ld_el(x + (size_t)(i + 1) * w, w) - v) & msk;
            const int sb = 8 * w - 1;
            const uint64_t s = (uint64_t)((int64_t)(d << (63 - sb)) >> 63);
            const uint64_t z = ((d << 1) ^ s) & msk;
            zeros += d == 0;
            const uint32_t b = (uint32_t)bits64(z);
            bz += b;
            mz = b > mz ? b : mz;
            for (int k = 0; k < w; k++) atomicAdd(&hd[k][(d >> (8 * k)) & 255], 1u);
        }
    }
    __syncthreads();
    // thread v owns byte value v
    const int v = threadIdx.x;
    uint32_t all = 0;
    for (int k = 0; k < w; k++) all += hx[k][v];
    float sAll = clog(all), sX = 0.f, sD = 0.f;
    for (int k = 0; k < w; k++) {
        sX += clog(hx[k][v]);
        sD += clog(hd[k][v]);
    }
    // 32-bin histogram: 8 consecutive values per bin
    uint32_t bin = all;
    for (int d = 1; d < 8; d <<= 1) bin += __shfl_xor(bin, d, 64);
    if ((v & 7) == 0) f[v >> 3] = (float)bin / (float)bytes;
    sAll = wg_sum<float>(sAll, lf);
    sX = wg_sum<float>(sX, lf);
    sD = wg_sum<float>(sD, lf);
    zeros = wg_sum<uint32_t>(zeros, lu);
    bz = wg_sum<unsigned long long>(bz, lz);
    mz = wg_max<uint32_t>(mz, lu);
    mx = wg_max<uint32_t>(mx, lu);
    if (threadIdx.x == 0) {
        // entropy of N samples = log2 N - (sum c log2 c) / N
        const float B = (float)bytes, N = (float)n, M = (float)m;
        f[32] = (__log2f(B) - sAll / B) / 8.f;
        f[33] = (float)ozl_next_pow2((uint64_t)w) / 3.f;
        f[34] = m ? (float)zeros / M : 0.f;
        f[35] = m ? (float)bz / M / 64.f : 0.f;
        f[36] = (float)mz / 64.f;
        f[37] = (float)mx / 64.f;
        // per plane: log2(count) - sum/count, averaged over the w planes
        f[38] = m ? (__log2f(M) - sD / ((float)w * M)) / 8.f : 0.f;
        f[39] = (__log2f(N) - sX / ((float)w * N)) / 8.f;
    }
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ inline uint16_t to_bf16(float x)
{
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// D[16 x 16] += A[16 x 32] . B[32 x 16]; lane l: A row l%16 / B column l%16,
// k = 8*(l/16) .. +8; D rows 4*(l/16) .. +4 of column l%16
__device__ inline f32x4 mfma(u16x8 a, u16x8 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// one layer: act (LDS, bf16 [16][64]) x W^T (global bf16 [nOut][64]) + b,
// relu when `relu`; result to LDS (bf16) or logits
template <int NB>
__device__ inline void layer(const uint16_t* act, const uint16_t* __restrict__ W, const float* __restrict__ b,
                             f32x4 out[NB])
{
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int nb = 0; nb < NB; nb++) {
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const int k0 = 32 * ks + 8 * kq;
            const u16x8 a = *reinterpret_cast<const u16x8*>(act + r * 64 + k0);
            const u16x8 wv = *reinterpret_cast<const u16x8*>(W + (size_t)(16 * nb + r) * 64 + k0);
            acc = mfma(a, wv, acc);
        }
        const float bb = b[16 * nb + r];
        for (int i = 0; i < 4; i++) acc[i] += bb;
        out[nb] = acc;
    }
}

__global__ void __launch_bounds__(64) k_mlp(const float* __restrict__ feats, uint32_t K,
                                            const uint8_t* __restrict__ params, uint8_t* __restrict__ choice,
                                            float* __restrict__ logits)
{
    __shared__ __attribute__((aligned(16))) uint16_t act[16 * 64];
    __shared__ float lg[16][16];
    const uint16_t* W1 = reinterpret_cast<const uint16_t*>(params);
    const uint16_t* W2 = W1 + 64 * 64;
    const uint16_t* W3 = W2 + 64 * 64;
    const float* b1 = reinterpret_cast<const float*>(W3 + 16 * 64);
    const float* b2 = b1 + 64;
    const float* b3 = b2 + 64;
    const int lane = threadIdx.x, r = lane & 15, kq = lane >> 4;
    const uint32_t row0 = blockIdx.x * 16;
    // features -> bf16 rows (inputs 40..63 zero, rows past K zero)
    for (int i = lane; i < 16 * 64; i += 64) {
        const int rr = i / 64, k = i % 64;
        const uint32_t c = row0 + rr;
        act[i] = (c < K && k < kF) ? to_bf16(feats[(size_t)c * kF + k]) : 0;
    }
    __syncthreads();
    f32x4 h[4];
    layer<4>(act, W1, b1, h);
    __syncthreads();
    for (int nb = 0; nb < 4; nb++)
        for (int i = 0; i < 4; i++) act[(4 * kq + i) * 64 + 16 * nb + r] = to_bf16(fmaxf(h[nb][i], 0.f));
    __syncthreads();
    layer<4>(act, W2, b2, h);
    __syncthreads();
    for (int nb = 0; nb < 4; nb++)
        for (int i = 0; i < 4; i++) act[(4 * kq + i) * 64 + 16 * nb + r] = to_bf16(fmaxf(h[nb][i], 0.f));
    __syncthreads();
    f32x4 o[1];
    layer<1>(act, W3, b3, o);
    for (int i = 0; i < 4; i++) lg[4 * kq + i][r] = o[0][i];
    __syncthreads();
    if (lane < 16) {
        const uint32_t c = row0 + lane;
        if (c < K) {
            int best = 0;
            float bv = lg[lane][0];
            for (int k = 1; k < 8; k++)
                if (lg[lane][k] > bv) { bv = lg[lane][k]; best = k; }
            choice[c] = (uint8_t)best;
            if (logits)
                for (int k = 0; k < 8; k++) logits[(size_t)c * 8 + k] = lg[lane][k];
        }
    }
}

}  // namespace

extern "C" {

// Parameter block bytes of the selector MLP.
size_t ozl_select_mlp_params_bytes(void) { return (64 * 64 + 64 * 64 + 16 * 64) * 2 + (64 + 64 + 16) * 4; }

// 40 features per chunk (layout in ozl_learned.hip's header) into feats
// (K x 40 floats). Device pointers.
int ozl_select_features(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t K, float* feats,
                        void* stream)
{
    if (K == 0) return OZL_OK;
    hipStream_t st = (hipStream_t)stream;
    OZL_LAUNCH("k_features", k_features, dim3(K), dim3(kT), 0, st, (const uint8_t*)src, spans, widths, K, feats);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// The MLP over K feature rows: choice[c] in 0..7, logits (K x 8, may be null).
// Device pointers.
int ozl_select_mlp(const float* feats, uint32_t K, const void* params, uint8_t* choice, float* logits, void* stream)
{
    if (K == 0) return OZL_OK;
    if (!params || !choice) return OZL_ERR_PARAMETER_INVALID;
    hipStream_t st = (hipStream_t)stream;
    OZL_LAUNCH("k_mlp", k_mlp, dim3((K + 15) / 16), dim3(64), 0, st, feats, K, (const uint8_t*)params, choice,
               logits);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"
