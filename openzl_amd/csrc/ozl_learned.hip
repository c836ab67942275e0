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
#include "ozl_gmem.h"

namespace {

constexpr int kF = 40;
constexpr int kT = 256;

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

template <int W> struct ElemT;
template <> struct ElemT<1> { using T = uint8_t; };
template <> struct ElemT<2> { using T = uint16_t; };
template <> struct ElemT<4> { using T = uint32_t; };
template <> struct ElemT<8> { using T = uint64_t; };

// Byte-plane histograms: W planes x 256 bins x (32 / W) copies in 32 KB of
// LDS, bin-major (the copy index is the low part of the address, so lanes of
// one wave that hit the same bin mostly hit different banks). Thread t owns
// the elements [16/W * v, 16/W * (v + 1)) of vectors v = t, t + 256, ...;
// equal bytes that follow each other in a plane are added as one count.
constexpr int kHistWords = 8192;

template <int W>
__device__ inline void plane_add(uint32_t* hc, int p, uint32_t b, uint32_t cnt)
{
    constexpr int C = 32 / W;
    atomicAdd(&hc[(p * 256 + (int)b) * C], cnt);
}

template <int W, bool DELTA>
__device__ void plane_hist(const uint8_t* x, uint32_t n, uint32_t* h, uint32_t& zeros, unsigned long long& bz,
                           uint32_t& mb)
{
    using T = typename ElemT<W>::T;
    constexpr int C = 32 / W, E = 16 / W;
    uint32_t* hc = h + (threadIdx.x & (C - 1));
    const uint32_t cnt = DELTA ? n - 1 : n;   // values: elements or deltas
    const uint32_t nv = (cnt + E - 1) / E;
    const bool al8 = ((uintptr_t)x & 7) == 0;
    for (uint32_t v = threadIdx.x; v < nv; v += kT) {
        const uint32_t e0 = v * E;
        T val[E];
        int nvalid = 0;
        if (al8 && e0 + E <= cnt) {
            // a whole vector: two aligned 8-byte loads (and, for deltas, the
            // element after it), the elements taken out of the words
            const uint64_t wa = ozl_gld8(x + (size_t)e0 * W), wb = ozl_gld8(x + (size_t)e0 * W + 8);
            T el[E + 1];
#pragma unroll
            for (int j = 0; j < E; j++) {
                const uint64_t word = j * W < 8 ? wa : wb;
                el[j] = (T)(word >> ((j * W) % 8 * 8));
            }
            if (DELTA) el[E] = (T)ozl_gld_w(x + (size_t)(e0 + E) * W, W);
#pragma unroll
            for (int j = 0; j < E; j++) val[j] = DELTA ? (T)(el[j + 1] - el[j]) : el[j];
            nvalid = E;
        } else {
#pragma unroll
            for (int j = 0; j < E; j++) {
                const uint32_t e = e0 + j;
                const uint32_t ee = e < cnt ? e : e0;
                const T a = (T)ozl_gld_w(x + (size_t)ee * W, W);
                if (DELTA) {
                    const T b = (T)ozl_gld_w(x + (size_t)(ee + 1) * W, W);
                    val[j] = (T)(b - a);
                } else {
                    val[j] = a;
                }
                nvalid += e < cnt;
            }
        }
#pragma unroll
        for (int j = 0; j < E; j++) {
            if (j < nvalid) {
                const uint64_t u = (uint64_t)val[j];
                if (DELTA) {
                    const uint64_t s = (uint64_t)((int64_t)(u << (64 - 8 * W)) >> 63);
                    const uint64_t z = ((u << 1) ^ s) & (W == 8 ? ~0ull : ((1ull << (8 * W)) - 1));
                    zeros += u == 0;
                    const uint32_t b = (uint32_t)bits64(z);
                    bz += b;
                    mb = b > mb ? b : mb;
                } else {
                    const uint32_t b = (uint32_t)bits64(u);
                    mb = b > mb ? b : mb;
                }
            }
        }
#pragma unroll
        for (int p = 0; p < W; p++) {
            uint32_t cur = (uint32_t)((uint64_t)val[0] >> (8 * p)) & 255u, run = 1;
#pragma unroll
            for (int j = 1; j < E; j++) {
                if (j < nvalid) {
                    const uint32_t b = (uint32_t)((uint64_t)val[j] >> (8 * p)) & 255u;
                    if (b == cur) {
                        run++;
                    } else {
                        plane_add<W>(hc, p, cur, run);
                        cur = b;
                        run = 1;
                    }
                }
            }
            plane_add<W>(hc, p, cur, run);
        }
    }
}

// sum of c log2 c over the W planes of h for bin v (and the total count of v)
template <int W>
__device__ inline float plane_clog(const uint32_t* h, int v, uint32_t& all)
{
    constexpr int C = 32 / W;
    float s = 0.f;
    all = 0;
    for (int p = 0; p < W; p++) {
        uint32_t t = 0;
#pragma unroll
        for (int c = 0; c < C; c++) t += h[(p * 256 + v) * C + c];
        s += clog(t);
        all += t;
    }
    return s;
}

template <int W>
__device__ void features_w(const uint8_t* x, uint32_t n, float* f, uint32_t* h, float* lf, uint32_t* lu,
                           unsigned long long* lz)
{
    const uint32_t m = n - 1;
    const uint64_t bytes = (uint64_t)n * W;
    const int v = threadIdx.x;
    uint32_t zeros = 0, mx = 0, mz = 0;
    unsigned long long bz = 0;
    // element planes
    for (int i = threadIdx.x; i < kHistWords; i += kT) h[i] = 0;
    __syncthreads();
    plane_hist<W, false>(x, n, h, zeros, bz, mx);
    __syncthreads();
    uint32_t all;
    float sX = plane_clog<W>(h, v, all);
    float sAll = clog(all);
    uint32_t bin = all;
    for (int d = 1; d < 8; d <<= 1) bin += __shfl_xor(bin, d, 64);
    if ((v & 7) == 0) f[v >> 3] = (float)bin / (float)bytes;
    __syncthreads();
    // delta planes
    float sD = 0.f;
    if (m) {
        for (int i = threadIdx.x; i < kHistWords; i += kT) h[i] = 0;
        __syncthreads();
        plane_hist<W, true>(x, n, h, zeros, bz, mz);
        __syncthreads();
        uint32_t dall;
        sD = plane_clog<W>(h, v, dall);
    }
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
        f[33] = (float)ozl_next_pow2((uint64_t)W) / 3.f;
        f[34] = m ? (float)zeros / M : 0.f;
        f[35] = m ? (float)bz / M / 64.f : 0.f;
        f[36] = (float)mz / 64.f;
        f[37] = (float)mx / 64.f;
        // per plane: log2(count) - sum/count, averaged over the W planes
        f[38] = m ? (__log2f(M) - sD / ((float)W * M)) / 8.f : 0.f;
        f[39] = (__log2f(N) - sX / ((float)W * N)) / 8.f;
    }
}

// one workgroup per chunk
__global__ void __launch_bounds__(kT) k_features(const uint8_t* __restrict__ src, const uint64_t* __restrict__ spans,
                                                 const uint8_t* __restrict__ widths, uint32_t K,
                                                 float* __restrict__ feats)
{
    __shared__ uint32_t h[kHistWords];
    __shared__ float lf[kT / 64];
    __shared__ uint32_t lu[kT / 64];
    __shared__ unsigned long long lz[kT / 64];
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    const int w = widths[c];
    const uint64_t bytes = spans[2 * c + 1];
    const uint8_t* x = src + spans[2 * c];
    float* f = feats + (size_t)c * kF;
    const bool ok = (w == 1 || w == 2 || w == 4 || w == 8) && bytes >= (uint64_t)w && bytes % w == 0 &&
                    bytes < (1ull << 31);
    if (!ok) {
        if (threadIdx.x < kF) f[threadIdx.x] = 0.f;
        return;
    }
    const uint32_t n = (uint32_t)(bytes / w);
    switch (w) {
    case 1: features_w<1>(x, n, f, h, lf, lu, lz); break;
    case 2: features_w<2>(x, n, f, h, lf, lu, lz); break;
    case 4: features_w<4>(x, n, f, h, lf, lu, lz); break;
    default: features_w<8>(x, n, f, h, lf, lu, lz); break;
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
