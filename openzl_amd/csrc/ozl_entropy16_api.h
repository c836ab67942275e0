// Internal interface of the 2-byte entropy kernels (ozl_entropy16.hip):
// ZL_GRAPH_HUFFMAN on struct / numeric inputs of 2-byte elements
// (src/openzl/codecs/entropy/encode_entropy_binding.c:612-682, 775-797) and
// the decoders of the nodes it creates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// One chunk of 2-byte elements (device address, element count).
struct ozl_c16 {
    uint64_t src;
    uint32_t n;
    uint32_t pad;
};

// Histogram statistics of a chunk: largest value, number of distinct values,
// the count of the largest value, and the int64 sum of the reference's
// entropy terms (data_stats.c:46-73).
struct ozl_s16 {
    uint32_t maxSym, card, maxCount, pad;
    int64_t termSum;
};

// Histograms (65,536 u32 per chunk, cleared by the call) and statistics of
// C chunks of at most maxN elements.
int ozl_e16_stats(const ozl_c16* dChunks, uint32_t C, uint32_t maxN, uint32_t* hist, ozl_s16* dStats,
                  hipStream_t st);

// Tokenize (alphabet in order of first occurrence, 1-byte indices) of the
// chunks of fewer than 256 distinct values: chunk job j reads chunk src with
// n elements and its histogram, writes its alphabet (u16) to alpha and its
// indices to idx. The histogram is overwritten.
struct ozl_tok16 {
    uint64_t src, hist, alpha, idx;
    uint32_t n, card;
};
int ozl_e16_tokenize(const ozl_tok16* dJobs, uint32_t T, hipStream_t st);

// 16-bit Huffman (huffman_struct_v2): table, weights (maxSym + 1 bytes) and
// streams (into out, cap = 2n + 32 bytes, zeroed by the caller) of a chunk.
struct ozl_h16job {
    uint64_t src, hist, weights, out, ws;
    uint32_t n, maxSym, card, x4;
};
struct ozl_h16res {
    uint32_t tableLog, size, status, pad;
};
// Workspace bytes of one job of `card` distinct values and largest value maxSym.
size_t ozl_e16_huf_ws(uint32_t card, uint32_t maxSym);
int ozl_e16_huf_encode(const ozl_h16job* dJobs, uint32_t H, ozl_h16res* dRes, hipStream_t st);

// Decoders. huffman_struct_v2: weights (nbWeights bytes), bits, n 2-byte
// elements out; ws holds nbWeights u16. status[i] = 0 or a ZL error code.
struct ozl_h16dec {
    uint64_t weights, bits, out, ws;
    uint32_t nbWeights, bitsLen, n, x4;
};
int ozl_e16_huf_decode(const ozl_h16dec* dItems, uint32_t D, int* dStatus, hipStream_t st);

// tokenize: out[i] = alpha[idx[i]] (element width w, index width iw);
// an index past the alphabet sets *status to corruption.
int ozl_tokenize_decode(const void* alpha, uint32_t card, int w, const void* idx, int iw, size_t n, void* out,
                        int* status, hipStream_t st);
// n elements of width w, each equal to the w bytes at val.
int ozl_fill_fixed(void* dst, const void* val, int w, size_t n, hipStream_t st);
