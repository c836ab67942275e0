/*
 * ozl_oracle.h -- CPU restatement of OpenZL's codec hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against; it is loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, never by the product. Every function states the
 * reference file:line whose behaviour it restates. All arithmetic is integer
 * except the two double-precision decision formulas the reference itself uses
 * (src/openzl/shared/data_stats.c:46-73, :288-299).
 */
#ifndef OZL_ORACLE_H
#define OZL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- small helpers ---- */
size_t orc_varint_write(uint64_t v, uint8_t* out);
int orc_next_pow2(uint64_t x);

/* ---- numeric transforms (widths 1,2,4,8) ---- */
void orc_delta_encode(void* deltas, void* first, const void* src, size_t n, int width);
void orc_delta_decode(void* dst, const void* first, const void* deltas, size_t n, int width);
void orc_zigzag_encode(void* dst, const void* src, size_t n, int width);
void orc_zigzag_decode(void* dst, const void* src, size_t n, int width);
void orc_transpose_split(uint8_t* planes, const uint8_t* src, size_t n, int width);
void orc_transpose_join(uint8_t* dst, const uint8_t* planes, size_t n, int width);

/* ---- bitpack / flatpack / constant ---- */
int orc_bitpack_nbbits(const void* src, size_t n, int width);
size_t orc_bitpack_encode(uint8_t* dst, const void* src, size_t n, int width, int nbBits);
size_t orc_bitpack_header(uint8_t hdr[2], size_t n, int width, int nbBits, size_t packedSize);
size_t orc_bitpack_decode(void* dst, size_t capacity, const uint8_t* src, size_t srcSize,
                          const uint8_t* hdr, size_t hdrSize, int* widthOut);
size_t orc_flatpack_encode(uint8_t* alphabet, uint8_t* packed, size_t* packedSize,
                           const uint8_t* src, size_t n);
size_t orc_flatpack_decode(uint8_t* dst, size_t capacity, const uint8_t* alphabet,
                           size_t alphabetSize, const uint8_t* packed, size_t packedSize);

/* ---- statistics and the per-chunk backend decision ---- */
void orc_histogram(uint32_t count[256], unsigned* maxSym, unsigned* cardinality,
                   const uint8_t* src, size_t n);
double orc_entropy(const uint32_t* count, size_t maxValue, size_t total);
extern const uint16_t orc_log2_table[256];
enum { ORC_BK_STORE = 0, ORC_BK_CONSTANT = 1, ORC_BK_FLATPACK = 2, ORC_BK_BITPACK = 3,
       ORC_BK_HUF = 4, ORC_BK_FSE = 5 };
int orc_entropy_decide(const uint8_t* src, size_t n, int hufMode);

/* ---- FSE (tANS) ---- */
typedef struct {
    unsigned tableLog;
    unsigned maxSym;
    uint16_t stateTable[4096];
    int32_t deltaFindState[256];
    uint32_t deltaNbBits[256];
} orc_fse_ctable;

unsigned orc_fse_optimal_tablelog(unsigned maxTableLog, size_t n, unsigned maxSym, unsigned minus);
int orc_fse_normalize(int16_t* norm, unsigned tableLog, const uint32_t* count, size_t total,
                      unsigned maxSym, int useLowProb);
int orc_fse_build_ctable(orc_fse_ctable* ct, const int16_t* norm, unsigned maxSym, unsigned tableLog);
size_t orc_fse_encode(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                      const orc_fse_ctable* ct, unsigned nbStates);
size_t orc_fse_write_ncount(uint8_t* dst, size_t cap, const int16_t* norm, unsigned maxSym,
                            unsigned tableLog);
long orc_fse_read_ncount(int16_t* norm, unsigned* maxSym, unsigned* tableLog,
                         const uint8_t* src, size_t size);
int orc_fse_decode(uint8_t* dst, size_t n, const uint8_t* src, size_t size,
                   const int16_t* norm, unsigned maxSym, unsigned tableLog, unsigned nbStates);

/* ---- Huffman ---- */
unsigned orc_huf_optimal_tablelog(unsigned maxTableLog, size_t n, unsigned maxSym);
int orc_huf_build_ctable(uint8_t* nbBits, uint16_t* values, const uint32_t* count,
                         unsigned maxSym, unsigned maxNbBits);
size_t orc_huf_encode_1x(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                         const uint8_t* nbBits, const uint16_t* values);
size_t orc_huf_encode_4x(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                         const uint8_t* nbBits, const uint16_t* values);
int orc_huf_decode(uint8_t* dst, size_t n, const uint8_t* src, size_t size,
                   const uint8_t* weights, size_t nbWeights, int x4);

/* ---- 2-byte inputs of ZL_GRAPH_HUFFMAN ---- */
enum { ORC_E2_STORE = 0, ORC_E2_CONSTANT = 1, ORC_E2_BITPACK = 3, ORC_E2_TOKENIZE = 6, ORC_E2_HUF16 = 7 };
void orc_histogram16(uint32_t* count, unsigned* maxSym, unsigned* card, const uint16_t* src, size_t n);
int orc_entropy2_decide(const uint16_t* src, size_t n, double* entropyOut);
size_t orc_tokenize2(uint16_t* alphabet, uint8_t* idx, const uint16_t* src, size_t n);
int orc_huf16_build(uint16_t* code, uint8_t* nbBits, const uint32_t* count, unsigned maxSym);
size_t orc_huf16_encode(uint8_t* dst, size_t cap, const uint16_t* src, size_t n, const uint16_t* code,
                        const uint8_t* nbBits, int x4);
int orc_huf16_decode(uint16_t* dst, size_t n, const uint8_t* src, size_t size, const uint8_t* weights,
                     size_t nbWeights, int x4);

/* ---- XXH3-64 (seed 0, default secret) ---- */
uint64_t orc_xxh3_64(const void* data, size_t len);

#ifdef __cplusplus
}
#endif
#endif
