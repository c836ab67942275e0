/*
 * ozl_oracle.c -- CPU restatement of OpenZL's codec hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see ozl_oracle.h).
 * Plain, single-threaded C written for clarity rather than speed. Each
 * function names the reference file:line it restates (paths relative to the
 * reference repository root). Each is pinned by tests/test_oracle.py against
 * the compiled reference library, against known-answer vectors, or against
 * closed forms; the header of that test file lists which pin covers what.
 */
#include "ozl_oracle.h"

#include <stdlib.h>
#include <string.h>

static int highbit32(uint32_t v) { return 31 - __builtin_clz(v); }
static int highbit64(uint64_t v) { return 63 - __builtin_clzll(v); }

/* src/openzl/shared/bits.h:230 ZL_nextPow2: smallest n with (1<<n) >= x. */
int orc_next_pow2(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }

/* src/openzl/shared/varint.h: LEB128, 7 bits per byte, low group first. */
size_t orc_varint_write(uint64_t v, uint8_t* out)
{
    size_t n = 0;
    while (v >= 0x80) { out[n++] = (uint8_t)(v | 0x80); v >>= 7; }
    out[n++] = (uint8_t)v;
    return n;
}

static uint64_t rdw(const void* p, size_t i, int w)
{
    const uint8_t* b = (const uint8_t*)p + i * (size_t)w;
    uint64_t v = 0;
    for (int k = 0; k < w; k++) v |= (uint64_t)b[k] << (8 * k);
    return v;
}
static void wrw(void* p, size_t i, int w, uint64_t v)
{
    uint8_t* b = (uint8_t*)p + i * (size_t)w;
    for (int k = 0; k < w; k++) b[k] = (uint8_t)(v >> (8 * k));
}
static uint64_t wmask(int w) { return w == 8 ? ~0ull : ((1ull << (8 * w)) - 1); }

/* src/openzl/codecs/delta/encode_delta_kernel.c:20-123 and
 * encode_delta_binding.c:35-43 (format >= 13): the first element goes to the
 * codec header, the stream holds d[i-1] = x[i] - x[i-1] mod 2^(8w). */
void orc_delta_encode(void* deltas, void* first, const void* src, size_t n, int width)
{
    if (n == 0) return;
    wrw(first, 0, width, rdw(src, 0, width));
    for (size_t i = 1; i < n; i++)
        wrw(deltas, i - 1, width, (rdw(src, i, width) - rdw(src, i - 1, width)) & wmask(width));
}

/* src/openzl/codecs/delta/decode_delta_kernel.c:14-267: inclusive prefix sum
 * seeded by the header value. */
void orc_delta_decode(void* dst, const void* first, const void* deltas, size_t n, int width)
{
    if (n == 0) return;
    uint64_t acc = rdw(first, 0, width);
    wrw(dst, 0, width, acc);
    for (size_t i = 1; i < n; i++) {
        acc = (acc + rdw(deltas, i - 1, width)) & wmask(width);
        wrw(dst, i, width, acc);
    }
}

/* src/openzl/codecs/zigzag/encode_zigzag_kernel.c:8-59: z = (v<<1) ^ (v>>(w-1)),
 * arithmetic shift of the signed value. */
void orc_zigzag_encode(void* dst, const void* src, size_t n, int width)
{
    int bits = 8 * width;
    for (size_t i = 0; i < n; i++) {
        uint64_t v = rdw(src, i, width);
        uint64_t sign = (v >> (bits - 1)) & 1 ? wmask(width) : 0;
        wrw(dst, i, width, ((v << 1) ^ sign) & wmask(width));
    }
}

/* src/openzl/codecs/zigzag/decode_zigzag_kernel.c:8-63: v = (z>>1) ^ -(z&1). */
void orc_zigzag_decode(void* dst, const void* src, size_t n, int width)
{
    for (size_t i = 0; i < n; i++) {
        uint64_t z = rdw(src, i, width);
        wrw(dst, i, width, ((z >> 1) ^ (0 - (z & 1))) & wmask(width));
    }
}

/* src/openzl/codecs/transpose/encode_transpose_kernel.c:63-107 (generic path):
 * plane p holds byte p of every element, planes laid out back to back. */
void orc_transpose_split(uint8_t* planes, const uint8_t* src, size_t n, int width)
{
    for (int p = 0; p < width; p++)
        for (size_t i = 0; i < n; i++) planes[(size_t)p * n + i] = src[i * (size_t)width + p];
}

/* src/openzl/codecs/transpose/decode_transpose_kernel.c:60-102 (generic path). */
void orc_transpose_join(uint8_t* dst, const uint8_t* planes, size_t n, int width)
{
    for (int p = 0; p < width; p++)
        for (size_t i = 0; i < n; i++) dst[i * (size_t)width + p] = planes[(size_t)p * n + i];
}

/* src/openzl/codecs/bitpack/encode_bitpack_binding.c:51-76 computeNbBits:
 * 1 + highbit(max), with max == 0 giving 1. */
int orc_bitpack_nbbits(const void* src, size_t n, int width)
{
    uint64_t mx = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t v = rdw(src, i, width);
        if (v > mx) mx = v;
    }
    return 1 + (mx == 0 ? 0 : highbit64(mx));
}

/* src/openzl/codecs/bitpack/common_bitpack_kernel.c:24-30, 131-187: LSB-first
 * packing into ceil(n*nbBits/8) bytes; nbBits == 8w is a plain copy. */
size_t orc_bitpack_encode(uint8_t* dst, const void* src, size_t n, int width, int nbBits)
{
    size_t size = (n * (size_t)nbBits + 7) / 8;
    memset(dst, 0, size);
    size_t bit = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t v = rdw(src, i, width);
        for (int b = 0; b < nbBits; b++, bit++)
            if ((v >> b) & 1) dst[bit >> 3] |= (uint8_t)(1u << (bit & 7));
    }
    return size;
}

/* src/openzl/codecs/bitpack/encode_bitpack_binding.c:120-143: byte0 =
 * log2(w)<<6 | (nbBits-1); byte1 (only when non-zero) = number of padding
 * elements the packed size could still hold. */
size_t orc_bitpack_header(uint8_t hdr[2], size_t n, int width, int nbBits, size_t packedSize)
{
    static const uint8_t lg[9] = { 0, 0, 1, 0, 2, 0, 0, 0, 3 };
    size_t maxElts = packedSize * 8 / (size_t)nbBits;
    hdr[0] = (uint8_t)((lg[width] << 6) | (nbBits - 1));
    if (maxElts > n) { hdr[1] = (uint8_t)(maxElts - n); return 2; }
    return 1;
}

/* src/openzl/codecs/bitpack/decode_bitpack_binding.c:10-60: n = floor(8L/nbBits)
 * minus the padding count. Returns the number of elements, or (size_t)-1. */
size_t orc_bitpack_decode(void* dst, size_t capacity, const uint8_t* src, size_t srcSize,
                          const uint8_t* hdr, size_t hdrSize, int* widthOut)
{
    if (hdrSize < 1 || hdrSize > 2) return (size_t)-1;
    int width = 1 << (hdr[0] >> 6);
    int nbBits = (hdr[0] & 63) + 1;
    if (nbBits > 8 * width) return (size_t)-1;
    size_t extra = hdrSize == 2 ? hdr[1] : 0;
    size_t maxElts = srcSize * 8 / (size_t)nbBits;
    if (extra > maxElts) return (size_t)-1;
    size_t n = maxElts - extra;
    if (n > capacity) return (size_t)-1;
    if ((n * (size_t)nbBits + 7) / 8 != srcSize) return (size_t)-1;
    size_t bit = 0;
    for (size_t i = 0; i < n; i++) {
        uint64_t v = 0;
        for (int b = 0; b < nbBits; b++, bit++) v |= (uint64_t)((src[bit >> 3] >> (bit & 7)) & 1) << b;
        wrw(dst, i, width, v);
    }
    if (widthOut) *widthOut = width;
    return n;
}

/* src/openzl/codecs/flatpack/encode_flatpack_kernel.c:8-33, 93-130 and
 * common_flatpack.h: sorted alphabet of present bytes; indices packed LSB-first
 * with nbBits = card<=1 ? card : 1+highbit(card-1), then a terminating 1 bit.
 * Returns the alphabet size. */
size_t orc_flatpack_encode(uint8_t* alphabet, uint8_t* packed, size_t* packedSize,
                           const uint8_t* src, size_t n)
{
    uint8_t map[256] = { 0 };
    if (n == 0) { *packedSize = 0; return 0; }
    for (size_t i = 0; i < n; i++) map[src[i]] = 1;
    size_t card = 0;
    for (int s = 0; s < 256; s++)
        if (map[s]) { map[s] = (uint8_t)card; alphabet[card++] = (uint8_t)s; }
    size_t nbBits = card <= 1 ? card : 1 + (size_t)highbit32((uint32_t)card - 1);
    size_t size = 1 + nbBits * n / 8;
    memset(packed, 0, size);
    size_t bit = 0;
    for (size_t i = 0; i < n; i++) {
        for (size_t b = 0; b < nbBits; b++, bit++)
            if ((map[src[i]] >> b) & 1) packed[bit >> 3] |= (uint8_t)(1u << (bit & 7));
    }
    packed[bit >> 3] |= (uint8_t)(1u << (bit & 7));
    *packedSize = size;
    return card;
}

/* src/openzl/codecs/flatpack/decode_flatpack_kernel.c:359: the element count
 * follows from the position of the terminating bit. Returns n or (size_t)-1. */
size_t orc_flatpack_decode(uint8_t* dst, size_t capacity, const uint8_t* alphabet,
                           size_t alphabetSize, const uint8_t* packed, size_t packedSize)
{
    if (packedSize == 0) return 0;
    if (alphabetSize == 0 || packed[packedSize - 1] == 0) return (size_t)-1;
    size_t nbBits = alphabetSize <= 1 ? alphabetSize : 1 + (size_t)highbit32((uint32_t)alphabetSize - 1);
    size_t totalBits = (packedSize - 1) * 8 + (size_t)highbit32(packed[packedSize - 1]);
    size_t n = totalBits / nbBits;
    if (n > capacity) return (size_t)-1;
    size_t bit = 0;
    for (size_t i = 0; i < n; i++) {
        size_t v = 0;
        for (size_t b = 0; b < nbBits; b++, bit++) v |= (size_t)((packed[bit >> 3] >> (bit & 7)) & 1) << b;
        if (v >= alphabetSize) return (size_t)-1;
        dst[i] = alphabet[v];
    }
    return n;
}

/* src/openzl/shared/histogram.c:84-118 / src/openzl/fse/compress/hist.c. */
void orc_histogram(uint32_t count[256], unsigned* maxSym, unsigned* cardinality,
                   const uint8_t* src, size_t n)
{
    memset(count, 0, 256 * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++) count[src[i]]++;
    unsigned m = 0, c = 0;
    for (unsigned s = 0; s < 256; s++)
        if (count[s]) { m = s; c++; }
    *maxSym = m;
    *cardinality = c;
}

/* src/openzl/shared/data_stats.c:15-44: entry i = floor(log2(1 + i/256) * 2^16),
 * generated here from that closed form in exact integer arithmetic: the largest
 * v with 2^(v/65536) <= 1 + i/256, found by bisection on 2^v <= ((256+i)/256)^65536
 * evaluated as a repeated-squaring comparison in 128-bit fixed point. */
const uint16_t orc_log2_table[256] = {
#include "log2_table.inc"
};

/* src/openzl/shared/data_stats.c:46-73 ZL_calculateEntropy: integer sum of
 * count * log2 in 16.16 fixed point with 2^62/total normalisation, then one
 * double division. */
double orc_entropy(const uint32_t* count, size_t maxValue, size_t total)
{
    if (total == 0) return 0;
    int64_t acc = 0;
    int64_t const norm = ((int64_t)1 << 62) / (int64_t)total;
    for (size_t i = 0; i <= maxValue; i++) {
        int64_t c = count[i];
        int64_t li = c * norm;
        int clz = __builtin_clzll((uint64_t)li | (1u << 9));
        int shift = (62 - 8) - clz + 1;
        li >>= shift;
        li &= 255;
        acc += -c * ((int64_t)orc_log2_table[li] - (int64_t)clz * 65536);
    }
    return (double)acc / (double)(total * 65536) - 1;
}

/* src/openzl/codecs/entropy/encode_entropy_binding.c:601-753 entropyCompressChunk
 * for a serial chunk, with the estimates of src/openzl/shared/data_stats.c:
 * 288-299 (Huffman fast estimate), 301-312 (bitpacked), 320-327 (flatpacked).
 * hufMode selects EBM_huf (1) or EBM_fse (0). */
int orc_entropy_decide(const uint8_t* src, size_t n, int hufMode)
{
    if (n <= 1) return ORC_BK_STORE;
    uint32_t count[256];
    unsigned maxSym, card;
    orc_histogram(count, &maxSym, &card, src, n);
    if (card == 1) return ORC_BK_CONSTANT;
    double h = orc_entropy(count, 255, n);
    size_t entropySize;
    if (hufMode) {
        if (h > 7) entropySize = n;
        else {
            double e = h < 1 ? 1 : h;
            entropySize = (size_t)((e * (double)n) / 8);
        }
    } else {
        entropySize = (size_t)(h * (double)n + 7) / 8;
    }
    size_t hdr = card / 4 > 10 ? card / 4 : 10;
    size_t baseline = entropySize + hdr < n ? entropySize + hdr : n;
    size_t flat = (n * (size_t)orc_next_pow2(card) + 7) / 8 + card;
    size_t bitp = (n * (size_t)orc_next_pow2((uint64_t)maxSym + 1) + 7) / 8;
    if (flat < bitp) {
        if (flat < baseline) return ORC_BK_FLATPACK;
    } else {
        if (bitp < baseline) return ORC_BK_BITPACK;
    }
    if (n <= baseline) return ORC_BK_STORE;
    return hufMode ? ORC_BK_HUF : ORC_BK_FSE;
}

/* ============================ FSE ============================ */

/* src/openzl/fse/compress/fse_compress.c:361-396 FSE_minTableLog /
 * FSE_optimalTableLog_internal; limits from src/openzl/fse/fse.h:702-736. */
static unsigned fse_min_tablelog(size_t n, unsigned maxSym)
{
    unsigned a = (unsigned)highbit32((uint32_t)n) + 1;
    unsigned b = (unsigned)highbit32(maxSym) + 2;
    return a < b ? a : b;
}
unsigned orc_fse_optimal_tablelog(unsigned maxTableLog, size_t n, unsigned maxSym, unsigned minus)
{
    unsigned maxBitsSrc = (unsigned)highbit32((uint32_t)(n - 1)) - minus;
    unsigned tl = maxTableLog ? maxTableLog : 11;
    unsigned minBits = fse_min_tablelog(n, maxSym);
    if (maxBitsSrc < tl) tl = maxBitsSrc;
    if (minBits > tl) tl = minBits;
    if (tl < 5) tl = 5;
    if (tl > 12) tl = 12;
    return tl;
}

/* src/openzl/fse/compress/fse_compress.c:401-485 FSE_normalizeM2 (fallback). */
static int fse_normalize_m2(int16_t* norm, unsigned tableLog, const uint32_t* count, size_t total,
                            unsigned maxSym, int16_t lowProb)
{
    const int16_t UNSET = -2;
    uint32_t distributed = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    uint32_t lowOne = (uint32_t)((total * 3) >> (tableLog + 1));
    for (unsigned s = 0; s <= maxSym; s++) {
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) { norm[s] = lowProb; distributed++; total -= count[s]; continue; }
        if (count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; continue; }
        norm[s] = UNSET;
    }
    uint32_t toDistribute = (1u << tableLog) - distributed;
    if (toDistribute == 0) return 0;
    if ((total / toDistribute) > lowOne) {
        lowOne = (uint32_t)((total * 3) / (toDistribute * 2));
        for (unsigned s = 0; s <= maxSym; s++)
            if (norm[s] == UNSET && count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; }
        toDistribute = (1u << tableLog) - distributed;
    }
    if (distributed == maxSym + 1) {
        uint32_t maxV = 0, maxC = 0;
        for (unsigned s = 0; s <= maxSym; s++)
            if (count[s] > maxC) { maxV = s; maxC = count[s]; }
        norm[maxV] = (int16_t)(norm[maxV] + (int16_t)toDistribute);
        return 0;
    }
    if (total == 0) {
        for (unsigned s = 0; toDistribute > 0; s = (s + 1) % (maxSym + 1))
            if (norm[s] > 0) { toDistribute--; norm[s]++; }
        return 0;
    }
    uint64_t vStepLog = 62 - tableLog;
    uint64_t mid = (1ull << (vStepLog - 1)) - 1;
    uint64_t rStep = (((1ull << vStepLog) * toDistribute) + mid) / (uint32_t)total;
    uint64_t tmp = mid;
    for (unsigned s = 0; s <= maxSym; s++) {
        if (norm[s] == UNSET) {
            uint64_t end = tmp + count[s] * rStep;
            uint32_t a = (uint32_t)(tmp >> vStepLog), b = (uint32_t)(end >> vStepLog);
            if (b - a < 1) return -1;
            norm[s] = (int16_t)(b - a);
            tmp = end;
        }
    }
    return 0;
}

/* src/openzl/fse/compress/fse_compress.c:487-549 FSE_normalizeCount. */
int orc_fse_normalize(int16_t* norm, unsigned tableLog, const uint32_t* count, size_t total,
                      unsigned maxSym, int useLowProb)
{
    static const uint32_t rtb[8] = { 0, 473195, 504333, 520860, 550000, 700000, 750000, 830000 };
    if (tableLog == 0) tableLog = 11;
    if (tableLog < 5 || tableLog > 12) return -1;
    if (tableLog < fse_min_tablelog(total, maxSym)) return -1;
    int16_t lowProb = useLowProb ? -1 : 1;
    uint64_t scale = 62 - tableLog;
    uint64_t step = (1ull << 62) / (uint32_t)total;
    uint64_t vStep = 1ull << (scale - 20);
    int still = 1 << tableLog;
    unsigned largest = 0;
    int16_t largestP = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tableLog);
    for (unsigned s = 0; s <= maxSym; s++) {
        if (count[s] == total) return 0;
        if (count[s] == 0) { norm[s] = 0; continue; }
        if (count[s] <= lowThreshold) { norm[s] = lowProb; still--; continue; }
        int16_t p = (int16_t)((count[s] * step) >> scale);
        if (p < 8) {
            uint64_t restToBeat = vStep * rtb[p];
            p = (int16_t)(p + ((count[s] * step) - ((uint64_t)p << scale) > restToBeat));
        }
        if (p > largestP) { largestP = p; largest = s; }
        norm[s] = p;
        still -= p;
    }
    if (-still >= (norm[largest] >> 1))
        return fse_normalize_m2(norm, tableLog, count, total, maxSym, lowProb);
    norm[largest] = (int16_t)(norm[largest] + still);
    return 0;
}

static unsigned fse_step(unsigned tableSize) { return (tableSize >> 1) + (tableSize >> 3) + 3; }

/* Symbol spreading shared by the encoder and decoder tables:
 * src/openzl/fse/compress/fse_compress.c:93-155 and
 * src/openzl/fse/decompress/fse_decompress.c:93-160. Low-probability (-1)
 * symbols take the top cells, the rest are laid out with the fixed step. */
static int fse_spread(uint8_t* cellSym, const int16_t* norm, unsigned maxSym, unsigned tableLog)
{
    unsigned tableSize = 1u << tableLog, mask = tableSize - 1, step = fse_step(tableSize);
    unsigned high = tableSize - 1;
    for (unsigned s = 0; s <= maxSym; s++)
        if (norm[s] == -1) cellSym[high--] = (uint8_t)s;
    unsigned pos = 0;
    for (unsigned s = 0; s <= maxSym; s++) {
        for (int k = 0; k < norm[s]; k++) {
            cellSym[pos] = (uint8_t)s;
            pos = (pos + step) & mask;
            while (pos > high) pos = (pos + step) & mask;
        }
    }
    return pos == 0 ? 0 : -1;
}

/* src/openzl/fse/compress/fse_compress.c:69-215 FSE_buildCTable_wksp. */
int orc_fse_build_ctable(orc_fse_ctable* ct, const int16_t* norm, unsigned maxSym, unsigned tableLog)
{
    unsigned tableSize = 1u << tableLog;
    uint8_t cellSym[4096];
    uint16_t cumul[258];
    if (fse_spread(cellSym, norm, maxSym, tableLog)) return -1;
    cumul[0] = 0;
    for (unsigned u = 1; u <= maxSym + 1; u++)
        cumul[u] = (uint16_t)(cumul[u - 1] + (norm[u - 1] == -1 ? 1 : norm[u - 1]));
    for (unsigned u = 0; u < tableSize; u++) ct->stateTable[cumul[cellSym[u]]++] = (uint16_t)(tableSize + u);
    unsigned total = 0;
    for (unsigned s = 0; s <= maxSym; s++) {
        int c = norm[s];
        if (c == 0) {
            ct->deltaNbBits[s] = ((tableLog + 1) << 16) - tableSize;
            ct->deltaFindState[s] = 0;
        } else if (c == -1 || c == 1) {
            ct->deltaNbBits[s] = (tableLog << 16) - tableSize;
            ct->deltaFindState[s] = (int32_t)total - 1;
            total++;
        } else {
            unsigned maxBitsOut = tableLog - (unsigned)highbit32((uint32_t)c - 1);
            unsigned minStatePlus = (unsigned)c << maxBitsOut;
            ct->deltaNbBits[s] = (maxBitsOut << 16) - minStatePlus;
            ct->deltaFindState[s] = (int32_t)total - c;
            total += (unsigned)c;
        }
    }
    ct->tableLog = tableLog;
    ct->maxSym = maxSym;
    return 0;
}

/* LSB-first bit writer with an end mark: src/openzl/fse/bitstream.h:154-250. */
typedef struct { uint8_t* dst; size_t cap; size_t bitPos; int overflow; } bitw;
static void bw_init(bitw* w, uint8_t* dst, size_t cap) { w->dst = dst; w->cap = cap; w->bitPos = 0; w->overflow = 0; memset(dst, 0, cap); }
static void bw_add(bitw* w, uint64_t v, unsigned nb)
{
    for (unsigned b = 0; b < nb; b++, w->bitPos++) {
        size_t byte = w->bitPos >> 3;
        if (byte >= w->cap) { w->overflow = 1; continue; }
        if ((v >> b) & 1) w->dst[byte] |= (uint8_t)(1u << (w->bitPos & 7));
    }
}
static size_t bw_close(bitw* w)
{
    bw_add(w, 1, 1);
    return (w->bitPos + 7) >> 3;
}

/* src/openzl/fse/compress/fse_compress.c:607-660 FSE_compress_usingCTable_generic
 * with the state helpers of src/openzl/fse/fse.h:527-569. Symbols are encoded
 * from the last to the first; symbol at position p uses state (nbStates-1-p)
 * mod nbStates. Returns 0 when n <= nbStates or the output does not fit. */
size_t orc_fse_encode(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                      const orc_fse_ctable* ct, unsigned nbStates)
{
    if (n <= nbStates) return 0;
    /* BIT_initCStream refuses buffers of 8 bytes or fewer; BIT_closeCStream
     * reports 0 when the write pointer passed cap - 8. */
    if (cap <= 8) return 0;
    uint32_t state[4];
    unsigned tl = ct->tableLog;
    bitw w;
    bw_init(&w, dst, cap);
    size_t s = nbStates - (n % nbStates);
    size_t p = n;
    for (unsigned i = 0; i < nbStates; i++) {
        unsigned sym = src[--p];
        uint32_t nbOut = (ct->deltaNbBits[sym] + (1u << 15)) >> 16;
        uint32_t value = (nbOut << 16) - ct->deltaNbBits[sym];
        state[(s++) % nbStates] = ct->stateTable[(value >> nbOut) + ct->deltaFindState[sym]];
    }
    (void)tl;
    while (p > 0) {
        unsigned sym = src[--p];
        uint32_t* st = &state[(s++) % nbStates];
        uint32_t nbOut = (*st + ct->deltaNbBits[sym]) >> 16;
        bw_add(&w, *st, nbOut);
        *st = ct->stateTable[(*st >> nbOut) + ct->deltaFindState[sym]];
    }
    for (unsigned i = 0; i < nbStates; i++) bw_add(&w, state[(s++) % nbStates], ct->tableLog);
    size_t size = bw_close(&w);
    if (w.overflow || size + 8 > cap) return 0;
    return size;
}

/* src/openzl/fse/compress/fse_compress.c:246-357 FSE_writeNCount_generic. */
size_t orc_fse_write_ncount(uint8_t* dst, size_t cap, const int16_t* norm, unsigned maxSym,
                            unsigned tableLog)
{
    (void)cap;
    uint8_t* out = dst;
    int tableSize = 1 << tableLog;
    uint32_t bits = 0;
    int nbits = 0;
    unsigned symbol = 0, alphabet = maxSym + 1;
    int previous0 = 0;
    bits += (tableLog - 5) << nbits;
    nbits += 4;
    int remaining = tableSize + 1, threshold = tableSize, nb = (int)tableLog + 1;
    while (symbol < alphabet && remaining > 1) {
        if (previous0) {
            unsigned start = symbol;
            while (symbol < alphabet && !norm[symbol]) symbol++;
            if (symbol == alphabet) break;
            while (symbol >= start + 24) {
                start += 24;
                bits += 0xFFFFu << nbits;
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16;
            }
            while (symbol >= start + 3) { start += 3; bits += 3u << nbits; nbits += 2; }
            bits += (symbol - start) << nbits;
            nbits += 2;
            if (nbits > 16) {
                out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
                bits >>= 16; nbits -= 16;
            }
        }
        {
            int count = norm[symbol++];
            int const max = (2 * threshold - 1) - remaining;
            remaining -= count < 0 ? -count : count;
            count++;
            if (count >= threshold) count += max;
            bits += (uint32_t)count << nbits;
            nbits += nb;
            nbits -= (count < max);
            previous0 = (count == 1);
            if (remaining < 1) return 0;
            while (remaining < threshold) { nb--; threshold >>= 1; }
        }
        if (nbits > 16) {
            out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8); out += 2;
            bits >>= 16; nbits -= 16;
        }
    }
    if (remaining != 1) return 0;
    out[0] = (uint8_t)bits; out[1] = (uint8_t)(bits >> 8);
    out += (nbits + 7) / 8;
    return (size_t)(out - dst);
}

/* Reader for the format above: src/openzl/fse/common/entropy_common.c
 * FSE_readNCount. Returns bytes consumed or -1. */
long orc_fse_read_ncount(int16_t* norm, unsigned* maxSymOut, unsigned* tableLogOut,
                         const uint8_t* src, size_t size)
{
    size_t bitPos = 0, totalBits = size * 8;
#define RD(nb) ({ uint32_t _v = 0; for (int _b = 0; _b < (nb); _b++, bitPos++) \
        _v |= (uint32_t)(bitPos < totalBits ? ((src[bitPos >> 3] >> (bitPos & 7)) & 1) : 0) << _b; _v; })
    unsigned tableLog = RD(4) + 5;
    if (tableLog > 12) return -1;
    int remaining = (1 << tableLog) + 1, threshold = 1 << tableLog, nb = (int)tableLog + 1;
    unsigned symbol = 0;
    int previous0 = 0;
    memset(norm, 0, 256 * sizeof(int16_t));
    while (remaining > 1 && symbol <= 255) {
        if (previous0) {
            unsigned n0 = symbol;
            for (;;) {
                uint32_t r = RD(2);
                n0 += r;
                if (r != 3) break;
            }
            if (n0 > 255) return -1;
            symbol = n0;
        }
        int max = (2 * threshold - 1) - remaining;
        int count;
        uint32_t low = RD(nb - 1);
        if ((int)low < max) count = (int)low;
        else {
            uint32_t hi = RD(1);
            count = (int)(low + (hi << (nb - 1)));
            if (count >= threshold) count -= max;
        }
        count--;
        remaining -= count < 0 ? -count : count;
        norm[symbol++] = (int16_t)count;
        previous0 = !count;
        while (remaining < threshold) { nb--; threshold >>= 1; }
    }
#undef RD
    if (remaining != 1) return -1;
    if (bitPos > totalBits) return -1;
    *maxSymOut = symbol - 1;
    *tableLogOut = tableLog;
    return (long)((bitPos + 7) / 8);
}

/* src/openzl/fse/decompress/fse_decompress.c:73-176 (table) and :255-338
 * (decode). With the symbol count known, symbol p is read from state p mod
 * nbStates; every symbol but the last one of each state advances its state by
 * nbBits read backwards from the stream. */
int orc_fse_decode(uint8_t* dst, size_t n, const uint8_t* src, size_t size,
                   const int16_t* norm, unsigned maxSym, unsigned tableLog, unsigned nbStates)
{
    unsigned tableSize = 1u << tableLog;
    uint8_t cellSym[4096];
    uint8_t cellBits[4096];
    uint16_t cellBase[4096];
    uint16_t next[256];
    if (fse_spread(cellSym, norm, maxSym, tableLog)) return -1;
    for (unsigned s = 0; s <= maxSym; s++) next[s] = (uint16_t)(norm[s] == -1 ? 1 : norm[s]);
    for (unsigned u = 0; u < tableSize; u++) {
        unsigned ns = next[cellSym[u]]++;
        cellBits[u] = (uint8_t)(tableLog - (unsigned)highbit32(ns));
        cellBase[u] = (uint16_t)((ns << cellBits[u]) - tableSize);
    }
    if (size == 0 || src[size - 1] == 0) return -1;
    /* position of the end mark: bits strictly below it are payload */
    long pos = (long)(size - 1) * 8 + highbit32(src[size - 1]);
#define RDB(nb) ({ uint32_t _v = 0; for (int _b = (int)(nb) - 1; _b >= 0; _b--) { pos--; \
        _v |= (uint32_t)(pos >= 0 ? ((src[pos >> 3] >> (pos & 7)) & 1) : 0) << _b; } _v; })
    uint32_t st[4];
    for (unsigned j = 0; j < nbStates; j++) st[j] = RDB(tableLog);
    for (size_t p = 0; p < n; p++) {
        unsigned j = (unsigned)(p % nbStates);
        dst[p] = cellSym[st[j]];
        if (p + nbStates < n) st[j] = cellBase[st[j]] + RDB(cellBits[st[j]]);
    }
#undef RDB
    return pos == 0 ? 0 : -1;
}

/* ============================ Huffman ============================ */

/* src/openzl/fse/compress/huf_compress.c:1218-1224 HUF_optimalTableLog. */
unsigned orc_huf_optimal_tablelog(unsigned maxTableLog, size_t n, unsigned maxSym)
{
    return orc_fse_optimal_tablelog(maxTableLog, n, maxSym, 1);
}

typedef struct { uint32_t count; uint16_t parent; uint8_t byte; uint8_t nbBits; } hnode;

/* src/openzl/fse/compress/huf_compress.c:487-495 HUF_getIndex bucket map. */
static uint32_t huf_bucket(uint32_t c) { return c < 166 ? c : (uint32_t)highbit32(c) + 158; }

static void huf_insertion(hnode* a, int lo, int hi)
{
    a += lo;
    int size = hi - lo + 1;
    for (int i = 1; i < size; i++) {
        hnode key = a[i];
        int j = i - 1;
        while (j >= 0 && a[j].count < key.count) { a[j + 1] = a[j]; j--; }
        a[j + 1] = key;
    }
}
static int huf_partition(hnode* a, int lo, int hi)
{
    uint32_t pivot = a[hi].count;
    int i = lo - 1;
    for (int j = lo; j < hi; j++)
        if (a[j].count > pivot) { i++; hnode t = a[i]; a[i] = a[j]; a[j] = t; }
    hnode t = a[i + 1]; a[i + 1] = a[hi]; a[hi] = t;
    return i + 1;
}
/* src/openzl/fse/compress/huf_compress.c:547-585 HUF_simpleQuickSort. */
static void huf_quicksort(hnode* a, int lo, int hi)
{
    if (hi - lo < 8) { huf_insertion(a, lo, hi); return; }
    while (lo < hi) {
        int idx = huf_partition(a, lo, hi);
        if (idx - lo < hi - idx) { huf_quicksort(a, lo, idx - 1); lo = idx + 1; }
        else { huf_quicksort(a, idx + 1, hi); hi = idx - 1; }
    }
}

/* src/openzl/fse/compress/huf_compress.c:596-641 HUF_sort (bucket placement by
 * symbol order, then quicksort of the log2 buckets), :657-704 HUF_buildTree,
 * :352-500 HUF_setMaxHeight, :706-729 HUF_buildCTableFromTree. Returns the final
 * max code length, or -1. */
int orc_huf_build_ctable(uint8_t* nbBits, uint16_t* values, const uint32_t* count,
                         unsigned maxSym, unsigned maxNbBits)
{
    hnode tbl[2 * 256 + 2];
    hnode* node = tbl + 1;
    memset(tbl, 0, sizeof(tbl));
    if (maxNbBits == 0) maxNbBits = 11;
    /* sort */
    {
        uint16_t base[192], curr[192];
        memset(base, 0, sizeof(base));
        for (unsigned s = 0; s <= maxSym; s++) base[huf_bucket(count[s])]++;
        for (int r = 191; r > 0; r--) { base[r - 1] = (uint16_t)(base[r - 1] + base[r]); }
        for (int r = 0; r < 192; r++) curr[r] = base[r];
        for (unsigned s = 0; s <= maxSym; s++) {
            uint32_t r = huf_bucket(count[s]) + 1;
            uint16_t pos = curr[r]++;
            node[pos].count = count[s];
            node[pos].byte = (uint8_t)s;
        }
        for (int r = 166; r < 191; r++) {
            int sz = curr[r] - base[r];
            if (sz > 1) huf_quicksort(node + base[r], 0, sz - 1);
        }
    }
    /* tree */
    int startNode = 256;
    int nonNull = (int)maxSym;
    while (node[nonNull].count == 0) nonNull--;
    int lowS = nonNull, nodeNb = startNode, nodeRoot = nodeNb + lowS - 1, lowN = nodeNb;
    node[nodeNb].count = node[lowS].count + node[lowS - 1].count;
    node[lowS].parent = node[lowS - 1].parent = (uint16_t)nodeNb;
    nodeNb++;
    lowS -= 2;
    for (int k = nodeNb; k <= nodeRoot; k++) node[k].count = 1u << 30;
    tbl[0].count = 1u << 31;
    while (nodeNb <= nodeRoot) {
        int n1 = (node[lowS].count < node[lowN].count) ? lowS-- : lowN++;
        int n2 = (node[lowS].count < node[lowN].count) ? lowS-- : lowN++;
        node[nodeNb].count = node[n1].count + node[n2].count;
        node[n1].parent = node[n2].parent = (uint16_t)nodeNb;
        nodeNb++;
    }
    node[nodeRoot].nbBits = 0;
    for (int k = nodeRoot - 1; k >= startNode; k--) node[k].nbBits = (uint8_t)(node[node[k].parent].nbBits + 1);
    for (int k = 0; k <= nonNull; k++) node[k].nbBits = (uint8_t)(node[node[k].parent].nbBits + 1);
    /* height limit */
    unsigned largestBits = node[nonNull].nbBits;
    unsigned target = maxNbBits;
    if (largestBits > target) {
        int totalCost = 0;
        const int baseCost = 1 << (largestBits - target);
        int n = nonNull;
        while (node[n].nbBits > target) {
            totalCost += baseCost - (1 << (largestBits - node[n].nbBits));
            node[n].nbBits = (uint8_t)target;
            n--;
        }
        while (node[n].nbBits == target) --n;
        totalCost >>= (largestBits - target);
        const uint32_t none = 0xF0F0F0F0u;
        uint32_t rankLast[14];
        for (int k = 0; k < 14; k++) rankLast[k] = none;
        {
            unsigned cur = target;
            for (int pos = n; pos >= 0; pos--) {
                if (node[pos].nbBits >= cur) continue;
                cur = node[pos].nbBits;
                rankLast[target - cur] = (uint32_t)pos;
            }
        }
        while (totalCost > 0) {
            unsigned nbDec = (unsigned)highbit32((uint32_t)totalCost) + 1;
            for (; nbDec > 1; nbDec--) {
                uint32_t highPos = rankLast[nbDec], lowPos = rankLast[nbDec - 1];
                if (highPos == none) continue;
                if (lowPos == none) break;
                if (node[highPos].count <= 2 * node[lowPos].count) break;
            }
            while (nbDec <= 12 && rankLast[nbDec] == none) nbDec++;
            totalCost -= 1 << (nbDec - 1);
            node[rankLast[nbDec]].nbBits++;
            if (rankLast[nbDec - 1] == none) rankLast[nbDec - 1] = rankLast[nbDec];
            if (rankLast[nbDec] == 0) rankLast[nbDec] = none;
            else {
                rankLast[nbDec]--;
                if (node[rankLast[nbDec]].nbBits != target - nbDec) rankLast[nbDec] = none;
            }
        }
        while (totalCost < 0) {
            if (rankLast[1] == none) {
                while (node[n].nbBits == target) n--;
                node[n + 1].nbBits--;
                rankLast[1] = (uint32_t)(n + 1);
                totalCost++;
                continue;
            }
            node[rankLast[1] + 1].nbBits--;
            rankLast[1]++;
            totalCost++;
        }
        largestBits = target;
    }
    if (largestBits > 12) return -1;
    /* canonical values */
    uint16_t nbPerRank[13] = { 0 }, valPerRank[13] = { 0 };
    for (int k = 0; k <= nonNull; k++) nbPerRank[node[k].nbBits]++;
    {
        uint16_t mn = 0;
        for (int r = (int)largestBits; r > 0; r--) {
            valPerRank[r] = mn;
            mn = (uint16_t)(mn + nbPerRank[r]);
            mn >>= 1;
        }
    }
    for (unsigned s = 0; s <= maxSym; s++) nbBits[s] = 0;
    for (unsigned k = 0; k <= maxSym; k++) nbBits[node[k].byte] = node[k].nbBits;
    /* unused symbols (length 0) carry no code value */
    for (unsigned s = 0; s <= maxSym; s++) values[s] = nbBits[s] ? valPerRank[nbBits[s]]++ : 0;
    return (int)largestBits;
}

/* src/openzl/fse/compress/huf_compress.c:808-948, 955-1086: one stream, codes
 * of symbols n-1 down to 0 in LSB-first order, a 1-bit end mark, 0 if the
 * stream cannot fit. */
size_t orc_huf_encode_1x(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                         const uint8_t* nbBits, const uint16_t* values)
{
    if (cap < 8) return 0;
    bitw w;
    bw_init(&w, dst, cap);
    for (size_t p = n; p > 0; p--) bw_add(&w, values[src[p - 1]], nbBits[src[p - 1]]);
    size_t size = bw_close(&w);
    /* HUF_closeCStream: overflow once the write pointer reaches cap - 8 */
    size_t fullBytes = w.bitPos >> 3;
    if (w.overflow || fullBytes >= cap - 8) return 0;
    return size;
}

/* src/openzl/fse/compress/huf_compress.c:1139-1186: 6-byte jump table of LE16
 * sizes of streams 1-3, four segments of (n+3)/4 symbols (last takes the rest);
 * 0 if any stream is empty or larger than 65535 bytes. */
size_t orc_huf_encode_4x(uint8_t* dst, size_t cap, const uint8_t* src, size_t n,
                         const uint8_t* nbBits, const uint16_t* values)
{
    if (cap < 6 + 1 + 1 + 1 + 8) return 0;
    if (n < 12) return 0;
    size_t seg = (n + 3) / 4;
    size_t op = 6;
    const uint8_t* ip = src;
    for (int k = 0; k < 4; k++) {
        size_t len = k < 3 ? seg : (size_t)(src + n - ip);
        size_t c = orc_huf_encode_1x(dst + op, cap - op, ip, len, nbBits, values);
        if (c == 0 || c > 65535) return 0;
        if (k < 3) { dst[2 * k] = (uint8_t)c; dst[2 * k + 1] = (uint8_t)(c >> 8); }
        op += c;
        ip += len;
    }
    return op;
}

/* Canonical decode from weights (w>0: length = tableLog + 1 - w, with tableLog
 * fixed by sum 2^(w-1) = 2^tableLog), src/openzl/codecs/entropy/
 * decode_entropy_binding.c:205-347. Streams are read backwards from their end
 * mark; any correct decoder of this format yields the same bytes. */
static int huf_decode_stream(uint8_t* dst, size_t n, const uint8_t* src, size_t size,
                             const uint8_t* len, const uint16_t* val, unsigned maxLen)
{
    (void)maxLen;
    if (size == 0 || src[size - 1] == 0) return -1;
    long pos = (long)(size - 1) * 8 + highbit32(src[size - 1]);
    for (size_t p = 0; p < n; p++) {
        /* read bit by bit (MSB of the code first) and match the canonical code */
        uint32_t code = 0;
        unsigned l = 0;
        int found = -1;
        while (found < 0) {
            pos--;
            if (pos < 0) return -1;
            code = (code << 1) | (uint32_t)((src[pos >> 3] >> (pos & 7)) & 1);
            l++;
            if (l > 12) return -1;
            for (int s = 0; s < 256; s++)
                if (len[s] == l && val[s] == code) { found = s; break; }
        }
        dst[p] = (uint8_t)found;
    }
    return pos == 0 ? 0 : -1;
}

int orc_huf_decode(uint8_t* dst, size_t n, const uint8_t* src, size_t size,
                   const uint8_t* weights, size_t nbWeights, int x4)
{
    uint8_t len[256];
    uint16_t val[256];
    memset(len, 0, sizeof(len));
    uint32_t total = 0;
    for (size_t s = 0; s < nbWeights; s++)
        if (weights[s]) total += 1u << (weights[s] - 1);
    if (total == 0 || (total & (total - 1))) return -1;
    unsigned tableLog = (unsigned)highbit32(total);
    unsigned perRank[14] = { 0 };
    for (size_t s = 0; s < nbWeights; s++)
        if (weights[s]) { len[s] = (uint8_t)(tableLog + 1 - weights[s]); perRank[len[s]]++; }
    /* canonical values: same assignment as HUF_buildCTableFromTree */
    uint16_t start[14] = { 0 };
    {
        uint16_t mn = 0;
        for (int r = (int)tableLog; r > 0; r--) { start[r] = mn; mn = (uint16_t)((mn + perRank[r]) >> 1); }
    }
    for (size_t s = 0; s < nbWeights; s++)
        if (len[s]) val[s] = start[len[s]]++;
    if (!x4) return huf_decode_stream(dst, n, src, size, len, val, tableLog);
    if (size < 6) return -1;
    size_t seg = (n + 3) / 4;
    size_t sz[4];
    sz[0] = src[0] | (src[1] << 8);
    sz[1] = src[2] | (src[3] << 8);
    sz[2] = src[4] | (src[5] << 8);
    if (6 + sz[0] + sz[1] + sz[2] > size) return -1;
    sz[3] = size - 6 - sz[0] - sz[1] - sz[2];
    size_t ip = 6, op = 0;
    for (int k = 0; k < 4; k++) {
        size_t cnt = k < 3 ? seg : n - 3 * seg;
        if (huf_decode_stream(dst + op, cnt, src + ip, sz[k], len, val, tableLog)) return -1;
        ip += sz[k];
        op += cnt;
    }
    return 0;
}

/* ===================== 2-byte entropy path ===================== */

/* src/openzl/shared/histogram.c:37-82 ZS_Histogram_count16: 65,536 counts,
 * the largest value present and the number of distinct values. */
void orc_histogram16(uint32_t* count, unsigned* maxSym, unsigned* card, const uint16_t* src, size_t n)
{
    memset(count, 0, 65536 * sizeof(uint32_t));
    unsigned mx = 0, cd = 0;
    for (size_t i = 0; i < n; i++) {
        if (count[src[i]]++ == 0) cd++;
        if (src[i] > mx) mx = src[i];
    }
    *maxSym = mx;
    *card = cd;
}

/* src/openzl/codecs/entropy/encode_entropy_binding.c:612-682, the choice for a
 * struct chunk of 2-byte elements inside ZL_GRAPH_HUFFMAN: STORE for <= 1
 * element, CONSTANT, BITPACK (via interpret-as-LE) when the packed size beats
 * both the entropy estimate plus header and the raw size, STORE when the
 * estimate is not below the raw size or the entropy exceeds 15 bits,
 * TOKENIZE below 256 distinct values, else huffman_struct_v2. */
int orc_entropy2_decide(const uint16_t* src, size_t n, double* entropyOut)
{
    static uint32_t count[65536];
    if (n <= 1) return ORC_E2_STORE;
    unsigned mx, card;
    orc_histogram16(count, &mx, &card, src, n);
    if (count[mx] == n) return ORC_E2_CONSTANT;
    double e = orc_entropy(count, mx, n);
    e = 1 > e ? 1 : e;
    if (entropyOut) *entropyOut = e;
    size_t entropySize = (size_t)(e * (double)n + 7) / 8;
    size_t hdr = mx / 4 > 100 ? mx / 4 : 100;
    size_t huf = entropySize + hdr, store = 2 * n;
    size_t nbBits = (size_t)orc_next_pow2((uint64_t)mx + 1);
    size_t bp = (n * nbBits + 7) / 8 + 2;
    if (bp <= huf && bp < store) return ORC_E2_BITPACK;
    if (e > 15 || huf >= store) return ORC_E2_STORE;
    if (card < 256) return ORC_E2_TOKENIZE;
    return ORC_E2_HUF16;
}

/* src/openzl/codecs/tokenize/encode_tokenize_binding.c:124-216 (tokenize of
 * a struct stream, never sorted): the alphabet lists the distinct values in
 * order of first occurrence, each element becomes its alphabet index (1 byte
 * wide for alphabets of at most 256 values). Returns the alphabet size, or 0
 * when it exceeds 256. */
size_t orc_tokenize2(uint16_t* alphabet, uint8_t* idx, const uint16_t* src, size_t n)
{
    static int32_t slot[65536];
    memset(slot, -1, sizeof(slot));
    size_t card = 0;
    for (size_t i = 0; i < n; i++) {
        if (slot[src[i]] < 0) {
            if (card == 256) return 0;
            slot[src[i]] = (int32_t)card;
            alphabet[card++] = src[i];
        }
        idx[i] = (uint8_t)slot[src[i]];
    }
    return card;
}

/* ---- 16-bit Huffman (src/openzl/codecs/entropy/encode_huffman_kernel.c) ---- */
typedef struct { uint32_t count, parent; uint16_t symbol, nbBits; } h16node;

static int h16_by_count(const void* a, const void* b)
{
    const h16node* x = (const h16node*)a;
    const h16node* y = (const h16node*)b;
    if (x->count != y->count) return x->count > y->count ? -1 : 1;
    return x->symbol < y->symbol ? -1 : x->symbol > y->symbol;
}

/* :36-140 ZS_largeHuffmanSetMaxHeight: the depths above maxNbBits are cut
 * down and the Kraft debt is repaid by deepening the cheapest shallower
 * symbols, then any overshoot is given back. Returns the final height. */
static int h16_set_max_height(h16node* node, int lastNonNull, int maxNbBits)
{
    const int largestBits = node[lastNonNull].nbBits;
    if (largestBits <= maxNbBits) return largestBits;
    int cost = 0;
    const int base = 1 << (largestBits - maxNbBits);
    int n = lastNonNull;
    for (; node[n].nbBits > maxNbBits; n--) {
        cost += base - (1 << (largestBits - node[n].nbBits));
        node[n].nbBits = (uint16_t)maxNbBits;
    }
    while (node[n].nbBits == maxNbBits) n--;
    cost >>= largestBits - maxNbBits;
    const uint32_t none = 0xF0F0F0F0u;
    uint32_t last[22];
    for (int r = 0; r < 22; r++) last[r] = none;
    {
        int cur = maxNbBits;
        for (int pos = n; pos >= 0; pos--) {
            if (node[pos].nbBits >= cur) continue;
            cur = node[pos].nbBits;
            last[maxNbBits - cur] = (uint32_t)pos;
        }
    }
    while (cost > 0) {
        int k = highbit32((uint32_t)cost) + 1;
        while (k > 1 && last[k] == none) k--;
        for (; k > 1; k--) {
            const uint32_t hi = last[k], lo = last[k - 1];
            if (lo == none) break;
            if (node[hi].count <= 2 * node[lo].count) break;
        }
        while (k <= 20 && last[k] == none) k++;
        cost -= 1 << (k - 1);
        if (last[k - 1] == none) last[k - 1] = last[k];
        node[last[k]].nbBits++;
        if (last[k] == 0) {
            last[k] = none;
        } else {
            last[k]--;
            if (node[last[k]].nbBits != maxNbBits - k) last[k] = none;
        }
    }
    while (cost < 0) {
        if (last[1] == none) {
            while (node[n].nbBits == maxNbBits) n--;
            node[n + 1].nbBits--;
            last[1] = (uint32_t)(n + 1);
            cost++;
            continue;
        }
        node[last[1] + 1].nbBits--;
        last[1]++;
        cost++;
    }
    return maxNbBits;
}

/* :274-329 ZS_largeHuffmanBuildCTable(ctable, count, maxSym, 0): maxNbBits 20
 * lowered to nextPow2(maxSym + 1) + 3; symbols sorted by decreasing count
 * (equal counts keep symbol order, :145-229); two-queue tree (:231-273);
 * height limit; canonical values per length in symbol order, in 16-bit
 * arithmetic as the reference's uint16_t counters (:275-301). code[s] and
 * nbBits[s] for s <= maxSym (nbBits 0 for absent symbols). Returns the table
 * log (the largest code length), or -1. */
int orc_huf16_build(uint16_t* code, uint8_t* nbBits, const uint32_t* count, unsigned maxSym)
{
    const int S = (int)maxSym + 1;
    h16node* tab = (h16node*)calloc((size_t)(2 * S + 3), sizeof(h16node));
    if (!tab) return -1;
    h16node* node = tab + 1;
    int maxNbBits = 20;
    {
        int bound = orc_next_pow2((uint64_t)S);
        if (bound + 3 < maxNbBits) maxNbBits = bound + 3;
    }
    for (int s = 0; s < S; s++) { node[s].count = count[s]; node[s].symbol = (uint16_t)s; }
    qsort(node, (size_t)S, sizeof(h16node), h16_by_count);
    int nonNull = S - 1;
    while (node[nonNull].count == 0) nonNull--;
    if (nonNull < 1) { free(tab); return -1; }
    {
        int nb = S, lowS = nonNull, lowN = nb;
        const int root = nb + lowS - 1;
        node[nb].count = node[lowS].count + node[lowS - 1].count;
        node[lowS].parent = node[lowS - 1].parent = (uint32_t)nb;
        nb++;
        lowS -= 2;
        for (int k = nb; k <= root; k++) node[k].count = 1u << 30;
        node[-1].count = 1u << 31;
        for (; nb <= root; nb++) {
            int a = node[lowS].count < node[lowN].count ? lowS-- : lowN++;
            int b = node[lowS].count < node[lowN].count ? lowS-- : lowN++;
            node[nb].count = node[a].count + node[b].count;
            node[a].parent = node[b].parent = (uint32_t)nb;
        }
        node[root].nbBits = 0;
        for (int k = root - 1; k >= S; k--) node[k].nbBits = (uint16_t)(node[node[k].parent].nbBits + 1);
        for (int k = 0; k <= nonNull; k++) node[k].nbBits = (uint16_t)(node[node[k].parent].nbBits + 1);
    }
    maxNbBits = h16_set_max_height(node, nonNull, maxNbBits);
    if (maxNbBits > 20) { free(tab); return -1; }
    {
        uint16_t perRank[21];
        uint16_t start[21];
        memset(perRank, 0, sizeof(perRank));
        memset(start, 0, sizeof(start));
        for (int k = 0; k <= nonNull; k++) perRank[node[k].nbBits]++;
        uint16_t mn = 0;
        for (int r = maxNbBits; r > 0; r--) { start[r] = mn; mn = (uint16_t)(mn + perRank[r]); mn >>= 1; }
        for (int s = 0; s < S; s++) nbBits[s] = 0;
        for (int k = 0; k < S; k++) nbBits[node[k].symbol] = (uint8_t)node[k].nbBits;
        for (int s = 0; s < S; s++) code[s] = start[nbBits[s]]++;
    }
    free(tab);
    return maxNbBits;
}

/* :331-452 ZS_largeHuffmanEncodeUsingCTable / X4 as called by
 * EI_huffman_struct_v2 (encode_entropy_binding.c:286-367) with a capacity of
 * 2n + 32 bytes: each stream is [LE32 element count][LE32 byte size][bits],
 * the bits written from the stream's last element to its first, LSB-first,
 * closed by a 1 bit; X4 cuts streams of n/4 + 1 elements (the last gets the
 * rest). A stream that does not fit its remaining capacity fails
 * (bitstream.h:154-245: at least 9 bytes of capacity, and the written bytes
 * stay 8 bytes below the end). Returns the total size or 0. */
size_t orc_huf16_encode(uint8_t* dst, size_t cap, const uint16_t* src, size_t n, const uint16_t* code,
                        const uint8_t* nbBits, int x4)
{
    size_t pos = 0, off = 0;
    const int ns = x4 ? 4 : 1;
    const size_t seg = n / 4 + 1;
    for (int k = 0; k < ns; k++) {
        size_t cnt = x4 ? (seg < n - off ? seg : n - off) : n;
        if (cap - pos < 8) return 0;
        size_t avail = cap - pos - 8;
        if (avail <= 8) return 0;
        uint8_t* b = dst + pos + 8;
        memset(b, 0, avail);
        uint64_t bit = 0;
        for (size_t i = cnt; i-- > 0;) {
            uint16_t s = src[off + i];
            unsigned nb = nbBits[s];
            uint32_t c = (uint32_t)code[s] & ((1u << nb) - 1);
            for (unsigned q = 0; q < nb; q++, bit++)
                if ((c >> q) & 1) {
                    if ((bit >> 3) >= avail) return 0;
                    b[bit >> 3] |= (uint8_t)(1u << (bit & 7));
                }
        }
        /* end mark; the stream's full bytes must stay below its end */
        if ((bit >> 3) >= avail) return 0;
        b[bit >> 3] |= (uint8_t)(1u << (bit & 7));
        bit++;
        if ((bit >> 3) >= avail - 8) return 0;
        size_t bytes = (size_t)((bit + 7) >> 3);
        for (int q = 0; q < 4; q++) {
            dst[pos + q] = (uint8_t)(cnt >> (8 * q));
            dst[pos + 4 + q] = (uint8_t)(bytes >> (8 * q));
        }
        pos += 8 + bytes;
        off += cnt;
    }
    return pos;
}

/* Decoder of the streams above with the table of the weights
 * (src/openzl/codecs/entropy/decode_entropy_binding.c:348-472 buildHUF16DTable
 * and DI_huffman_struct_v2, decode_huffman_kernel.c:13-254): weights of at
 * least two symbols, each <= 20, whose 2^(w-1) sum to a power of two (the
 * table log); length = tableLog + 1 - w; symbols of one weight in symbol
 * order. Streams are read from their end mark down; bits below a stream's
 * start read as zero. Returns 0, or -1 on corruption. */
int orc_huf16_decode(uint16_t* dst, size_t n, const uint8_t* src, size_t size, const uint8_t* weights,
                     size_t nbWeights, int x4)
{
    if (nbWeights < 2 || nbWeights > 65536) return -1;
    uint64_t sum = 0;
    size_t nz = 0;
    for (size_t s = 0; s < nbWeights; s++) {
        if (weights[s] > 20) return -1;
        if (weights[s]) { sum += (1ull << weights[s]) >> 1; nz++; }
    }
    if (nz < 2 || (sum & (sum - 1))) return -1;
    const int tl = highbit64(sum);
    if (tl > 20) return -1;
    uint32_t* dt = (uint32_t*)malloc(sizeof(uint32_t) << tl);   /* symbol | nbBits << 16 */
    if (!dt) return -1;
    {
        uint32_t start[21] = { 0 };
        for (size_t s = 0; s < nbWeights; s++) start[weights[s]]++;
        uint32_t next = 0;
        for (int w = 1; w <= tl; w++) { uint32_t c = start[w]; start[w] = next; next += c << (w - 1); }
        for (size_t s = 0; s < nbWeights; s++) {
            const int w = weights[s];
            if (!w) continue;
            const uint32_t len = 1u << (w - 1);
            for (uint32_t i = 0; i < len; i++) dt[start[w] + i] = (uint32_t)s | (uint32_t)(tl + 1 - w) << 16;
            start[w] += len;
        }
    }
    size_t ip = 0, op = 0;
    const int ns = x4 ? 4 : 1;
    int rc = 0;
    for (int k = 0; k < ns && !rc; k++) {
        if (size - ip < 8) { rc = -1; break; }
        uint32_t cnt = 0, bytes = 0;
        for (int q = 0; q < 4; q++) {
            cnt |= (uint32_t)src[ip + q] << (8 * q);
            bytes |= (uint32_t)src[ip + 4 + q] << (8 * q);
        }
        ip += 8;
        if (bytes > size - ip || cnt > n - op || bytes == 0 || src[ip + bytes - 1] == 0) { rc = -1; break; }
        const uint8_t* b = src + ip;
        int64_t bit = (int64_t)(bytes - 1) * 8 + highbit32(b[bytes - 1]);
        for (uint32_t i = 0; i < cnt; i++) {
            uint32_t v = 0;
            for (int q = 0; q < tl; q++) {
                int64_t p = bit - 1 - q;
                v = (v << 1) | (p >= 0 ? (uint32_t)((b[p >> 3] >> (p & 7)) & 1) : 0u);
            }
            dst[op + i] = (uint16_t)(dt[v] & 0xFFFF);
            bit -= (int64_t)(dt[v] >> 16);
        }
        ip += bytes;
        op += cnt;
    }
    if (!rc && (op != n || ip != size)) rc = -1;
    free(dt);
    return rc;
}

/* ============================ XXH3-64 ============================ */
/* Restated from the published XXH3 algorithm (xxHash specification, as vendored
 * at src/openzl/shared/xxhash.h; default secret at :4403-4416). seed = 0. */
static const uint8_t kSecret[192] = {
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c,
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f,
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21,
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c,
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3,
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8,
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d,
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64,
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb,
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e,
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce,
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
};
#define P32_1 0x9E3779B1u
#define P32_2 0x85EBCA77u
#define P32_3 0xC2B2AE3Du
#define P64_1 0x9E3779B185EBCA87ull
#define P64_2 0xC2B2AE3D27D4EB4Full
#define P64_3 0x165667B19E3779F9ull
#define P64_4 0x85EBCA77C2B2AE63ull
#define P64_5 0x27D4EB2F165667C5ull
#define PMX1 0x165667919E3779F9ull
#define PMX2 0x9FB21C651E98DF25ull

static uint64_t r64(const uint8_t* p) { uint64_t v; memcpy(&v, p, 8); return v; }
static uint32_t r32(const uint8_t* p) { uint32_t v; memcpy(&v, p, 4); return v; }
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
static uint64_t fold128(uint64_t a, uint64_t b)
{
    unsigned __int128 m = (unsigned __int128)a * b;
    return (uint64_t)m ^ (uint64_t)(m >> 64);
}
static uint64_t avalanche64(uint64_t h)
{
    h ^= h >> 33; h *= P64_2; h ^= h >> 29; h *= P64_3; h ^= h >> 32; return h;
}
static uint64_t avalanche3(uint64_t h) { h ^= h >> 37; h *= PMX1; h ^= h >> 32; return h; }
static uint64_t rrmxmx(uint64_t h, uint64_t len)
{
    h ^= rotl64(h, 49) ^ rotl64(h, 24);
    h *= PMX2;
    h ^= (h >> 35) + len;
    h *= PMX2;
    h ^= h >> 28;
    return h;
}
static uint64_t mix16(const uint8_t* in, const uint8_t* sec)
{
    return fold128(r64(in) ^ r64(sec), r64(in + 8) ^ r64(sec + 8));
}
static void acc512(uint64_t acc[8], const uint8_t* in, const uint8_t* sec)
{
    for (int i = 0; i < 8; i++) {
        uint64_t d = r64(in + 8 * i);
        uint64_t k = d ^ r64(sec + 8 * i);
        acc[i ^ 1] += d;
        acc[i] += (uint64_t)(uint32_t)k * (k >> 32);
    }
}

uint64_t orc_xxh3_64(const void* data, size_t len)
{
    const uint8_t* in = (const uint8_t*)data;
    const uint8_t* sec = kSecret;
    if (len == 0) return avalanche64(r64(sec + 56) ^ r64(sec + 64));
    if (len <= 3) {
        uint32_t c1 = in[0], c2 = in[len >> 1], c3 = in[len - 1];
        uint32_t comb = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
        uint64_t flip = (uint64_t)(r32(sec) ^ r32(sec + 4));
        return avalanche64((uint64_t)comb ^ flip);
    }
    if (len <= 8) {
        uint32_t i1 = r32(in), i2 = r32(in + len - 4);
        uint64_t flip = r64(sec + 8) ^ r64(sec + 16);
        uint64_t v = i2 + ((uint64_t)i1 << 32);
        return rrmxmx(v ^ flip, len);
    }
    if (len <= 16) {
        uint64_t f1 = r64(sec + 24) ^ r64(sec + 32);
        uint64_t f2 = r64(sec + 40) ^ r64(sec + 48);
        uint64_t lo = r64(in) ^ f1, hi = r64(in + len - 8) ^ f2;
        uint64_t acc = len + __builtin_bswap64(lo) + hi + fold128(lo, hi);
        return avalanche3(acc);
    }
    if (len <= 128) {
        uint64_t acc = len * P64_1;
        if (len > 32) {
            if (len > 64) {
                if (len > 96) { acc += mix16(in + 48, sec + 96); acc += mix16(in + len - 64, sec + 112); }
                acc += mix16(in + 32, sec + 64); acc += mix16(in + len - 48, sec + 80);
            }
            acc += mix16(in + 16, sec + 32); acc += mix16(in + len - 32, sec + 48);
        }
        acc += mix16(in, sec); acc += mix16(in + len - 16, sec + 16);
        return avalanche3(acc);
    }
    if (len <= 240) {
        uint64_t acc = len * P64_1;
        size_t rounds = len / 16;
        for (size_t i = 0; i < 8; i++) acc += mix16(in + 16 * i, sec + 16 * i);
        acc = avalanche3(acc);
        for (size_t i = 8; i < rounds; i++) acc += mix16(in + 16 * i, sec + 16 * (i - 8) + 3);
        acc += mix16(in + len - 16, sec + 136 - 17);
        return avalanche3(acc);
    }
    uint64_t acc[8] = { P32_3, P64_1, P64_2, P64_3, P64_4, P32_2, P64_5, P32_1 };
    const size_t stripesPerBlock = (192 - 64) / 8;
    const size_t blockLen = 64 * stripesPerBlock;
    size_t nbBlocks = (len - 1) / blockLen;
    for (size_t b = 0; b < nbBlocks; b++) {
        for (size_t s = 0; s < stripesPerBlock; s++) acc512(acc, in + b * blockLen + s * 64, sec + s * 8);
        for (int i = 0; i < 8; i++) {
            uint64_t a = acc[i];
            a ^= a >> 47;
            a ^= r64(sec + 192 - 64 + 8 * i);
            a *= P32_1;
            acc[i] = a;
        }
    }
    size_t nbStripes = ((len - 1) - blockLen * nbBlocks) / 64;
    for (size_t s = 0; s < nbStripes; s++) acc512(acc, in + nbBlocks * blockLen + s * 64, sec + s * 8);
    acc512(acc, in + len - 64, sec + 192 - 64 - 7);
    uint64_t res = len * P64_1;
    for (int i = 0; i < 4; i++) res += fold128(acc[2 * i] ^ r64(sec + 11 + 16 * i), acc[2 * i + 1] ^ r64(sec + 11 + 16 * i + 8));
    return avalanche3(res);
}
