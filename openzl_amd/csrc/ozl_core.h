// ozl_core.h -- per-chunk table construction, decisions and small-stream
// coders of the entropy hot path, written once for the device (called by one
// lane of a workgroup, tables in LDS or global memory) and compiled for the
// host only by the CPU unit build used in tests.
//
// Everything here is integer arithmetic except ozl_entropy(), which ends with
// the reference's single double division, and the double size estimates of
// ozl_decide(); the build uses -ffp-contract=off so products and sums round
// exactly as in the reference C code.
#pragma once
#include <stddef.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OZL_HD __host__ __device__ __forceinline__
#else
#define OZL_HD static inline
#endif

// Constant tables exist twice under hipcc: a __constant__ copy for device
// code and a plain copy for host code (the host-side shadow of a __constant__
// variable does not hold its initializer). OZL_TAB(name) picks the copy of
// the side being compiled.
#if defined(__HIPCC__) && defined(__HIP_DEVICE_COMPILE__)
#define OZL_TAB(name) name##_dev
#else
#define OZL_TAB(name) name##_host
#endif

// ------------------------------------------------------------------ limits
#define OZL_FSE_MAX_TABLELOG 12
#define OZL_FSE_MIN_TABLELOG 5
#define OZL_FSE_DEFAULT_TABLELOG 11
#define OZL_HUF_DEFAULT_TABLELOG 11

enum { OZL_BK_STORE = 0, OZL_BK_CONSTANT = 1, OZL_BK_FLATPACK = 2, OZL_BK_BITPACK = 3, OZL_BK_HUF = 4,
       OZL_BK_FSE = 5 };

OZL_HD int ozl_hb32(uint32_t v) { return 31 - __builtin_clz(v); }
OZL_HD int ozl_hb64(uint64_t v) { return 63 - __builtin_clzll(v); }
// smallest n with (1 << n) >= x   (src/openzl/shared/bits.h:230)
OZL_HD int ozl_next_pow2(uint64_t x) { return x <= 1 ? 0 : 64 - __builtin_clzll(x - 1); }

OZL_HD size_t ozl_varint(uint64_t v, uint8_t* out)
{
    size_t n = 0;
    while (v >= 0x80) { out[n++] = (uint8_t)(v | 0x80); v >>= 7; }
    out[n++] = (uint8_t)v;
    return n;
}
OZL_HD size_t ozl_varint_size(uint64_t v)
{
    size_t n = 1;
    while (v >= 0x80) { n++; v >>= 7; }
    return n;
}

// floor(log2(1 + i/256) * 2^16) -- src/openzl/shared/data_stats.c:15-44.
// Generated from the closed form with exact integer arithmetic.
static const uint16_t ozl_log2_tab_host[256] = {
#include "ozl_log2_table.inc"
};
#if defined(__HIPCC__)
static __device__ __constant__ uint16_t ozl_log2_tab_dev[256] = {
#include "ozl_log2_table.inc"
};
#endif

// ------------------------------------------------------------------ stats
// One symbol's term of ZL_calculateEntropy (data_stats.c:46-73): the int64
// sum of these over all symbols is order independent, so lanes may compute
// terms in parallel and reduce.
OZL_HD int64_t ozl_entropy_term(uint32_t count, int64_t norm)
{
    int64_t c = count;
    int64_t li = c * norm;
    int clz = __builtin_clzll((uint64_t)li | (1u << 9));
    int shift = 55 - clz;
    li >>= shift;
    li &= 255;
    return -c * ((int64_t)OZL_TAB(ozl_log2_tab)[li] - (int64_t)clz * 65536);
}
OZL_HD double ozl_entropy_finish(int64_t sum, size_t total)
{
    return (double)sum / (double)(total * 65536) - 1;
}
OZL_HD double ozl_entropy(const uint32_t* count, unsigned maxValue, size_t total)
{
    if (total == 0) return 0;
    int64_t norm = ((int64_t)1 << 62) / (int64_t)total;
    int64_t acc = 0;
    for (unsigned s = 0; s <= maxValue; s++)
        if (count[s]) acc += ozl_entropy_term(count[s], norm);  // a zero count's term is 0
    return ozl_entropy_finish(acc, total);
}

// entropyCompressChunk's choice for a serial chunk of n >= 2 bytes with
// cardinality >= 2 (encode_entropy_binding.c:684-752) given its entropy.
OZL_HD int ozl_decide(double h, unsigned maxSym, unsigned card, size_t n, int hufMode)
{
    size_t entropySize;
    if (hufMode) {  // DataStatsU8_estimateHuffmanSizeFast, data_stats.c:288-299
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
    size_t flat = (n * (size_t)ozl_next_pow2(card) + 7) / 8 + card;          // data_stats.c:320-327
    size_t bitp = (n * (size_t)ozl_next_pow2((uint64_t)maxSym + 1) + 7) / 8;  // data_stats.c:301-312
    if (flat < bitp) {
        if (flat < baseline) return OZL_BK_FLATPACK;
    } else {
        if (bitp < baseline) return OZL_BK_BITPACK;
    }
    if (n <= baseline) return OZL_BK_STORE;
    return hufMode ? OZL_BK_HUF : OZL_BK_FSE;
}

// Full decision for a serial chunk from its histogram.
OZL_HD int ozl_decide_from_hist(const uint32_t* count, size_t n, int hufMode, unsigned* maxSymOut,
                                unsigned* cardOut)
{
    unsigned maxSym = 0, card = 0;
    for (unsigned s = 0; s < 256; s++)
        if (count[s]) { maxSym = s; card++; }
    *maxSymOut = maxSym;
    *cardOut = card;
    if (n <= 1) return OZL_BK_STORE;
    if (card == 1) return OZL_BK_CONSTANT;
    return ozl_decide(ozl_entropy(count, maxSym, n), maxSym, card, n, hufMode);
}
// The same decision from the histogram's statistics: maxSym, cardinality and
// the int64 sum of ozl_entropy_term over the symbols (computed in parallel).
OZL_HD int ozl_decide_from_stats(int64_t termSum, size_t n, unsigned maxSym, unsigned card, int hufMode)
{
    if (n <= 1) return OZL_BK_STORE;
    if (card == 1) return OZL_BK_CONSTANT;
    return ozl_decide(ozl_entropy_finish(termSum, n), maxSym, card, n, hufMode);
}

// ------------------------------------------------------------------ FSE
// src/openzl/fse/compress/fse_compress.c:361-396
OZL_HD unsigned ozl_fse_min_tablelog(size_t n, unsigned maxSym)
{
    unsigned a = (unsigned)ozl_hb32((uint32_t)n) + 1, b = (unsigned)ozl_hb32(maxSym) + 2;
    return a < b ? a : b;
}
OZL_HD unsigned ozl_fse_optimal_tablelog(unsigned maxTableLog, size_t n, unsigned maxSym, unsigned minus)
{
    unsigned maxBitsSrc = (unsigned)ozl_hb32((uint32_t)(n - 1)) - minus;
    unsigned tl = maxTableLog ? maxTableLog : OZL_FSE_DEFAULT_TABLELOG;
    unsigned minBits = ozl_fse_min_tablelog(n, maxSym);
    if (maxBitsSrc < tl) tl = maxBitsSrc;
    if (minBits > tl) tl = minBits;
    if (tl < OZL_FSE_MIN_TABLELOG) tl = OZL_FSE_MIN_TABLELOG;
    if (tl > OZL_FSE_MAX_TABLELOG) tl = OZL_FSE_MAX_TABLELOG;
    return tl;
}

// FSE_normalizeM2, fse_compress.c:401-485 (fallback for skewed histograms)
OZL_HD int ozl_fse_normalize_m2(int16_t* norm, unsigned tl, const uint32_t* count, size_t total, unsigned maxSym)
{
    const int16_t UNSET = -2;
    uint32_t distributed = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tl);
    uint32_t lowOne = (uint32_t)((total * 3) >> (tl + 1));
    for (unsigned s = 0; s <= maxSym; s++) {
        uint32_t c = count[s];
        if (c == 0) { norm[s] = 0; continue; }
        if (c <= lowThreshold) { norm[s] = -1; distributed++; total -= c; continue; }
        if (c <= lowOne) { norm[s] = 1; distributed++; total -= c; continue; }
        norm[s] = UNSET;
    }
    uint32_t toDist = (1u << tl) - distributed;
    if (toDist == 0) return 0;
    if ((total / toDist) > lowOne) {
        lowOne = (uint32_t)((total * 3) / (toDist * 2));
        for (unsigned s = 0; s <= maxSym; s++)
            if (norm[s] == UNSET && count[s] <= lowOne) { norm[s] = 1; distributed++; total -= count[s]; }
        toDist = (1u << tl) - distributed;
    }
    if (distributed == maxSym + 1) {
        uint32_t maxV = 0, maxC = 0;
        for (unsigned s = 0; s <= maxSym; s++)
            if (count[s] > maxC) { maxV = s; maxC = count[s]; }
        norm[maxV] = (int16_t)(norm[maxV] + (int16_t)toDist);
        return 0;
    }
    if (total == 0) {
        for (unsigned s = 0; toDist > 0; s = (s + 1) % (maxSym + 1))
            if (norm[s] > 0) { toDist--; norm[s]++; }
        return 0;
    }
    uint64_t vStepLog = 62 - tl;
    uint64_t mid = (1ull << (vStepLog - 1)) - 1;
    uint64_t rStep = (((1ull << vStepLog) * toDist) + mid) / (uint32_t)total;
    uint64_t acc = mid;
    for (unsigned s = 0; s <= maxSym; s++) {
        if (norm[s] == UNSET) {
            uint64_t end = acc + count[s] * rStep;
            uint32_t a = (uint32_t)(acc >> vStepLog), b = (uint32_t)(end >> vStepLog);
            if (b - a < 1) return -1;
            norm[s] = (int16_t)(b - a);
            acc = end;
        }
    }
    return 0;
}

// FSE_normalizeCount with useLowProbCount=1, fse_compress.c:487-549
// rtbTable of FSE_normalizeCount (selects, not an indexed local array)
OZL_HD uint32_t ozl_fse_rtb(int p)
{
    return p == 1 ? 473195u : p == 2 ? 504333u : p == 3 ? 520860u : p == 4 ? 550000u
         : p == 5 ? 700000u : p == 6 ? 750000u : p == 7 ? 830000u : 0u;
}

OZL_HD int ozl_fse_normalize(int16_t* norm, unsigned tl, const uint32_t* count, size_t total, unsigned maxSym)
{
    if (tl < OZL_FSE_MIN_TABLELOG || tl > OZL_FSE_MAX_TABLELOG) return -1;
    if (tl < ozl_fse_min_tablelog(total, maxSym)) return -1;
    uint64_t scale = 62 - tl;
    uint64_t step = (1ull << 62) / (uint32_t)total;
    uint64_t vStep = 1ull << (scale - 20);
    int still = 1 << tl;
    unsigned largest = 0;
    int16_t largestP = 0;
    uint32_t lowThreshold = (uint32_t)(total >> tl);
    for (unsigned s = 0; s <= maxSym; s++) {
        uint32_t c = count[s];
        if (c == total) return 1;  // constant input: nothing to normalize
        if (c == 0) { norm[s] = 0; continue; }
        if (c <= lowThreshold) { norm[s] = -1; still--; continue; }
        int16_t p = (int16_t)((c * step) >> scale);
        if (p < 8) {
            uint64_t restToBeat = vStep * ozl_fse_rtb(p);
            p = (int16_t)(p + ((c * step) - ((uint64_t)p << scale) > restToBeat));
        }
        if (p > largestP) { largestP = p; largest = s; }
        norm[s] = p;
        still -= p;
    }
    if (-still >= (norm[largest] >> 1)) return ozl_fse_normalize_m2(norm, tl, count, total, maxSym);
    norm[largest] = (int16_t)(norm[largest] + still);
    return 0;
}

OZL_HD unsigned ozl_fse_step(unsigned ts) { return (ts >> 1) + (ts >> 3) + 3; }

// Symbol spread shared by encoder and decoder tables (fse_compress.c:93-155,
// fse_decompress.c:93-160): low-probability symbols take the top cells in
// symbol order, the others follow the fixed step, skipping the top area.
OZL_HD int ozl_fse_spread(uint8_t* cellSym, const int16_t* norm, unsigned maxSym, unsigned tl)
{
    unsigned ts = 1u << tl, mask = ts - 1, step = ozl_fse_step(ts), high = ts - 1;
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

// Packed per-symbol encoding transform: bits 0-3 maxBitsOut, 4-17
// minStatePlus (<= 2^13), 18-31 deltaFindState + 4096 (deltaFindState lies in
// [-2^12, 2^12) for tableLog <= 12). deltaNbBits = (maxBitsOut<<16) -
// minStatePlus (fse_compress.c:160-185).
OZL_HD uint32_t ozl_tt_pack(unsigned mb, unsigned msp, int dfs) { return mb | (msp << 4) | ((uint32_t)(dfs + 4096) << 18); }
OZL_HD uint32_t ozl_tt_dnb(uint32_t p) { return ((p & 15) << 16) - ((p >> 4) & 0x3FFF); }
OZL_HD int ozl_tt_dfs(uint32_t p) { return (int)(p >> 18) - 4096; }

// FSE_buildCTable_wksp, fse_compress.c:69-215. stateTable[1<<tl], tt[maxSym+1],
// cellSym is scratch of 1<<tl bytes, cumul of maxSym+2 entries.
OZL_HD int ozl_fse_build_ctable(uint16_t* stateTable, uint32_t* tt, uint8_t* cellSym, const int16_t* norm,
                                unsigned maxSym, unsigned tl, uint16_t* cumul)
{
    unsigned ts = 1u << tl;
    if (ozl_fse_spread(cellSym, norm, maxSym, tl)) return -1;
    cumul[0] = 0;
    for (unsigned s = 0; s <= maxSym; s++) cumul[s + 1] = (uint16_t)(cumul[s] + (norm[s] == -1 ? 1 : norm[s]));
    for (unsigned s = 0; s <= maxSym; s++) {
        int c = norm[s];
        if (c == 0) tt[s] = ozl_tt_pack(tl + 1, ts, -1);
        else if (c == -1 || c == 1) tt[s] = ozl_tt_pack(tl, ts, (int)cumul[s] - 1);
        else {
            unsigned mb = tl - (unsigned)ozl_hb32((uint32_t)c - 1);
            tt[s] = ozl_tt_pack(mb, (unsigned)c << mb, (int)cumul[s] - c);
        }
    }
    for (unsigned u = 0; u < ts; u++) stateTable[cumul[cellSym[u]]++] = (uint16_t)(ts + u);
    return 0;
}

// Decoding table entry: symbol | nbBits << 8 | newState << 16
// (fse_decompress.c:160-172).
// `next` is workspace for maxSym+1 entries.
OZL_HD int ozl_fse_build_dtable(uint32_t* dt, uint8_t* cellSym, const int16_t* norm, unsigned maxSym, unsigned tl,
                                uint16_t* next)
{
    unsigned ts = 1u << tl;
    if (ozl_fse_spread(cellSym, norm, maxSym, tl)) return -1;
    for (unsigned s = 0; s <= maxSym; s++) next[s] = (uint16_t)(norm[s] == -1 ? 1 : norm[s]);
    for (unsigned u = 0; u < ts; u++) {
        unsigned sym = cellSym[u];
        unsigned ns = next[sym]++;
        unsigned nb = tl - (unsigned)ozl_hb32(ns);
        dt[u] = sym | (nb << 8) | (((ns << nb) - ts) << 16);
    }
    return 0;
}

// FSE_writeNCount_generic, fse_compress.c:246-357. dst needs 2*256+4 bytes.
OZL_HD size_t ozl_fse_write_ncount(uint8_t* dst, const int16_t* norm, unsigned maxSym, unsigned tl)
{
    uint8_t* out = dst;
    int ts = 1 << tl;
    uint32_t bits = (tl - OZL_FSE_MIN_TABLELOG);
    int nbits = 4;
    unsigned symbol = 0, alphabet = maxSym + 1;
    int prev0 = 0;
    int remaining = ts + 1, threshold = ts, nb = (int)tl + 1;
    while (symbol < alphabet && remaining > 1) {
        if (prev0) {
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
        int count = norm[symbol++];
        int const max = (2 * threshold - 1) - remaining;
        remaining -= count < 0 ? -count : count;
        count++;
        if (count >= threshold) count += max;
        bits += (uint32_t)count << nbits;
        nbits += nb;
        nbits -= (count < max);
        prev0 = (count == 1);
        if (remaining < 1) return 0;
        while (remaining < threshold) { nb--; threshold >>= 1; }
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

// Reader of the same format (src/openzl/fse/common/entropy_common.c
// FSE_readNCount). Returns bytes used, or -1.
OZL_HD long ozl_fse_read_ncount(int16_t* norm, unsigned* maxSymOut, unsigned* tlOut, const uint8_t* src, size_t size)
{
    // LSB-first fields through a bit buffer refilled a byte at a time (zero
    // bits past the end)
    size_t pos = 0, total = size * 8, next = 0;
    uint64_t buf = 0;
    int have = 0;
    auto rd = [&](int nb) -> uint32_t {
        while (have < nb) {
            buf |= (uint64_t)(next < size ? src[next] : 0) << have;
            next++;
            have += 8;
        }
        uint32_t v = (uint32_t)(buf & ((1ull << nb) - 1));
        buf >>= nb;
        have -= nb;
        pos += (size_t)nb;
        return v;
    };
    unsigned tl = rd(4) + OZL_FSE_MIN_TABLELOG;
    if (tl > OZL_FSE_MAX_TABLELOG) return -1;
    int remaining = (1 << tl) + 1, threshold = 1 << tl, nb = (int)tl + 1;
    unsigned symbol = 0;
    int prev0 = 0;
    for (int s = 0; s < 256; s++) norm[s] = 0;
    while (remaining > 1 && symbol <= 255) {
        if (prev0) {
            unsigned n0 = symbol;
            for (;;) {
                uint32_t r = rd(2);
                n0 += r;
                if (r != 3) break;
                if (n0 > 255) return -1;
            }
            if (n0 > 255) return -1;
            symbol = n0;
        }
        int max = (2 * threshold - 1) - remaining;
        int count;
        uint32_t low = rd(nb - 1);
        if ((int)low < max) count = (int)low;
        else {
            uint32_t hi = rd(1);
            count = (int)(low + (hi << (nb - 1)));
            if (count >= threshold) count -= max;
        }
        count--;
        remaining -= count < 0 ? -count : count;
        norm[symbol++] = (int16_t)count;
        prev0 = !count;
        while (remaining < threshold) { nb--; threshold >>= 1; }
    }
    if (remaining != 1 || pos > total || symbol == 0) return -1;
    *maxSymOut = symbol - 1;
    *tlOut = tl;
    return (long)((pos + 7) / 8);
}

// Sequential FSE encoder (fse_compress.c:607-660) for small streams and the
// CPU unit build; the large-chunk encoder is the lane-parallel kernel.
OZL_HD size_t ozl_fse_encode_seq(uint8_t* dst, size_t cap, const uint8_t* src, size_t n, const uint16_t* stateTable,
                                 const uint32_t* tt, unsigned tl, unsigned nbStates)
{
    if (n <= nbStates || nbStates < 1 || nbStates > 4 || cap <= 8) return 0;
    // state (s + j) % nbStates lives in x[j]; every step uses x0 and rotates
    // (no dynamically indexed local array)
    uint32_t x0 = 0, x1 = 0, x2 = 0, x3 = 0;
    auto rot = [&]() {
        uint32_t t = x0;
        if (nbStates == 4) { x0 = x1; x1 = x2; x2 = x3; x3 = t; }
        else if (nbStates == 3) { x0 = x1; x1 = x2; x2 = t; }
        else if (nbStates == 2) { x0 = x1; x1 = t; }
    };
    uint64_t acc = 0;
    unsigned accBits = 0;
    size_t op = 0;
    size_t p = n;
    // state indices start at nbStates - (n % nbStates): rotate into place
    for (size_t k = 0; k < (nbStates - (n % nbStates)) % nbStates; k++) rot();
    for (unsigned i = 0; i < nbStates; i++) {
        uint32_t t = tt[src[--p]];
        uint32_t dnb = ozl_tt_dnb(t);
        uint32_t nbOut = (dnb + (1u << 15)) >> 16;
        uint32_t v = (nbOut << 16) - dnb;
        x0 = stateTable[(int)(v >> nbOut) + ozl_tt_dfs(t)];
        rot();
    }
    auto put = [&](uint32_t v, unsigned nb) {
        acc |= (uint64_t)(v & ((1u << nb) - 1)) << accBits;
        accBits += nb;
        while (accBits >= 8) {
            if (op < cap) dst[op] = (uint8_t)acc;
            op++;
            acc >>= 8;
            accBits -= 8;
        }
    };
    while (p > 0) {
        uint32_t t = tt[src[--p]];
        uint32_t nbOut = (x0 + ozl_tt_dnb(t)) >> 16;
        put(x0, nbOut);
        x0 = stateTable[(int)(x0 >> nbOut) + ozl_tt_dfs(t)];
        rot();
    }
    for (unsigned i = 0; i < nbStates; i++) {
        put(x0, tl);
        rot();
    }
    put(1, 1);
    if (accBits) {
        if (op < cap) dst[op] = (uint8_t)acc;
        op++;
    }
    if (op + 8 > cap) return 0;
    return op;
}

// Sequential FSE decoder given the decoding table; symbol p is produced by
// state p mod nbStates (fse_decompress.c:255-338).
OZL_HD int ozl_fse_decode_seq(uint8_t* dst, size_t n, const uint8_t* src, size_t size, const uint32_t* dt,
                              unsigned tl, unsigned nbStates)
{
    if (size == 0 || src[size - 1] == 0) return -1;
    long pos = (long)(size - 1) * 8 + ozl_hb32(src[size - 1]);
    auto rdb = [&](unsigned nb) -> uint32_t {
        uint32_t v = 0;
        for (int b = (int)nb - 1; b >= 0; b--) {
            pos--;
            if (pos < 0) return 0xFFFFFFFFu;
            v |= (uint32_t)((src[pos >> 3] >> (pos & 7)) & 1) << b;
        }
        return v;
    };
    uint32_t st[4];
    for (unsigned j = 0; j < nbStates; j++) st[j] = rdb(tl);
    if (pos < 0) return -1;
    for (size_t p = 0; p < n; p++) {
        unsigned j = (unsigned)(p % nbStates);
        uint32_t e = dt[st[j]];
        dst[p] = (uint8_t)e;
        if (p + nbStates < n) {
            uint32_t r = rdb((e >> 8) & 255);
            if (pos < 0) return -1;
            st[j] = (e >> 16) + r;
        }
    }
    return pos == 0 ? 0 : -1;
}

// ------------------------------------------------------------------ Huffman
struct ozl_hnode {
    uint32_t count;
    uint16_t parent;
    uint8_t byte;
    uint8_t nbBits;
};

// Working memory of ozl_huf_build: the tree and the small arrays the build
// indexes at run time (kept together so that a kernel can place them in LDS).
struct ozl_huf_scratch {
    ozl_hnode tbl[2 * 256 + 2];
    uint16_t base[192], curr[192];
    int stk[32];
    uint32_t rankLast[14];
    uint16_t perRank[13], valPerRank[13];
    int nonNull;    // last leaf with a nonzero count (set by ozl_huf_build)
};

OZL_HD uint32_t ozl_huf_bucket(uint32_t c) { return c < 166 ? c : (uint32_t)ozl_hb32(c) + 158; }

// HUF_insertionSort (huf_compress.c:507-520): descending, stable.
OZL_HD void ozl_huf_insertion(ozl_hnode* a, int lo, int hi)
{
    ozl_hnode* b = a + lo;
    int size = hi - lo + 1;
    for (int i = 1; i < size; i++) {
        ozl_hnode key = b[i];
        int j = i - 1;
        while (j >= 0 && b[j].count < key.count) { b[j + 1] = b[j]; j--; }
        b[j + 1] = key;
    }
}

// HUF_simpleQuickSort (huf_compress.c:547-585) with an explicit stack. A call
// on fewer than 8 elements is an insertion sort; otherwise it partitions in a
// loop, recursing on the smaller side. Ranges handled by different calls are
// disjoint, so running them in another order gives the same array; small
// calls are therefore run at once and only large ones are stacked.
OZL_HD void ozl_huf_sort_range(ozl_hnode* a, int lo0, int hi0, int* stk)
{
    int sp = 0;
    stk[sp++] = lo0;
    stk[sp++] = hi0;
    while (sp) {
        int hi = stk[--sp];
        int lo = stk[--sp];
        if (hi - lo < 8) { ozl_huf_insertion(a, lo, hi); continue; }
        while (lo < hi) {
            uint32_t pivot = a[hi].count;
            int i = lo - 1;
            for (int j = lo; j < hi; j++)
                if (a[j].count > pivot) { i++; ozl_hnode t = a[i]; a[i] = a[j]; a[j] = t; }
            ozl_hnode t = a[i + 1]; a[i + 1] = a[hi]; a[hi] = t;
            int idx = i + 1;
            int clo, chi;  // the recursive call
            if (idx - lo < hi - idx) { clo = lo; chi = idx - 1; lo = idx + 1; }
            else { clo = idx + 1; chi = hi; hi = idx - 1; }
            if (chi - clo < 8) ozl_huf_insertion(a, clo, chi);
            else { stk[sp++] = clo; stk[sp++] = chi; }
        }
    }
}

// HUF_buildCTable_wksp (huf_compress.c:731-764): sort (:596-641), tree
// (:657-704), height limit (:352-500), canonical codes (:706-729), in phases
// that a kernel can interleave with wave-parallel versions of the sort's
// placement and of the depths (they give the same arrays).

// Sort, first part: every symbol's count and value to its place (buckets in
// descending order, symbol order within a bucket); leaves W->base / W->curr
// as the quicksort pass expects them. W->tbl and W->base start at zero.
OZL_HD void ozl_huf_place(const uint32_t* count, unsigned maxSym, ozl_huf_scratch* W)
{
    ozl_hnode* node = W->tbl + 1;
    uint16_t* base = W->base;
    uint16_t* curr = W->curr;
    for (unsigned s = 0; s <= maxSym; s++) base[ozl_huf_bucket(count[s])]++;
    for (int r = 191; r > 0; r--) base[r - 1] = (uint16_t)(base[r - 1] + base[r]);
    for (int r = 0; r < 192; r++) curr[r] = base[r];
    for (unsigned s = 0; s <= maxSym; s++) {
        uint32_t r = ozl_huf_bucket(count[s]) + 1;
        uint16_t pos = curr[r]++;
        node[pos].count = count[s];
        node[pos].byte = (uint8_t)s;
    }
}

// Sort, second part: the buckets of large counts sorted by count
// (HUF_simpleQuickSort); then the tree of the nonzero leaves. Returns the
// last nonzero leaf.
OZL_HD int ozl_huf_sort_tree(unsigned maxSym, ozl_huf_scratch* W)
{
    ozl_hnode* tbl = W->tbl;
    ozl_hnode* node = tbl + 1;
    for (int r = 166; r < 191; r++) {
        int sz = W->curr[r] - W->base[r];
        if (sz > 1) ozl_huf_sort_range(node + W->base[r], 0, sz - 1, W->stk);
    }
    const int startNode = 256;
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
    W->nonNull = nonNull;
    return nonNull;
}

// Depth of every node of the tree (root 256 + nonNull - 1).
OZL_HD void ozl_huf_depths(ozl_huf_scratch* W)
{
    ozl_hnode* node = W->tbl + 1;
    const int nonNull = W->nonNull, startNode = 256, nodeRoot = startNode + nonNull - 1;
    node[nodeRoot].nbBits = 0;
    for (int k = nodeRoot - 1; k >= startNode; k--) node[k].nbBits = (uint8_t)(node[node[k].parent].nbBits + 1);
    for (int k = 0; k <= nonNull; k++) node[k].nbBits = (uint8_t)(node[node[k].parent].nbBits + 1);
}

// Height limit (HUF_setMaxHeight) on the leaves' depths; returns the largest
// code length, or -1.
OZL_HD int ozl_huf_limit(unsigned maxNbBits, ozl_huf_scratch* W)
{
    ozl_hnode* node = W->tbl + 1;
    const int nonNull = W->nonNull;
    unsigned largest = node[nonNull].nbBits;
    unsigned target = maxNbBits;
    if (largest > target) {
        int totalCost = 0;
        const int baseCost = 1 << (largest - target);
        int n = nonNull;
        while (node[n].nbBits > target) {
            totalCost += baseCost - (1 << (largest - node[n].nbBits));
            node[n].nbBits = (uint8_t)target;
            n--;
        }
        while (node[n].nbBits == target) --n;
        totalCost >>= (largest - target);
        const uint32_t none = 0xF0F0F0F0u;
        uint32_t* rankLast = W->rankLast;
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
            unsigned nbDec = (unsigned)ozl_hb32((uint32_t)totalCost) + 1;
            for (; nbDec > 1; nbDec--) {
                uint32_t hp = rankLast[nbDec], lp = rankLast[nbDec - 1];
                if (hp == none) continue;
                if (lp == none) break;
                if (node[hp].count <= 2 * node[lp].count) break;
            }
            while (nbDec <= 12 && rankLast[nbDec] == none) nbDec++;
            if (nbDec > 12) return -1;
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
        largest = target;
    }
    if (largest > 12) return -1;
    return (int)largest;
}

// Returns the max code length or -1.
// `zeroed`: the caller has cleared W->tbl and W->base (a wave does it in
// parallel on the device). `lengthsOnly`: stop once the leaves W->tbl[1 + k]
// (k <= W->nonNull, in sorted order) hold their final code lengths; the
// caller derives nbBits and values from them (canonical codes, :706-729).
OZL_HD int ozl_huf_build(uint8_t* nbBits, uint16_t* values, const uint32_t* count, unsigned maxSym,
                         unsigned maxNbBits, ozl_huf_scratch* W, bool zeroed = false, bool lengthsOnly = false)
{
    ozl_hnode* tbl = W->tbl;
    ozl_hnode* node = tbl + 1;
    if (!zeroed) {
        for (int k = 0; k < 2 * 256 + 2; k++) { tbl[k].count = 0; tbl[k].parent = 0; tbl[k].byte = 0; tbl[k].nbBits = 0; }
        for (int r = 0; r < 192; r++) W->base[r] = 0;
    }
    ozl_huf_place(count, maxSym, W);
    const int nonNull = ozl_huf_sort_tree(maxSym, W);
    ozl_huf_depths(W);
    const int lim = ozl_huf_limit(maxNbBits, W);
    if (lim < 0) return -1;
    const unsigned largest = (unsigned)lim;
    if (lengthsOnly) return (int)largest;
    uint16_t* perRank = W->perRank;
    uint16_t* valPerRank = W->valPerRank;
    for (int r = 0; r < 13; r++) { perRank[r] = 0; valPerRank[r] = 0; }
    for (int k = 0; k <= nonNull; k++) perRank[node[k].nbBits]++;
    {
        uint16_t mn = 0;
        for (int r = (int)largest; r > 0; r--) {
            valPerRank[r] = mn;
            mn = (uint16_t)(mn + perRank[r]);
            mn >>= 1;
        }
    }
    for (unsigned s = 0; s <= maxSym; s++) nbBits[s] = 0;
    for (unsigned k = 0; k <= maxSym; k++) nbBits[node[k].byte] = node[k].nbBits;
    for (unsigned s = 0; s <= maxSym; s++) values[s] = nbBits[s] ? valPerRank[nbBits[s]]++ : 0;
    return (int)largest;
}

// Canonical code lengths/values from weights (the decoder side of
// decode_entropy_binding.c:205-284). Returns tableLog or -1.
// `rk` is workspace for 28 entries.
OZL_HD int ozl_huf_from_weights(uint8_t* len, uint16_t* val, const uint8_t* w, unsigned nbW, uint16_t* rk)
{
    uint32_t total = 0;
    for (unsigned s = 0; s < nbW; s++) {
        if (w[s] > 12) return -1;
        if (w[s]) total += 1u << (w[s] - 1);
    }
    if (total == 0 || (total & (total - 1))) return -1;
    unsigned tl = (unsigned)ozl_hb32(total);
    if (tl > 12) return -1;
    uint16_t *perRank = rk, *start = rk + 14;
    for (int r = 0; r < 14; r++) perRank[r] = start[r] = 0;
    for (unsigned s = 0; s < 256; s++) len[s] = 0;
    for (unsigned s = 0; s < nbW; s++)
        if (w[s]) { len[s] = (uint8_t)(tl + 1 - w[s]); perRank[len[s]]++; }
    uint16_t mn = 0;
    for (int r = (int)tl; r > 0; r--) { start[r] = mn; mn = (uint16_t)((mn + perRank[r]) >> 1); }
    for (unsigned s = 0; s < nbW; s++) val[s] = len[s] ? start[len[s]]++ : 0;
    return (int)tl;
}

// ------------------------------------------------------------------ small packers
// flatpack (encode_flatpack_kernel.c:8-33, 93-130); returns alphabet size.
// `map` is workspace of 256 entries.
OZL_HD size_t ozl_flatpack_encode_seq(uint8_t* alphabet, uint8_t* packed, size_t* packedSize, const uint8_t* src,
                                      size_t n, const uint32_t* count, uint8_t* map)
{
    size_t card = 0;
    for (int s = 0; s < 256; s++)
        if (count[s]) { map[s] = (uint8_t)card; alphabet[card++] = (uint8_t)s; }
    size_t nb = card <= 1 ? card : 1 + (size_t)ozl_hb32((uint32_t)card - 1);
    size_t bits = 0, op = 0;
    uint32_t acc = 0;
    for (size_t i = 0; i < n; i++) {
        acc |= (uint32_t)map[src[i]] << bits;
        bits += nb;
        if (bits >= 8) { packed[op++] = (uint8_t)acc; acc >>= 8; bits -= 8; }
    }
    acc |= 1u << bits;
    packed[op++] = (uint8_t)acc;
    *packedSize = op;
    return card;
}

// LSB-first bit packing of bytes (common_bitpack_kernel.c:131-187)
OZL_HD size_t ozl_bitpack8_seq(uint8_t* dst, const uint8_t* src, size_t n, int nbBits)
{
    size_t size = (n * (size_t)nbBits + 7) / 8, op = 0;
    uint32_t acc = 0;
    int bits = 0;
    for (size_t i = 0; i < n; i++) {
        acc |= (uint32_t)src[i] << bits;
        bits += nbBits;
        while (bits >= 8) { dst[op++] = (uint8_t)acc; acc >>= 8; bits -= 8; }
    }
    if (bits) dst[op++] = (uint8_t)acc;
    return size;
}
