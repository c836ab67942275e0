// CPU unit build of openzl_amd/csrc/ozl_core.h (test-only). The same source
// that the HIP kernels run is compiled for the host so its table builders and
// decisions can be checked against the oracle without a GPU.
#include "ozl_core.h"
#include <string.h>
#include "ozl_dparse.h"

extern "C" {
int cu_decide(const uint8_t* src, size_t n, int hufMode)
{
    uint32_t count[256] = { 0 };
    for (size_t i = 0; i < n; i++) count[src[i]]++;
    unsigned ms, card;
    return ozl_decide_from_hist(count, n, hufMode, &ms, &card);
}
double cu_entropy(const uint32_t* count, unsigned maxValue, size_t total) { return ozl_entropy(count, maxValue, total); }
int cu_fse_normalize(int16_t* norm, unsigned tl, const uint32_t* count, size_t total, unsigned maxSym)
{
    return ozl_fse_normalize(norm, tl, count, total, maxSym);
}
unsigned cu_fse_optimal_tablelog(unsigned maxTL, size_t n, unsigned maxSym, unsigned minus)
{
    return ozl_fse_optimal_tablelog(maxTL, n, maxSym, minus);
}
int cu_fse_build_ctable(uint16_t* st, uint32_t* tt, const int16_t* norm, unsigned maxSym, unsigned tl)
{
    uint8_t cells[4096];
    uint16_t cumul[257];
    return ozl_fse_build_ctable(st, tt, cells, norm, maxSym, tl, cumul);
}
size_t cu_fse_encode(uint8_t* dst, size_t cap, const uint8_t* src, size_t n, const uint16_t* st, const uint32_t* tt,
                     unsigned tl, unsigned nbStates)
{
    return ozl_fse_encode_seq(dst, cap, src, n, st, tt, tl, nbStates);
}
int cu_fse_decode(uint8_t* dst, size_t n, const uint8_t* src, size_t size, const int16_t* norm, unsigned maxSym,
                  unsigned tl, unsigned nbStates)
{
    static uint32_t dt[4096];
    uint8_t cells[4096];
    uint16_t next[256];
    if (ozl_fse_build_dtable(dt, cells, norm, maxSym, tl, next)) return -2;
    return ozl_fse_decode_seq(dst, n, src, size, dt, tl, nbStates);
}
size_t cu_fse_write_ncount(uint8_t* dst, const int16_t* norm, unsigned maxSym, unsigned tl)
{
    return ozl_fse_write_ncount(dst, norm, maxSym, tl);
}
long cu_fse_read_ncount(int16_t* norm, unsigned* ms, unsigned* tl, const uint8_t* src, size_t size)
{
    return ozl_fse_read_ncount(norm, ms, tl, src, size);
}
int cu_huf_build(uint8_t* nb, uint16_t* val, const uint32_t* count, unsigned maxSym, unsigned maxNbBits)
{
    static ozl_huf_scratch w;
    return ozl_huf_build(nb, val, count, maxSym, maxNbBits, &w);
}
int cu_huf_from_weights(uint8_t* len, uint16_t* val, const uint8_t* w, unsigned nbW)
{
    uint16_t rk[28];
    return ozl_huf_from_weights(len, val, w, nbW, rk);
}
size_t cu_flatpack(uint8_t* alphabet, uint8_t* packed, size_t* packedSize, const uint8_t* src, size_t n)
{
    uint32_t count[256] = { 0 };
    for (size_t i = 0; i < n; i++) count[src[i]]++;
    uint8_t map[256];
    return ozl_flatpack_encode_seq(alphabet, packed, packedSize, src, n, count, map);
}

// Frame recognition of the batched entropy decoder (ozl_dparse.h): returns
// the plan's status and fills `plan` (the ozl_dplan bytes) and `side` (FSE
// normalized counts int16[256] or Huffman code lengths u8[256]).
// words = 1: the frame's first OZL_DPREFIX bytes are seen as aligned words
// (zeros past the frame) and the NCount / nested FSE streams inside them are
// read 32 bits at a time, as on the device; 0: byte by byte.
int cu_dparse(const uint8_t* fr, size_t flen, void* plan, uint8_t* side, int words)
{
    static ozl_dlds L;
    alignas(16) static uint8_t pre[OZL_DPREFIX + 16];
    ozl_dplan d;
    memset(&L, 0x5A, sizeof(L));
    if (words) {
        size_t pn = flen < OZL_DPREFIX ? flen : OZL_DPREFIX;
        memset(pre, 0, sizeof(pre));
        memcpy(pre, fr, pn);
        for (int s = 0; s < 256; s++) L.norm[s] = 0;
        ozl_fview v = { fr, pre, pn, reinterpret_cast<const uint32_t*>(pre) };
        ozl_dparse_frame(v, flen, L, d, true);
    } else {
        ozl_fview v = { fr, fr, flen };
        ozl_dparse_frame(v, flen, L, d);
    }
    memcpy(plan, &d, sizeof(d));
    if (d.status == 0 && d.bk == OZL_BK_FSE) memcpy(side, L.norm, 512);
    if (d.status == 0 && d.bk == OZL_BK_HUF) memcpy(side, L.wlen, 256);
    return (int)d.status;
}
size_t cu_dplan_size(void) { return sizeof(ozl_dplan); }
}

// Frame and chunk headers read through windows of `window` bytes (0: the
// whole frame), as the universal decoder reads device frames: out receives
// [nbChunks, headerEnd, endMarker, then per chunk start, dataStart, csumPos,
// end, nbNodes, nbStored]; returns 0 or the error code.
#include <algorithm>
#include "ozl_hparse.h"
extern "C" int cu_parse_frame(const uint8_t* f, size_t size, size_t window, uint64_t* out, size_t outCap)
{
    using namespace ozl_hparse;
    ParsedFrame P;
    size_t pos = 0, win = window ? window : size;
    bool done = false;
    std::vector<uint8_t> buf;
    while (!done) {
        if (pos >= size) return OZL_ERR_CORRUPTION;
        const size_t len = std::min(win, size - pos), base = pos;
        buf.assign(f + base, f + base + len);
        bool longer = false;
        int rc = parse_window(buf.data(), base, len, size, P, pos, done, longer);
        if (rc == kNeedMore) {
            if (len == size - base && window == 0) return OZL_ERR_CORRUPTION;
            if (longer) {
                if (len >= size - base) return OZL_ERR_CORRUPTION;
                win *= 4;
            }
            continue;
        }
        if (rc) return rc;
    }
    size_t k = 0;
    auto put = [&](uint64_t v) { if (k < outCap) out[k] = v; k++; };
    put(P.chunks.size());
    put(P.headerEnd);
    put(P.end);
    for (const ParsedChunk& C : P.chunks) {
        put(C.start);
        put(C.dataStart);
        put(C.csumPos);
        put(C.end);
        put(C.nodes.size());
        put(C.storedId.size());
    }
    return 0;
}
