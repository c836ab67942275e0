// ozl_hparse.h -- host-side reading of frame and chunk headers (format
// version 23; several chunks per frame from version 21 on), shared by the
// universal decoder (ozl_decoder.hip) and its CPU unit build.
// Reference: src/openzl/decompress/decode_frameheader.c, decompress2.c:1442-1560, 1690-1720.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "openzl_amd.h"
#include "ozl_errors.h"
#include "ozl_frame.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"

namespace ozl_hparse {

// Reads frame bytes [base, end) held at p (a window of the frame on the host).
struct HostReader {
    const uint8_t* p;
    size_t base, pos, end;
    bool bad;
    uint32_t u8()
    {
        if (pos < base || pos >= end) { bad = true; return 0; }
        return p[pos++ - base];
    }
    uint64_t varint()
    {
        uint64_t v = 0;
        for (int s = 0; s < 64; s += 7) {
            uint32_t x = u8();
            v |= (uint64_t)(x & 127) << s;
            if (x < 128) return v;
        }
        bad = true;
        return 0;
    }
    void unpack(std::vector<uint32_t>& out, size_t cnt, int nb)
    {
        out.assign(cnt, 0);
        size_t nbytes = (cnt * (size_t)nb + 7) / 8;
        if (pos < base || pos + nbytes > end) { bad = true; return; }
        const uint8_t* q = p + (pos - base);
        size_t bit = 0;
        for (size_t i = 0; i < cnt; i++) {
            uint32_t v = 0;
            for (int b = 0; b < nb; b++, bit++) v |= (uint32_t)((q[bit >> 3] >> (bit & 7)) & 1) << b;
            out[i] = v;
        }
        pos += nbytes;
    }
};

// A node's header bytes: held in place up to 22 bytes (every header of the
// codecs here), in a vector beyond (a frame carries thousands of nodes).
struct NodeHeader {
    uint8_t n = 0;
    uint8_t b[22];
    std::vector<uint8_t> big;
    size_t size() const { return n < sizeof(b) + 1 ? n : big.size(); }
    bool empty() const { return size() == 0; }
    const uint8_t* data() const { return n <= sizeof(b) ? b : big.data(); }
    uint8_t operator[](size_t k) const { return data()[k]; }
    void assign(const uint8_t* first, const uint8_t* last)
    {
        const size_t len = (size_t)(last - first);
        if (len <= sizeof(b)) {
            n = (uint8_t)len;
            for (size_t k = 0; k < len; k++) b[k] = first[k];
            big.clear();
        } else {
            n = (uint8_t)(sizeof(b) + 1);
            big.assign(first, last);
        }
    }
};

struct DNode {
    uint8_t trid;
    NodeHeader hdr;
    uint32_t nbVO;
    uint32_t firstOut, nbOut, in;
};

inline int nb_singletons(uint8_t trid)
{
    switch (trid) {
    case OZL_TR_DELTA_INT: case OZL_TR_ZIGZAG: case OZL_TR_CONV_STRUCT_TO_SERIAL: case OZL_TR_CONV_NUM_TO_STRUCT_LE:
    case OZL_TR_CONV_STRUCT_TO_NUM_LE: case OZL_TR_CONV_NUM_TO_SERIAL_LE: case OZL_TR_BITPACK_SERIAL:
    case OZL_TR_BITPACK_INT: case OZL_TR_CONSTANT_SERIAL: case OZL_TR_CONSTANT_FIXED: case OZL_TR_FSE_NCOUNT: return 1;
    case OZL_TR_FLATPACK: case OZL_TR_FSE_V2: case OZL_TR_HUFFMAN_V2: case OZL_TR_TOKENIZE_FIXED:
    case OZL_TR_HUFFMAN_STRUCT_V2: return 2;
    case OZL_TR_TRANSPOSE_SPLIT: case OZL_TR_SPLITN: case OZL_TR_SPLITN_STRUCT: case OZL_TR_SPLITN_NUM: return 0;
    default: return -1;
    }
}

// One chunk of a frame (format version 21+: a frame holds chunks, each with
// its own graph, stored streams and checksums, then a zero byte;
// decompress2.c:1442-1560, 1690-1720).
struct ParsedChunk {
    size_t start = 0, dataStart = 0, csumPos = 0, end = 0;   // end: after its checksums
    std::vector<DNode> nodes;
    std::vector<uint64_t> storedSize;
    std::vector<uint32_t> storedId;
    uint32_t nbStreams = 0;
};

struct ParsedFrame {
    int flags = 0, type = 0;
    uint64_t size = 0;
    size_t headerEnd = 0;
    std::vector<ParsedChunk> chunks;
    size_t end = 0;   // position of the end-of-frame marker
    bool endDeferred = false;   // the marker (the frame's last byte) is still to be checked as zero
};

// a read ran past the host window before the end of the frame
constexpr int kNeedMore = -2;

// Frame header (decode_frameheader.c); the window starts at the frame start.
inline int parse_frame_header(HostReader& r, ParsedFrame& P)
{
    uint32_t magic = 0;
    for (int k = 0; k < 4; k++) magic |= r.u8() << (8 * k);
    if (r.bad || magic < OZL_MAGIC_BASE || magic > OZL_MAGIC_BASE + 64) return OZL_ERR_HEADER_UNKNOWN;
    if (magic != OZL_MAGIC_BASE + OZL_FORMAT_VERSION) return OZL_ERR_FORMAT_VERSION_UNSUPPORTED;
    P.flags = (int)r.u8();
    if (P.flags & ~3) return ZL_ErrorCode_frameParameter_unsupported;
    uint32_t tok = r.u8();
    if ((tok & 15) != 1) return ZL_ErrorCode_frameParameter_unsupported;  // one input
    P.type = (int)((tok >> 4) & 3);
    uint64_t s1 = r.varint();
    if (r.bad || s1 == 0) return OZL_ERR_CORRUPTION;
    P.size = s1 - 1;
    if (P.flags & 2) {
        uint8_t hc = (uint8_t)(ozl_xxh3_short(r.p, r.pos) & 255);
        if (r.u8() != hc) return OZL_ERR_CORRUPTION;
    }
    if (r.bad) return OZL_ERR_CORRUPTION;
    P.headerEnd = r.pos;
    return 0;
}

// The headers of the chunk at r.pos of a frame of frameSize bytes.
inline int parse_chunk(HostReader& r, size_t frameSize, int flags, ParsedChunk& C)
{
    const bool windowed = r.end < frameSize;
    auto fail = [&]() { return windowed ? kNeedMore : OZL_ERR_CORRUPTION; };
    C.start = r.pos;
    uint64_t nbDec1 = r.varint();
    if (r.bad) return fail();
    if (nbDec1 == 0) return OZL_ERR_CORRUPTION;
    size_t nn = (size_t)nbDec1 - 1;
    uint64_t nbStored = r.varint();
    if (r.bad) return fail();
    if (nn > 20000 || nbStored > 110000) return OZL_ERR_CORRUPTION;
    std::vector<uint32_t> trt, ids, hf, vf, inf, dist;
    r.unpack(trt, nn, 1);
    for (uint32_t t : trt)
        if (t) return ZL_ErrorCode_frameParameter_unsupported;  // custom transforms
    r.unpack(ids, nn, 6);
    C.nodes.assign(nn, DNode());
    for (size_t k = 0; k < nn; k++) C.nodes[k].trid = (uint8_t)ids[nn - 1 - k];
    r.unpack(hf, nn, 1);
    std::vector<size_t> hlen(nn, 0);
    for (size_t i = 0; i < nn; i++)
        if (hf[i]) hlen[nn - 1 - i] = (size_t)r.varint() + 1;
    r.unpack(vf, nn, 1);
    for (size_t k = 0; k < nn; k++) C.nodes[k].nbVO = 0;
    for (size_t i = 0; i < nn; i++)
        if (vf[i]) C.nodes[nn - 1 - i].nbVO = (uint32_t)r.varint() + 1;
    r.unpack(inf, nn, 1);
    for (uint32_t x : inf)
        if (x) return ZL_ErrorCode_frameParameter_unsupported;  // multi-input transforms
    r.unpack(dist, nn, ozl_next_pow2((uint64_t)nn + nbStored));
    C.storedSize.resize(nbStored);
    for (size_t k = 0; k < nbStored; k++) C.storedSize[k] = r.varint();
    if (r.bad) return fail();
    // node headers (decode order)
    size_t hp = r.pos;
    for (size_t i = 0; i < nn; i++) {
        DNode& nd = C.nodes[nn - 1 - i];
        const size_t l = hlen[nn - 1 - i];
        if (hp + l > r.end) return fail();
        nd.hdr.assign(r.p + (hp - r.base), r.p + (hp - r.base) + l);
        hp += l;
    }
    C.dataStart = hp;
    // stream numbering (creation order)
    uint32_t next = 1;
    for (size_t k = 0; k < nn; k++) {
        DNode& nd = C.nodes[k];
        int s1n = nb_singletons(nd.trid);
        if (s1n < 0) return ZL_ErrorCode_frameParameter_unsupported;
        nd.firstOut = next;
        nd.nbOut = (uint32_t)s1n + nd.nbVO;
        next += nd.nbOut;
        uint32_t dd = dist[nn - 1 - k] + 1;
        if (dd > nd.firstOut) return OZL_ERR_CORRUPTION;
        nd.in = nd.firstOut - dd;
    }
    C.nbStreams = next;
    std::vector<uint8_t> consumed(next, 0);
    for (auto& nd : C.nodes) {
        if (consumed[nd.in]) return OZL_ERR_CORRUPTION;
        consumed[nd.in] = 1;
    }
    for (int64_t t = (int64_t)next - 1; t >= 0; t--)
        if (!consumed[t]) C.storedId.push_back((uint32_t)t);
    if (C.storedId.size() != nbStored) return OZL_ERR_CORRUPTION;
    size_t pos = C.dataStart;
    for (size_t k = 0; k < nbStored; k++) {
        if (C.storedSize[k] > frameSize) return OZL_ERR_CORRUPTION;
        pos += C.storedSize[k];
    }
    C.csumPos = pos;
    C.end = pos + ((flags & 1) ? 4 : 0) + ((flags & 2) ? 4 : 0);
    if (C.end + 1 > frameSize) return OZL_ERR_CORRUPTION;   // the end marker follows at the latest
    return 0;
}

// The chunks of one frame whose bytes [base, base + len) are on the host
// (the whole frame, or a window of a device frame): parses the header when
// pos is 0, then chunks until the end marker, or until the window runs out
// (kNeedMore: the next window starts at pos; `longer` when a chunk's headers
// did not fit a window that starts at the chunk). With deferEnd, a chunk that
// ends on the frame's last byte, outside the window, ends the frame: that
// byte can only be the end marker (a chunk there would not fit), and the
// caller checks that it is zero (P.endDeferred).
inline int parse_window(const uint8_t* w, size_t base, size_t len, size_t frameSize, ParsedFrame& P, size_t& pos,
                 bool& done, bool& longer, bool deferEnd = false)
{
    longer = false;
    HostReader r = { w, base, 0, base + len, false };
    if (pos == 0) {
        if (base != 0) return OZL_ERR_LOGIC;
        r.pos = 0;
        int rc = parse_frame_header(r, P);
        if (rc) return (r.end < frameSize && r.bad) ? (longer = true, kNeedMore) : rc;
        pos = P.headerEnd;
    }
    for (;;) {
        if (pos >= base + len) return kNeedMore;
        if (w[pos - base] == 0) {   // end of frame
            P.end = pos;
            done = true;
            return 0;
        }
        ParsedChunk C;
        r.pos = pos;
        r.bad = false;
        int rc = parse_chunk(r, frameSize, P.flags, C);
        if (rc == kNeedMore) {
            longer = pos == base;
            return kNeedMore;
        }
        if (rc) return rc;
        pos = C.end;
        P.chunks.push_back(std::move(C));
        if (deferEnd && pos >= base + len && pos + 1 == frameSize) {
            P.end = pos;
            P.endDeferred = true;
            done = true;
            return 0;
        }
    }
}

}  // namespace ozl_hparse
