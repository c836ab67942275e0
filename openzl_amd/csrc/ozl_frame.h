// ozl_frame.h -- frame and chunk header serialisation (host and device).
//
// Restates the wire format of src/openzl/common/wire_format.h:16-128 as
// written by src/openzl/compress/encode_frameheader.c:441-635 (frame header)
// and :667-832 (chunk header), with the final-graph ordering rules of
// src/openzl/compress/cctx.c:1593-1718 and rtgraphs.c:444-468:
//   * nodes are listed in reverse creation order,
//   * each node's input distance is firstOutputStream - inputStream,
//   * stored streams are listed in reverse stream-creation order, preceded by
//     the transform-header stream (node headers, reverse node order).
// Format version 23 (include/openzl/zl_version.h:57).
#pragma once
#include "ozl_core.h"
#include "ozl_xxh3.h"

#define OZL_FORMAT_VERSION 23u
#define OZL_MAGIC_BASE 0xD7B1A5C0u

// Standard transform IDs (src/openzl/common/wire_format.h:214-307)
enum {
    OZL_TR_DELTA_INT = 1,
    OZL_TR_ZIGZAG = 3,
    OZL_TR_TRANSPOSE_SPLIT = 4,
    OZL_TR_CONV_STRUCT_TO_SERIAL = 6,
    OZL_TR_CONV_STRUCT_TO_NUM_LE = 7,
    OZL_TR_CONV_NUM_TO_STRUCT_LE = 8,
    OZL_TR_CONV_NUM_TO_SERIAL_LE = 10,
    OZL_TR_BITPACK_SERIAL = 27,
    OZL_TR_BITPACK_INT = 28,
    OZL_TR_FLATPACK = 29,
    OZL_TR_TOKENIZE_FIXED = 36,
    OZL_TR_SPLITN = 40,
    OZL_TR_CONSTANT_SERIAL = 44,
    OZL_TR_CONSTANT_FIXED = 45,
    OZL_TR_SPLITN_STRUCT = 47,
    OZL_TR_SPLITN_NUM = 48,
    OZL_TR_FSE_V2 = 49,
    OZL_TR_HUFFMAN_V2 = 50,
    OZL_TR_HUFFMAN_STRUCT_V2 = 51,
    OZL_TR_FSE_NCOUNT = 52,
};

// Stream types as encoded in the frame header (encode_frameheader.c:368-381)
enum { OZL_T_SERIAL = 0, OZL_T_STRUCT = 1, OZL_T_NUMERIC = 2 };

struct ozl_fnode {
    uint8_t trid;
    uint8_t hdrLen;
    uint16_t nbVO;
    uint32_t in;        // input stream id (all nodes here have one input)
    uint32_t firstOut;  // id of the node's first output stream
    uint8_t hdr[12];
};

// Where a stream's bytes live; interpreted by the frame assembler.
struct ozl_fstream {
    uint32_t size;
    uint8_t store;
    uint8_t src;     // OZL_SRC_*
    uint16_t srcOff; // offset inside an auxiliary area
};
enum { OZL_SRC_NONE = 0, OZL_SRC_INPUT = 1, OZL_SRC_PAYLOAD = 2, OZL_SRC_AUX = 3 };

// LSB-first packing of nb-bit fields (ZS_bitpackEncode32 on header fields)
struct ozl_bitsink {
    uint8_t* out;
    size_t pos;
    uint64_t acc;
    int bits;
};
OZL_HD void ozl_bs_put(ozl_bitsink* s, uint32_t v, int nb)
{
    s->acc |= (uint64_t)v << s->bits;
    s->bits += nb;
    while (s->bits >= 8) { s->out[s->pos++] = (uint8_t)s->acc; s->acc >>= 8; s->bits -= 8; }
}
OZL_HD void ozl_bs_flush(ozl_bitsink* s)
{
    if (s->bits) { s->out[s->pos++] = (uint8_t)s->acc; s->acc = 0; s->bits = 0; }
}

// Frame header for single-input frames (encode_frameheader.c:441-635).
// Returns its size (<= 16 bytes).
OZL_HD size_t ozl_write_frame_header(uint8_t* out, int type, uint64_t inputSize, int contentChecksum,
                                     int compressedChecksum)
{
    uint32_t magic = OZL_MAGIC_BASE + OZL_FORMAT_VERSION;
    size_t p = 0;
    out[p++] = (uint8_t)magic; out[p++] = (uint8_t)(magic >> 8);
    out[p++] = (uint8_t)(magic >> 16); out[p++] = (uint8_t)(magic >> 24);
    out[p++] = (uint8_t)((contentChecksum ? 1 : 0) | (compressedChecksum ? 2 : 0));
    out[p++] = (uint8_t)(1 | (type << 4));
    p += ozl_varint(inputSize + 1, out + p);
    if (compressedChecksum) out[p] = (uint8_t)(ozl_xxh3_short(out, p) & 255), p++;
    return p;
}

// Chunk header + transform-header stream for a finished graph
// (encode_frameheader.c:667-832, cctx.c:1593-1718). `order` receives the ids
// of stored streams in frame order. Returns the number of bytes written.
OZL_HD size_t ozl_write_chunk_header(uint8_t* out, const ozl_fnode* nodes, int nbNodes, const ozl_fstream* streams,
                                     int nbStreams, int* order, int* nbStoredOut)
{
    int nbStored = 0;
    for (int s = nbStreams - 1; s >= 0; s--)
        if (streams[s].store) order[nbStored++] = s;
    *nbStoredOut = nbStored;
    size_t p = 0;
    p += ozl_varint((uint64_t)nbNodes + 1, out + p);
    p += ozl_varint((uint64_t)nbStored, out + p);
    ozl_bitsink bs = { out, p, 0, 0 };
    for (int k = 0; k < nbNodes; k++) ozl_bs_put(&bs, 0, 1);  // all standard transforms
    ozl_bs_flush(&bs);
    for (int k = nbNodes - 1; k >= 0; k--) ozl_bs_put(&bs, nodes[k].trid, 6);
    ozl_bs_flush(&bs);
    for (int k = nbNodes - 1; k >= 0; k--) ozl_bs_put(&bs, nodes[k].hdrLen > 0, 1);
    ozl_bs_flush(&bs);
    p = bs.pos;
    for (int k = nbNodes - 1; k >= 0; k--)
        if (nodes[k].hdrLen) p += ozl_varint(nodes[k].hdrLen - 1u, out + p);
    bs.pos = p;
    for (int k = nbNodes - 1; k >= 0; k--) ozl_bs_put(&bs, nodes[k].nbVO > 0, 1);
    ozl_bs_flush(&bs);
    p = bs.pos;
    for (int k = nbNodes - 1; k >= 0; k--)
        if (nodes[k].nbVO) p += ozl_varint(nodes[k].nbVO - 1u, out + p);
    bs.pos = p;
    for (int k = 0; k < nbNodes; k++) ozl_bs_put(&bs, 0, 1);  // every node has exactly one input
    ozl_bs_flush(&bs);
    int nbBits = ozl_next_pow2((uint64_t)nbNodes + (uint64_t)nbStored);
    for (int k = nbNodes - 1; k >= 0; k--) ozl_bs_put(&bs, nodes[k].firstOut - nodes[k].in - 1, nbBits);
    ozl_bs_flush(&bs);
    p = bs.pos;
    for (int k = 0; k < nbStored; k++) p += ozl_varint(streams[order[k]].size, out + p);
    for (int k = nbNodes - 1; k >= 0; k--)
        for (int b = 0; b < nodes[k].hdrLen; b++) out[p++] = nodes[k].hdr[b];
    return p;
}

// -------------------------------------------------------------------------
// Per-chunk plan of the FSE / Huffman dynamic graphs for one serial input of
// fewer than 100000 bytes (no splitting: encode_entropy_binding.c:491-519).
struct ozl_eplan {
    uint32_t n;         // input bytes
    uint32_t payload;   // main payload size (FSE/HUF bits, flatpack packed, bitpacked)
    uint8_t bk;         // OZL_BK_*
    uint8_t maxSym;
    uint8_t tl;         // FSE tableLog / Huffman max code length
    uint8_t nbStates;   // FSE states (2 or 4) / Huffman 4-stream flag
    uint16_t card;
    uint16_t aux;       // FSE: ncount bytes; flatpack: alphabet size; bitpack: nbBits; constant: value
    uint8_t wbk;        // Huffman weights: backend, or 255 when stored raw (minStreamSize)
    uint8_t wMaxSym;
    uint8_t wTl;
    uint8_t wAux8;      // weights flatpack/bitpack nbBits, constant value
    uint16_t wAux;      // weights: FSE ncount bytes / flatpack alphabet size
    uint16_t wPayload;  // weights: FSE bits / flatpack packed / bitpacked size
    uint32_t contentHash;
    uint32_t status;    // 0 ok, else error code
    uint32_t pad[8];
};

// Auxiliary area per chunk (bytes): main ncount/alphabet at 0; Huffman
// weights at 0, their nested ncount/alphabet at 256, nested payload at 512.
#define OZL_AUX_BYTES 1024
#define OZL_AUX_W_SIDE 256
#define OZL_AUX_W_PAYLOAD 512

#define OZL_MIN_STREAM_SIZE 10  // default ZL_CParam_minStreamSize (zl_compress.h:137)

// Nodes and streams that the reference engine creates for one entropy chunk
// (encode_entropy_binding.c:601-753 plus the successors of each node, the
// minStreamSize auto-store of cctx.c:1127-1140, the num->serial conversion of
// cctx.c:563-605, and the standard graphs of graph_registry.c:125-181).
// Stream 0 is the chunk input. Returns the number of nodes.
OZL_HD int ozl_entropy_layout(const ozl_eplan* p, ozl_fnode* nodes, ozl_fstream* st, int* nbStreamsOut)
{
    int nn = 0, ns = 1;
    st[0] = { p->n, 0, OZL_SRC_INPUT, 0 };
    auto node = [&](uint8_t trid, uint32_t in, int nbOut, uint16_t nbVO) -> ozl_fnode* {
        ozl_fnode* q = &nodes[nn++];
        q->trid = trid;
        q->hdrLen = 0;
        q->nbVO = nbVO;
        q->in = in;
        q->firstOut = (uint32_t)ns;
        (void)nbOut;
        return q;
    };
    auto hdr_n = [](ozl_fnode* q, uint8_t first, uint32_t n) {
        // [first, LE n in ceil(bits(n+1)/8) bytes] (encode_entropy_binding.c:73-81, 247-255)
        int nbBytes = (ozl_next_pow2((uint64_t)n + 1) + 7) / 8;
        q->hdr[0] = first;
        for (int b = 0; b < nbBytes; b++) q->hdr[1 + b] = (uint8_t)(n >> (8 * b));
        q->hdrLen = (uint8_t)(1 + nbBytes);
    };
    if (p->n < OZL_MIN_STREAM_SIZE || p->bk == OZL_BK_STORE) {
        st[0].store = 1;
        *nbStreamsOut = ns;
        return nn;
    }
    switch (p->bk) {
    case OZL_BK_CONSTANT: {
        ozl_fnode* q = node(OZL_TR_CONSTANT_SERIAL, 0, 1, 0);
        q->hdrLen = (uint8_t)ozl_varint(p->n, q->hdr);
        st[ns++] = { 1, 1, OZL_SRC_AUX, 0 };
        break;
    }
    case OZL_BK_FLATPACK: {
        node(OZL_TR_FLATPACK, 0, 2, 0);
        st[ns++] = { p->aux, 1, OZL_SRC_AUX, 0 };       // alphabet
        st[ns++] = { p->payload, 1, OZL_SRC_PAYLOAD, 0 }; // packed
        break;
    }
    case OZL_BK_BITPACK: {
        ozl_fnode* q = node(OZL_TR_BITPACK_SERIAL, 0, 1, 0);
        size_t maxElts = (size_t)p->payload * 8 / p->aux;
        q->hdr[0] = (uint8_t)(p->aux - 1);
        q->hdrLen = 1;
        if (maxElts > p->n) { q->hdr[1] = (uint8_t)(maxElts - p->n); q->hdrLen = 2; }
        st[ns++] = { p->payload, 1, OZL_SRC_PAYLOAD, 0 };
        break;
    }
    case OZL_BK_FSE: {
        ozl_fnode* q = node(OZL_TR_FSE_V2, 0, 2, 0);
        hdr_n(q, p->nbStates, p->n);
        int ncountId = ns;
        st[ns++] = { 2u * (p->maxSym + 1u), 0, OZL_SRC_AUX, 0 };  // normalized counts (int16)
        int bitsId = ns;
        st[ns++] = { p->payload, 1, OZL_SRC_PAYLOAD, 0 };
        if (st[ncountId].size < OZL_MIN_STREAM_SIZE) {
            // stored through STORE1 -> serial_store: num->serial conversion first
            ozl_fnode* c = node(OZL_TR_CONV_NUM_TO_SERIAL_LE, (uint32_t)ncountId, 1, 0);
            c->hdr[0] = 1;  // ZL_nextPow2(eltWidth = 2)
            c->hdrLen = 1;
            st[ns++] = { st[ncountId].size, 1, OZL_SRC_AUX, 0 };  // raw int16 counts, in aux at 0
        } else {
            node(OZL_TR_FSE_NCOUNT, (uint32_t)ncountId, 1, 0);
            st[ns++] = { p->aux, 1, OZL_SRC_AUX, 0 };
        }
        (void)bitsId;
        break;
    }
    case OZL_BK_HUF: {
        ozl_fnode* q = node(OZL_TR_HUFFMAN_V2, 0, 2, 0);
        hdr_n(q, p->nbStates, p->n);
        uint32_t m = p->maxSym + 1u;
        int wId = ns;
        st[ns++] = { m, 0, OZL_SRC_AUX, 0 };                  // weights (numeric u8)
        st[ns++] = { p->payload, 1, OZL_SRC_PAYLOAD, 0 };     // bits
        // both the minStreamSize store and the FSE graph take the weights
        // through a num->serial conversion
        ozl_fnode* c = node(OZL_TR_CONV_NUM_TO_SERIAL_LE, (uint32_t)wId, 1, 0);
        c->hdr[0] = 0;  // ZL_nextPow2(eltWidth = 1)
        c->hdrLen = 1;
        int sId = ns;
        st[ns++] = { m, 0, OZL_SRC_AUX, 0 };
        if (m < OZL_MIN_STREAM_SIZE) {
            st[sId].store = 1;
            break;
        }
        switch (p->wbk) {
        case OZL_BK_STORE: st[sId].store = 1; break;
        case OZL_BK_CONSTANT: {
            ozl_fnode* k = node(OZL_TR_CONSTANT_SERIAL, (uint32_t)sId, 1, 0);
            k->hdrLen = (uint8_t)ozl_varint(m, k->hdr);
            st[ns++] = { 1, 1, OZL_SRC_AUX, OZL_AUX_W_PAYLOAD };
            break;
        }
        case OZL_BK_FLATPACK:
            node(OZL_TR_FLATPACK, (uint32_t)sId, 2, 0);
            st[ns++] = { p->wAux, 1, OZL_SRC_AUX, OZL_AUX_W_SIDE };
            st[ns++] = { p->wPayload, 1, OZL_SRC_AUX, OZL_AUX_W_PAYLOAD };
            break;
        case OZL_BK_BITPACK: {
            ozl_fnode* k = node(OZL_TR_BITPACK_SERIAL, (uint32_t)sId, 1, 0);
            size_t maxElts = (size_t)p->wPayload * 8 / p->wAux8;
            k->hdr[0] = (uint8_t)(p->wAux8 - 1);
            k->hdrLen = 1;
            if (maxElts > m) { k->hdr[1] = (uint8_t)(maxElts - m); k->hdrLen = 2; }
            st[ns++] = { p->wPayload, 1, OZL_SRC_AUX, OZL_AUX_W_PAYLOAD };
            break;
        }
        case OZL_BK_FSE: {
            ozl_fnode* k = node(OZL_TR_FSE_V2, (uint32_t)sId, 2, 0);
            hdr_n(k, 2, m);  // fewer than 1000 symbols: 2 states
            int nc = ns;
            st[ns++] = { 2u * (p->wMaxSym + 1u), 0, OZL_SRC_AUX, OZL_AUX_W_SIDE };
            st[ns++] = { p->wPayload, 1, OZL_SRC_AUX, OZL_AUX_W_PAYLOAD };
            if (st[nc].size < OZL_MIN_STREAM_SIZE) {
                ozl_fnode* cv = node(OZL_TR_CONV_NUM_TO_SERIAL_LE, (uint32_t)nc, 1, 0);
                cv->hdr[0] = 1;
                cv->hdrLen = 1;
                st[ns++] = { st[nc].size, 1, OZL_SRC_AUX, OZL_AUX_W_SIDE };
            } else {
                node(OZL_TR_FSE_NCOUNT, (uint32_t)nc, 1, 0);
                st[ns++] = { p->wAux, 1, OZL_SRC_AUX, OZL_AUX_W_SIDE };
            }
            break;
        }
        default: break;
        }
        break;
    }
    default: break;
    }
    *nbStreamsOut = ns;
    return nn;
}
