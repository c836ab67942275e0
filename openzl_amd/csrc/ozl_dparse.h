// ozl_dparse.h -- recognition of one batched ZL_GRAPH_FSE / ZL_GRAPH_HUFFMAN
// frame and the set-up of its decoding tables (host and device).
//
// The frame and chunk headers are parsed (encode_frameheader.c:441-832 in the
// reverse direction), the node layout is recognised from its transform ids,
// the expected chunk header is re-serialised from what was recognised with
// ozl_entropy_layout / ozl_write_chunk_header and compared byte for byte
// (any difference -> the frame is reported unsupported and is left to the
// generic decoder), and the NCount / Huffman weights (with their nested
// graphs) are decoded into normalized counts or Huffman code lengths.
// The device runs it on one lane per frame; the CPU unit build runs the same
// code against reference frames.
#pragma once
#include "ozl_core.h"
#include "ozl_errors.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"

#if defined(__HIPCC__)
#define OZL_HDN static __host__ __device__
#define OZL_HDM __host__ __device__ inline
#else
#define OZL_HDN static
#define OZL_HDM inline
#endif



struct ozl_dplan {
    uint32_t n;
    uint8_t bk, tl, nbStates, pad0;
    uint32_t payOff, payLen;   // main payload inside the frame
    uint32_t auxOff, auxLen;   // flatpack alphabet / constant byte inside the frame
    uint32_t aux;              // bitpack / flatpack nbBits
    uint32_t chunkOff;         // start of the chunk inside the frame
    uint32_t csumOff;          // position of the content checksum
    uint32_t status;
    uint32_t pad0b;
    uint64_t srcPtr;           // main payload (absolute device address)
    uint64_t outPtr;           // decoded output (absolute device address)
};

// Bytes of a frame kept in the parser's copy (pn <= OZL_DPREFIX).
#define OZL_DPREFIX 1024

// A frame seen through a copy of its first `pn` bytes (in LDS on the device):
// reads inside the copy use it, the rest go to the frame itself. When `w` is
// set, the copy is 4-byte aligned with OZL_DPREFIX + 8 readable bytes, and
// the bit streams that lie inside it are read 32 bits at a time.
struct ozl_fview {
    const uint8_t* g;  // the frame
    const uint8_t* p;  // copy of its first pn bytes
    size_t pn;
    const uint32_t* w = nullptr;   // p as words
    OZL_HDM uint8_t operator[](size_t i) const { return i < pn ? p[i] : g[i]; }
    OZL_HDM const uint8_t* at(size_t off, size_t len) const { return off + len <= pn ? p + off : g + off; }
    OZL_HDM bool words(size_t off, size_t len) const { return w && off + len <= pn; }
    // the 32 bits of the copy from bit b on (LSB-first); the word index wraps
    // inside the copy, so reads past a stream's end stay in bounds
    OZL_HDM uint32_t bits32(size_t b) const
    {
        const uint32_t i = (uint32_t)(b >> 5) & (OZL_DPREFIX / 4 - 1);
        const uint64_t x = (uint64_t)w[i] | ((uint64_t)w[i + 1] << 32);
        return (uint32_t)(x >> (b & 31));
    }
};

// ozl_fse_read_ncount (ozl_core.h) for an NCount lying inside the copy:
// the same fields and checks, each field read from a 32-bit window (the
// reference reads past the end as zeros, here as the frame's next bytes;
// either way a read past the end fails the final check). norm must hold
// zeros on entry.
OZL_HDN long ozl_read_ncount_w(int16_t* norm, unsigned* maxSymOut, unsigned* tlOut, const ozl_fview& fr, size_t off,
                               size_t size)
{
    const size_t b0 = off * 8, total = size * 8;
    size_t pos = 0;
    const unsigned tl = (fr.bits32(b0) & 15u) + OZL_FSE_MIN_TABLELOG;
    pos = 4;
    if (tl > OZL_FSE_MAX_TABLELOG) return -1;
    int remaining = (1 << tl) + 1, threshold = 1 << tl, nb = (int)tl + 1;
    unsigned symbol = 0;
    bool prev0 = false;
    while (remaining > 1 && symbol <= 255) {
        if (prev0) {
            unsigned n0 = symbol;
            for (;;) {
                const uint32_t r = fr.bits32(b0 + pos) & 3u;
                pos += 2;
                n0 += r;
                if (r != 3) break;
                if (n0 > 255) return -1;
            }
            if (n0 > 255) return -1;
            symbol = n0;
        }
        // nb - 1 low bits, and the next bit when they reach max
        const int max = (2 * threshold - 1) - remaining;
        const uint32_t win = fr.bits32(b0 + pos);
        int count = (int)(win & ((1u << (nb - 1)) - 1u));
        if (count < max) pos += (size_t)(nb - 1);
        else {
            count = (int)(win & ((1u << nb) - 1u));
            if (count >= threshold) count -= max;
            pos += (size_t)nb;
        }
        count--;
        remaining -= count < 0 ? -count : count;
        norm[symbol++] = (int16_t)count;
        prev0 = count == 0;
        while (remaining < threshold) { nb--; threshold >>= 1; }
    }
    if (remaining != 1 || pos > total || symbol == 0) return -1;
    *maxSymOut = symbol - 1;
    *tlOut = tl;
    return (long)((pos + 7) / 8);
}

// ozl_fse_decode_seq (ozl_core.h) for a stream lying inside the copy, with
// NS interleaved states: bits [pos - nb, pos) read as one field.
template <int NS>
OZL_HDN int ozl_fse_decode_w(uint8_t* dst, size_t n, const ozl_fview& fr, size_t off, size_t size, const uint32_t* dt,
                             unsigned tl)
{
    if (size == 0) return -1;
    const uint32_t last = fr[off + size - 1];
    if (last == 0) return -1;
    const size_t b0 = off * 8;
    long pos = (long)(size - 1) * 8 + ozl_hb32(last);
    auto rdb = [&](unsigned nb) -> uint32_t {
        pos -= (long)nb;
        return pos < 0 ? 0u : fr.bits32(b0 + (size_t)pos) & ((1u << nb) - 1u);
    };
    // state p mod NS of symbol p sits in x0 when p is decoded
    uint32_t x0 = rdb(tl), x1 = NS > 1 ? rdb(tl) : 0u, x2 = NS > 2 ? rdb(tl) : 0u, x3 = NS > 3 ? rdb(tl) : 0u;
    if (pos < 0) return -1;
    for (size_t p = 0; p < n; p++) {
        const uint32_t e = dt[x0];
        dst[p] = (uint8_t)e;
        if (p + NS < n) {
            const uint32_t r = rdb((e >> 8) & 255);
            if (pos < 0) return -1;
            x0 = (e >> 16) + r;
        }
        const uint32_t t = x0;
        if (NS == 2) { x0 = x1; x1 = t; }
        else if (NS == 3) { x0 = x1; x1 = x2; x2 = t; }
        else if (NS == 4) { x0 = x1; x1 = x2; x2 = x3; x3 = t; }
    }
    return pos == 0 ? 0 : -1;
}

struct ozl_reader {
    ozl_fview p;
    size_t pos, end;
    bool bad;
    OZL_HDM uint32_t u8() { if (pos >= end) { bad = true; return 0; } return p[pos++]; }
    OZL_HDM uint64_t varint()
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
    // LSB-first unpack of cnt fields of nb bits
    OZL_HDM void unpack(uint32_t* out, int cnt, int nb)
    {
        size_t nbytes = ((size_t)cnt * nb + 7) / 8;
        if (pos + nbytes > end) { bad = true; return; }
        size_t bit = 0;
        for (int i = 0; i < cnt; i++) {
            uint32_t v = 0;
            for (int b = 0; b < nb; b++, bit++) v |= (uint32_t)((p[pos + (bit >> 3)] >> (bit & 7)) & 1) << b;
            out[i] = v;
        }
        pos += nbytes;
    }
};

// The nested FSE graph of the Huffman weights codes at most 256 symbols of
// value <= 12: its tableLog never exceeds 6 (fse_compress.c:379-396), so its
// decoding table is bounded here at 2^8 cells.
#define OZL_WEIGHTS_MAX_TABLELOG 8

// Per-frame parsing state kept with the tables (in LDS on the device, where
// dynamically indexed locals would live in scratch memory).
struct ozl_dscratch {
    uint8_t fh[24];
    ozl_fnode nodes[8];
    ozl_fnode en[8];
    ozl_fstream st[12];
    uint32_t soff[12];
    int order[12];
    uint32_t f[8], ids[8], hf[8], vf[8], inf[8], dist[8];
    uint32_t storedSize[4];
    uint16_t next[256];  // table build workspace
    uint16_t rk[28];
};

struct ozl_dlds {
    ozl_dscratch x;
    int16_t norm[256];
    uint8_t cells[1 << OZL_WEIGHTS_MAX_TABLELOG];
    uint8_t w[256];
    uint8_t wlen[256];
    uint16_t wval[256];
    uint32_t wdt[1 << OZL_WEIGHTS_MAX_TABLELOG];
    uint8_t expect[512];
};

// An NCount read while the header was parsed (kept so that the tables
// pass does not read it again).
struct ozl_ncread {
    size_t off, size;
    long used;
    unsigned ms, tl;
};

// NCount of `size` bytes at `off` into norm: through the copy's words when it
// lies there (norm then holds zeros on entry), else byte by byte.
OZL_HDN long ozl_dread_ncount(int16_t* norm, unsigned* ms, unsigned* tl, const ozl_fview& fr, size_t off, size_t size,
                              bool normZero)
{
    if (normZero && fr.words(off, size)) return ozl_read_ncount_w(norm, ms, tl, fr, off, size);
    return ozl_fse_read_ncount(norm, ms, tl, fr.at(off, size), size);
}

// Decodes the Huffman weights stream described by the recognised layout.
OZL_HDN int ozl_decode_weights(const ozl_fview& fr, const ozl_fnode* nodes, int nn, const ozl_fstream* st,
                                const uint32_t* soff, uint32_t m, ozl_dlds& L, const ozl_ncread* nc)
{
    // stream ids follow ozl_entropy_layout: 1 weights, 3 serial weights, 4.. nested outputs
    if (nn == 2) {  // conversion then STORE
        for (uint32_t i = 0; i < m; i++) L.w[i] = fr[soff[3] + i];
        return 0;
    }
    uint8_t t = nodes[2].trid;
    if (t == OZL_TR_CONSTANT_SERIAL) {
        for (uint32_t i = 0; i < m; i++) L.w[i] = fr[soff[4]];
        return 0;
    }
    if (t == OZL_TR_FLATPACK) {
        uint32_t card = st[4].size, plen = st[5].size;
        const uint8_t* alpha = fr.at(soff[4], st[4].size);
        const uint8_t* pk = fr.at(soff[5], st[5].size);
        if (card == 0 || plen == 0 || pk[plen - 1] == 0) return -1;
        uint32_t nb = card <= 1 ? card : 1 + (uint32_t)ozl_hb32(card - 1);
        uint64_t totalBits = (uint64_t)(plen - 1) * 8 + ozl_hb32(pk[plen - 1]);
        if (nb == 0 || totalBits / nb != m) return -1;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t v = 0;
            for (uint32_t b = 0; b < nb; b++) {
                uint64_t bit = (uint64_t)i * nb + b;
                v |= (uint32_t)((pk[bit >> 3] >> (bit & 7)) & 1) << b;
            }
            if (v >= card) return -1;
            L.w[i] = alpha[v];
        }
        return 0;
    }
    if (t == OZL_TR_BITPACK_SERIAL) {
        uint32_t nb = (nodes[2].hdr[0] & 63) + 1;
        const uint8_t* pk = fr.at(soff[4], st[4].size);
        if (nb > 8) return -1;
        for (uint32_t i = 0; i < m; i++) {
            uint32_t v = 0;
            for (uint32_t b = 0; b < nb; b++) {
                uint64_t bit = (uint64_t)i * nb + b;
                v |= (uint32_t)((pk[bit >> 3] >> (bit & 7)) & 1) << b;
            }
            L.w[i] = (uint8_t)v;
        }
        return 0;
    }
    if (t == OZL_TR_FSE_V2) {
        unsigned ms, tl;
        int16_t* wn = L.norm;  // free in Huffman frames
        if (nn == 4 && nodes[3].trid == OZL_TR_FSE_NCOUNT) {  // NCount encoded
            if (nc && nc->off == soff[6] && nc->size == st[6].size) {   // read by the header pass
                if (nc->used < 0) return -1;
                ms = nc->ms;
                tl = nc->tl;
            } else if (ozl_fse_read_ncount(wn, &ms, &tl, fr.at(soff[6], st[6].size), st[6].size) < 0) return -1;
        } else {        // raw int16 counts, stored after a num->serial conversion
            uint32_t cnt = st[6].size / 2;
            if (cnt == 0) return -1;
            uint32_t sum = 0;
            for (uint32_t s = 0; s < 256; s++) wn[s] = 0;
            for (uint32_t s = 0; s < cnt; s++) {
                wn[s] = (int16_t)(fr[soff[6] + 2 * s] | (fr[soff[6] + 2 * s + 1] << 8));
                sum += wn[s] == -1 ? 1 : (uint32_t)wn[s];
            }
            if (sum == 0 || (sum & (sum - 1))) return -1;
            ms = cnt - 1;
            tl = (unsigned)ozl_hb32(sum);
        }
        if (tl < OZL_FSE_MIN_TABLELOG || tl > OZL_FSE_MAX_TABLELOG) return -1;
        if (tl > OZL_WEIGHTS_MAX_TABLELOG) return -1;
        if (ozl_fse_build_dtable(L.wdt, L.cells, wn, ms, tl, L.x.next)) return -1;
        const unsigned ns = nodes[2].hdr[0];
        if (fr.words(soff[5], st[5].size) && ns >= 1 && ns <= 4) {
            const size_t o = soff[5], z = st[5].size;
            return ns == 1 ? ozl_fse_decode_w<1>(L.w, m, fr, o, z, L.wdt, tl)
                 : ns == 2 ? ozl_fse_decode_w<2>(L.w, m, fr, o, z, L.wdt, tl)
                 : ns == 3 ? ozl_fse_decode_w<3>(L.w, m, fr, o, z, L.wdt, tl)
                           : ozl_fse_decode_w<4>(L.w, m, fr, o, z, L.wdt, tl);
        }
        return ozl_fse_decode_seq(L.w, m, fr.at(soff[5], st[5].size), st[5].size, L.wdt, tl, ns);
    }
    return -1;
}

// Parses frame `fr` (flen bytes). On success d.status is 0; for FSE, L.norm
// holds the normalized counts (d.aux = maxSymbol, d.tl = tableLog); for
// Huffman, L.wlen holds the code lengths (d.tl = maximum length). The tables
// themselves are built by the decoding kernels.
// `normZero`: L.norm holds zeros on entry. `weightsOut`: when set, a Huffman
// frame stops once L.w holds its weights, their number goes to *weightsOut and
// the caller derives the code lengths (ozl_huf_from_weights) and d.tl.
OZL_HDN void ozl_dparse_frame(const ozl_fview& fr, size_t flen, ozl_dlds& L, ozl_dplan& d, bool normZero = false,
                              uint32_t* weightsOut = nullptr)
{
    d = ozl_dplan{};
    d.status = OZL_ERR_CORRUPTION;
    ozl_reader r = { fr, 0, flen, false };
    uint32_t magic = 0;
    for (int k = 0; k < 4; k++) magic |= r.u8() << (8 * k);
    uint32_t flags = r.u8();
    uint32_t tok = r.u8();
    uint64_t n1 = r.varint();
    bool ok = !r.bad && magic == OZL_MAGIC_BASE + OZL_FORMAT_VERSION && flags == 3 && tok == 1 && n1 >= 1;
    uint8_t* fh = L.x.fh;
    size_t fhl = 0;
    if (ok) {
        fhl = ozl_write_frame_header(fh, OZL_T_SERIAL, n1 - 1, 1, 1);
        for (size_t k = 0; k + 1 < fhl; k++) ok &= fr[k] == fh[k];
        ok &= r.u8() == fh[fhl - 1];  // header checksum
    }
    d.n = (uint32_t)(n1 - 1);
    d.chunkOff = (uint32_t)r.pos;
    ozl_fnode* nodes = L.x.nodes;
    ozl_fstream* st = L.x.st;
    uint32_t* soff = L.x.soff;
    int nn = 0, nbStored = 0;
    if (ok) {
        uint64_t nbDec = r.varint();
        uint64_t nbSt = r.varint();
        ok = !r.bad && nbDec >= 1 && nbDec - 1 <= 4 && nbSt >= 1 && nbSt <= 4;
        if (ok) {
            nn = (int)nbDec - 1;
            nbStored = (int)nbSt;
            uint32_t* f = L.x.f;
            uint32_t* ids = L.x.ids;
            r.unpack(f, nn, 1);
            r.unpack(ids, nn, 6);
            // node k (creation order) is decoder nn-1-k
            for (int k = 0; k < nn; k++) {
                nodes[k].trid = (uint8_t)ids[nn - 1 - k];
                nodes[k].hdrLen = 0;
                nodes[k].nbVO = 0;
            }
            uint32_t* hf = L.x.hf;
            r.unpack(hf, nn, 1);
            for (int k = nn - 1; k >= 0; k--)
                if (hf[nn - 1 - k]) nodes[k].hdrLen = (uint8_t)(r.varint() + 1);
            ok = !r.bad;
        }
    }
    // Recognise the layout from the creation-order transform ids and rebuild
    // the plan; the layout function then yields the expected stream list.
    ozl_eplan p = {};
    p.n = d.n;
    int ns = 0;
    if (ok) {
        uint8_t t0 = nn > 0 ? nodes[0].trid : 0;
        if (nn == 0) p.bk = OZL_BK_STORE;
        else if (t0 == OZL_TR_CONSTANT_SERIAL && nn == 1) p.bk = OZL_BK_CONSTANT;
        else if (t0 == OZL_TR_FLATPACK && nn == 1) p.bk = OZL_BK_FLATPACK;
        else if (t0 == OZL_TR_BITPACK_SERIAL && nn == 1) p.bk = OZL_BK_BITPACK;
        else if (t0 == OZL_TR_FSE_V2 && nn == 2) p.bk = OZL_BK_FSE;
        else if (t0 == OZL_TR_HUFFMAN_V2) p.bk = OZL_BK_HUF;
        else ok = false;
    }
    // The stored sizes come after the VO/input flags and distances; parse
    // the rest of the header generically first.
    uint32_t* storedSize = L.x.storedSize;
    for (int k = 0; k < 4; k++) storedSize[k] = 0;
    if (ok) {
        uint32_t* vf = L.x.vf;
        r.unpack(vf, nn, 1);
        for (int k = 0; k < nn; k++) ok &= vf[k] == 0;
        uint32_t* inf = L.x.inf;
        r.unpack(inf, nn, 1);
        for (int k = 0; k < nn; k++) ok &= inf[k] == 0;
        uint32_t* dist = L.x.dist;
        r.unpack(dist, nn, ozl_next_pow2((uint64_t)nn + (uint64_t)nbStored));
        for (int k = 0; k < nbStored; k++) storedSize[k] = (uint32_t)r.varint();
        ok &= !r.bad;
    }
    size_t hsPos = r.pos;  // header stream
    uint32_t hsLen = 0;
    if (ok) {
        for (int k = 0; k < nn; k++) hsLen += nodes[k].hdrLen;
        // node headers, decode order (reverse creation)
        size_t q = hsPos;
        for (int k = nn - 1; k >= 0; k--) {
            if (nodes[k].hdrLen > 12 || q + nodes[k].hdrLen > flen) { ok = false; break; }
            for (int i = 0; i < nodes[k].hdrLen; i++) nodes[k].hdr[i] = fr[q + i];
            q += nodes[k].hdrLen;
        }
    }
    // fill the plan fields the layout depends on
    if (ok) {
        switch (p.bk) {
        case OZL_BK_FLATPACK: p.aux = (uint16_t)storedSize[1]; p.payload = storedSize[0]; break;
        case OZL_BK_BITPACK: p.aux = (uint16_t)((nodes[0].hdr[0] & 63) + 1); p.payload = storedSize[0]; break;
        case OZL_BK_FSE:
            p.nbStates = nodes[0].hdr[0];
            // stored list: [ncount or raw counts, bits]
            p.payload = storedSize[1];
            if (nodes[1].trid == OZL_TR_FSE_NCOUNT) { p.aux = (uint16_t)storedSize[0]; p.maxSym = 4; }
            else p.maxSym = (uint8_t)(storedSize[0] / 2 - 1);
            break;
        case OZL_BK_HUF: {
            p.nbStates = nodes[0].hdr[0];
            // maxSym from the weights stream size (stored raw, or the conversion's stream)
            // stored list: [bits, weights] when the weights are raw, else
            // [nested streams..., bits]
            p.payload = storedSize[nbStored - 1];
            p.maxSym = 255;  // fixed below from the nested header
            if (nn < 2 || nodes[1].trid != OZL_TR_CONV_NUM_TO_SERIAL_LE) ok = false;
            if (ok) {
                uint8_t t2 = nn > 2 ? nodes[2].trid : 0;
                if (nn == 2) { p.wbk = OZL_BK_STORE; p.maxSym = (uint8_t)(storedSize[0] - 1); }
                else if (t2 == OZL_TR_CONSTANT_SERIAL) {
                    p.wbk = OZL_BK_CONSTANT;
                    uint64_t m = 0;
                    for (int i = 0, s = 0; i < nodes[2].hdrLen; i++, s += 7) m |= (uint64_t)(nodes[2].hdr[i] & 127) << s;
                    p.maxSym = (uint8_t)(m - 1);
                } else if (t2 == OZL_TR_FLATPACK) {
                    p.wbk = OZL_BK_FLATPACK; p.wPayload = (uint16_t)storedSize[0]; p.wAux = (uint16_t)storedSize[1];
                } else if (t2 == OZL_TR_BITPACK_SERIAL) {
                    p.wbk = OZL_BK_BITPACK; p.wAux8 = (uint8_t)((nodes[2].hdr[0] & 63) + 1); p.wPayload = (uint16_t)storedSize[0];
                } else if (t2 == OZL_TR_FSE_V2) {
                    p.wbk = OZL_BK_FSE;
                    // stored list: [ncount or raw counts, nested bits, bits]
                    p.wPayload = (uint16_t)storedSize[1];
                    if (nn == 4 && nodes[3].trid == OZL_TR_FSE_NCOUNT) { p.wAux = (uint16_t)storedSize[0]; p.wMaxSym = 4; }
                    else p.wMaxSym = (uint8_t)(storedSize[0] / 2 - 1);
                } else ok = false;
                if (t2 == OZL_TR_FLATPACK || t2 == OZL_TR_BITPACK_SERIAL || t2 == OZL_TR_FSE_V2) {
                    uint64_t m = 0;  // from the nested node header or sizes
                    if (t2 == OZL_TR_FSE_V2) {
                        for (int i = 1; i < nodes[2].hdrLen; i++) m |= (uint64_t)nodes[2].hdr[i] << (8 * (i - 1));
                    } else if (t2 == OZL_TR_BITPACK_SERIAL) {
                        uint32_t nb = p.wAux8;
                        uint32_t extra = nodes[2].hdrLen == 2 ? nodes[2].hdr[1] : 0;
                        m = (uint64_t)p.wPayload * 8 / nb - extra;
                    } else {
                        uint32_t card = p.wAux, plen = p.wPayload;
                        uint32_t nb = card <= 1 ? card : 1 + (uint32_t)ozl_hb32(card - 1);
                        const uint8_t* pk = nullptr;
                        (void)pk;
                        // terminating bit position gives the count; needs the stream
                        m = 0;
                        if (nb) {
                            // payload position is known after layout; computed below
                            m = (uint64_t)-1;
                        }
                        (void)plen;
                    }
                    if (m != (uint64_t)-1) p.maxSym = (uint8_t)(m - 1);
                }
            }
            break;
        }
        default: break;
        }
    }
    // stored stream positions: header stream first, then stored streams in list order
    size_t dataPos = hsPos + hsLen;
    ozl_ncread nc = {};
    bool haveNc = false;
    if (ok) {
        // flatpack weights: recover m from the terminating bit now that positions are known
        if (p.bk == OZL_BK_HUF && p.wbk == OZL_BK_FLATPACK) {
            // stored list for this shape: [packed, alphabet, bits]
            uint32_t plen = storedSize[0];
            const uint8_t* pk = fr.at(dataPos, plen);
            uint32_t card = p.wAux;
            uint32_t nb = card <= 1 ? card : 1 + (uint32_t)ozl_hb32(card - 1);
            if (plen == 0 || dataPos + plen > flen || pk[plen - 1] == 0 || nb == 0) ok = false;
            else p.maxSym = (uint8_t)(((uint64_t)(plen - 1) * 8 + ozl_hb32(pk[plen - 1])) / nb - 1);
        }
        bool mainNc = p.bk == OZL_BK_FSE && nodes[1].trid == OZL_TR_FSE_NCOUNT;
        bool nestedNc = p.bk == OZL_BK_HUF && p.wbk == OZL_BK_FSE && nn == 4 && nodes[3].trid == OZL_TR_FSE_NCOUNT;
        if ((mainNc || nestedNc) && dataPos + storedSize[0] > flen) ok = false;
        if (ok && (mainNc || nestedNc)) {
            // maxSym from the NCount itself; the counts stay for the tables
            nc.off = dataPos;
            nc.size = storedSize[0];
            nc.used = ozl_dread_ncount(L.norm, &nc.ms, &nc.tl, fr, dataPos, storedSize[0], normZero);
            haveNc = true;
            if (nc.used < 0) ok = false;
            else if (mainNc) p.maxSym = (uint8_t)nc.ms;
            else p.wMaxSym = (uint8_t)nc.ms;
        }
    }
    int* order = L.x.order;
    int nbSt2 = 0;
    ozl_fnode* en = L.x.en;
    if (ok) {
        int enn = ozl_entropy_layout(&p, en, st, &ns);
        ok = enn == nn;
        for (int k = 0; ok && k < nn; k++) ok = en[k].trid == nodes[k].trid;
        if (ok) {
            size_t chl = ozl_write_chunk_header(L.expect, en, enn, st, ns, order, &nbSt2);
            ok = nbSt2 == nbStored && d.chunkOff + chl <= flen;
            for (size_t k = 0; ok && k < chl; k++) ok = L.expect[k] == fr[d.chunkOff + k];
            if (ok) {
                size_t q = d.chunkOff + chl;
                for (int k = 0; k < 12; k++) soff[k] = 0;
                for (int k = 0; k < nbStored; k++) { soff[order[k]] = (uint32_t)q; q += st[order[k]].size; }
                d.csumOff = (uint32_t)q;
                ok = q + 8 + 1 == flen && fr[flen - 1] == 0;
            }
        }
    }
    // tables
    if (ok) {
        d.bk = p.bk;
        d.status = 0;
        switch (p.bk) {
        case OZL_BK_STORE: d.payOff = soff[0]; d.payLen = d.n; break;
        case OZL_BK_CONSTANT: d.auxOff = soff[1]; d.auxLen = 1; break;
        case OZL_BK_FLATPACK:
            d.auxOff = soff[1]; d.auxLen = st[1].size; d.payOff = soff[2]; d.payLen = st[2].size;
            if (d.auxLen < 2 || d.auxLen > 256) d.status = OZL_ERR_CORRUPTION;
            break;
        case OZL_BK_BITPACK:
            d.aux = p.aux; d.payOff = soff[1]; d.payLen = st[1].size;
            if (d.aux < 1 || d.aux > 8 || d.payLen != ((uint64_t)d.n * d.aux + 7) / 8) d.status = OZL_ERR_CORRUPTION;
            break;
        case OZL_BK_FSE: {
            d.payOff = soff[2]; d.payLen = st[2].size; d.nbStates = p.nbStates;
            unsigned ms, tl;
            if (nodes[1].trid == OZL_TR_FSE_NCOUNT) {
                if (haveNc && nc.off == soff[3] && nc.size == st[3].size) {
                    ms = nc.ms;   // read by the header pass (which succeeded)
                    tl = nc.tl;
                } else if (ozl_fse_read_ncount(L.norm, &ms, &tl, fr.at(soff[3], st[3].size), st[3].size) < 0) {
                    d.status = OZL_ERR_CORRUPTION;
                    break;
                }
            } else {  // raw int16 counts after a num->serial conversion
                uint32_t cnt = st[3].size / 2, sum = 0;
                for (int s = 0; s < 256; s++) L.norm[s] = 0;
                for (uint32_t s = 0; s < cnt; s++) {
                    L.norm[s] = (int16_t)(fr[soff[3] + 2 * s] | (fr[soff[3] + 2 * s + 1] << 8));
                    if (L.norm[s] < -1) sum = 3;  // not a normalized count: rejected below
                    sum += L.norm[s] == -1 ? 1 : (uint32_t)L.norm[s];
                }
                if (cnt == 0 || cnt > 256 || sum == 0 || (sum & (sum - 1))) { d.status = OZL_ERR_CORRUPTION; break; }
                ms = cnt - 1;
                tl = (unsigned)ozl_hb32(sum);
            }
            if (tl < OZL_FSE_MIN_TABLELOG || tl > 11 || (d.nbStates != 2 && d.nbStates != 4)) { d.status = OZL_ERR_CORRUPTION; break; }
            d.tl = (uint8_t)tl;
            d.aux = ms;
            break;
        }
        case OZL_BK_HUF: {
            d.payOff = soff[2]; d.payLen = st[2].size; d.nbStates = p.nbStates;
            if (d.nbStates > 1 || (d.nbStates == 1 && d.n < 12)) { d.status = OZL_ERR_CORRUPTION; break; }
            uint32_t m = p.maxSym + 1u;
            if (ozl_decode_weights(fr, nodes, nn, st, soff, m, L, haveNc ? &nc : nullptr)) { d.status = OZL_ERR_CORRUPTION; break; }
            if (weightsOut) { *weightsOut = m; break; }
            int tl = ozl_huf_from_weights(L.wlen, L.wval, L.w, m, L.x.rk);
            if (tl < 1 || tl > 11) { d.status = OZL_ERR_CORRUPTION; break; }
            d.tl = (uint8_t)tl;
            break;
        }
        default: break;
        }
    }
}
