// Batched decoding of FSE / Huffman graph frames (the universal decompressor
// path of src/openzl/decompress/decompress2.c:1442-1580 restricted to the
// single-chunk frames ZL_GRAPH_FSE and ZL_GRAPH_HUFFMAN produce for inputs
// below 100000 bytes).
//
//   k_dparse       one workgroup per frame: lane 0 parses the frame and chunk
//                  headers, recognises the node layout, re-serialises the
//                  expected chunk header from what it recognised and compares
//                  (any difference -> frame reported unsupported), decodes the
//                  NCount / Huffman weights (including the nested weight
//                  graph) and builds the FSE decoding table or the Huffman
//                  lookup table
//   k_fse_decode   one 4-lane group per frame, one lane per interleaved state
//                  (fse_decompress.c:255-338), table in LDS
//   k_huf_decode   one lane per Huffman stream, lookup table in LDS
//   k_pack_decode  store / constant / flatpack / bitpack
//   k_dverify      one wavefront per frame: content and compressed checksums
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_dparse.h"
#include "ozl_profile.h"


struct ozl_dws {
    ozl_dplan* plans;
    uint8_t* tables;
};

namespace {

constexpr int kT = 256;

__global__ void __launch_bounds__(kT) k_dparse(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                               uint32_t B, uint8_t* __restrict__ dst, const uint64_t* __restrict__ doff,
                                               ozl_dws ws)
{
    __shared__ ozl_dlds L;
    __shared__ uint16_t lut[4096];
    __shared__ int sFill[3];  // tl, filled flag
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    const uint8_t* fr = frames + foff[b];
    size_t flen = foff[b + 1] - foff[b];
    uint8_t* tab = ws.tables + (size_t)b * OZL_DTAB_BYTES;
    if (threadIdx.x == 0) {
        ozl_dplan d;
        int hufTl;
        ozl_dparse_frame(fr, flen, tab, L, d, &hufTl);
        sFill[0] = hufTl;
        sFill[1] = hufTl > 0;
        d.srcPtr = (uint64_t)(uintptr_t)(fr + d.payOff);
        d.outPtr = (uint64_t)(uintptr_t)(dst + doff[b]);
        ws.plans[b] = d;
        if (d.status == 0 && doff[b + 1] - doff[b] != d.n) ws.plans[b].status = OZL_ERR_DSTCAPACITY_TOOSMALL;
    }
    __syncthreads();
    if (sFill[1]) {  // Huffman lookup: entry = symbol | length << 8, indexed by the next tl bits
        int tl = sFill[0];
        for (int s = threadIdx.x; s < 256; s += kT) {
            uint32_t l = L.wlen[s];
            if (!l) continue;
            uint32_t first = (uint32_t)L.wval[s] << (tl - l), cnt = 1u << (tl - l);
            for (uint32_t k = 0; k < cnt; k++) lut[first + k] = (uint16_t)(s | (l << 8));
        }
        __syncthreads();
        for (int k = threadIdx.x; k < (1 << tl); k += kT) reinterpret_cast<uint16_t*>(tab)[k] = lut[k];
    }
}

// reads nb (<= 16) bits ending (exclusive) at bit position `end` of a stream,
// zero-filled below bit 0 (the reference's reader pads the same way)
__device__ inline uint32_t read_bits_back(const uint8_t* s, long end, uint32_t nb)
{
    long start = end - (long)nb;
    uint32_t v = 0;
    long b0 = start >> 3;
    for (long byte = b0; byte <= ((end - 1) >> 3); byte++) {
        uint32_t x = byte >= 0 ? s[byte] : 0;
        long sh = byte * 8 - start;
        v |= sh >= 0 ? (x << sh) : (x >> (-sh));
    }
    return nb ? (v & ((1u << nb) - 1)) : 0;
}

constexpr int kDecGroups = 16;

__global__ void __launch_bounds__(64) k_fse_decode(uint32_t B, ozl_dws ws)
{
    __shared__ uint32_t sDT[kDecGroups][2048];
    const int lane = threadIdx.x & 63, g = lane >> 2, q = lane & 3;
    const uint32_t b = blockIdx.x * kDecGroups + g;
    bool active = b < B;
    ozl_dplan d;
    if (active) {
        d = ws.plans[b];
        active = d.status == 0 && d.bk == OZL_BK_FSE;
    }
    if (active) {
        const uint32_t* dt = reinterpret_cast<const uint32_t*>(ws.tables + (size_t)b * OZL_DTAB_BYTES);
        for (unsigned i = q; i < (1u << d.tl); i += 4) sDT[g][i] = dt[i];
    }
    __syncthreads();
    if (!active) return;
    const uint32_t* DT = sDT[g];
    const uint8_t* s = (const uint8_t*)(uintptr_t)d.srcPtr;
    uint8_t* out = (uint8_t*)(uintptr_t)d.outPtr;
    const uint32_t S = d.nbStates, n = d.n, tl = d.tl;
    const int base = lane & ~3;
    if (d.payLen == 0 || s[d.payLen - 1] == 0) { if (q == 0) ws.plans[b].status = OZL_ERR_CORRUPTION; return; }
    long pos = (long)(d.payLen - 1) * 8 + ozl_hb32(s[d.payLen - 1]);
    // initial states: state j is the j-th tl-bit field read backwards
    uint32_t state = 0;
    if (q < (int)S) state = read_bits_back(s, pos - (long)q * tl, tl);
    pos -= (long)S * tl;
    const uint32_t rounds = (n + S - 1) / S;
    for (uint32_t r = 0; r < rounds; r++) {
        uint32_t p = r * S + q;
        bool mine = q < (int)S && p < n;
        uint32_t e = mine ? DT[state] : 0;
        // every symbol but the last S moves its state, also with 0 bits to read
        bool next = mine && p + S < n;
        uint32_t nb = next ? (e >> 8) & 255 : 0;
        // read order inside the round = lane order
        uint32_t n0 = __shfl(nb, base + 0, 64), n1 = __shfl(nb, base + 1, 64);
        uint32_t n2 = __shfl(nb, base + 2, 64);
        uint32_t before = q == 0 ? 0 : q == 1 ? n0 : q == 2 ? n0 + n1 : n0 + n1 + n2;
        uint32_t all = __shfl(before + nb, base + (int)S - 1, 64);
        if (mine) {
            out[p] = (uint8_t)e;
            if (next) state = (e >> 16) + read_bits_back(s, pos - (long)before, nb);
        }
        pos -= (long)all;
    }
    if (q == 0 && pos != 0) ws.plans[b].status = OZL_ERR_CORRUPTION;
}

__global__ void __launch_bounds__(64) k_huf_decode(uint32_t B, ozl_dws ws)
{
    __shared__ uint16_t sL[kDecGroups][2048];
    const int lane = threadIdx.x & 63, g = lane >> 2, q = lane & 3;
    const uint32_t b = blockIdx.x * kDecGroups + g;
    bool active = b < B;
    ozl_dplan d;
    if (active) {
        d = ws.plans[b];
        active = d.status == 0 && d.bk == OZL_BK_HUF && d.tl <= 11;
    }
    if (active) {
        const uint16_t* lt = reinterpret_cast<const uint16_t*>(ws.tables + (size_t)b * OZL_DTAB_BYTES);
        for (unsigned i = q; i < (1u << d.tl); i += 4) sL[g][i] = lt[i];
    }
    __syncthreads();
    if (!active) return;
    const uint8_t* s = (const uint8_t*)(uintptr_t)d.srcPtr;
    uint8_t* out = (uint8_t*)(uintptr_t)d.outPtr;
    const uint32_t n = d.n, tl = d.tl;
    const int x4 = d.nbStates;
    if (!x4 && q > 0) return;
    uint32_t seg = x4 ? (n + 3) / 4 : n;
    uint32_t cnt = x4 ? (q < 3 ? seg : n - 3 * seg) : n;
    uint32_t sz[4] = { d.payLen, 0, 0, 0 };
    uint32_t start = 0;
    if (x4) {
        if (d.payLen < 6) { if (q == 0) ws.plans[b].status = OZL_ERR_CORRUPTION; return; }
        sz[0] = s[0] | (s[1] << 8);
        sz[1] = s[2] | (s[3] << 8);
        sz[2] = s[4] | (s[5] << 8);
        if (6u + sz[0] + sz[1] + sz[2] > d.payLen) { ws.plans[b].status = OZL_ERR_CORRUPTION; return; }
        sz[3] = d.payLen - 6 - sz[0] - sz[1] - sz[2];
        start = 6;
        for (int k = 0; k < q; k++) start += sz[k];
    }
    const uint8_t* st = s + start;
    uint32_t len = sz[q];
    if (len == 0 || st[len - 1] == 0) { ws.plans[b].status = OZL_ERR_CORRUPTION; return; }
    long pos = (long)(len - 1) * 8 + ozl_hb32(st[len - 1]);
    uint8_t* o = out + (x4 ? (size_t)q * seg : 0);
    const uint16_t* LT = sL[g];
    for (uint32_t i = 0; i < cnt; i++) {
        uint32_t e = LT[read_bits_back(st, pos, tl)];
        o[i] = (uint8_t)e;
        pos -= (long)(e >> 8);
    }
    if (pos != 0) ws.plans[b].status = OZL_ERR_CORRUPTION;
}

__global__ void __launch_bounds__(kT) k_pack_decode(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                    uint32_t B, uint8_t* __restrict__ dst, const uint64_t* __restrict__ doff,
                                                    ozl_dws ws)
{
    __shared__ uint8_t alpha[256];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_dplan d = ws.plans[b];
    if (d.status != 0 || d.bk == OZL_BK_FSE || d.bk == OZL_BK_HUF) return;
    const uint8_t* fr = frames + foff[b];
    uint8_t* out = dst + doff[b];
    const uint32_t n = d.n;
    if (d.bk == OZL_BK_STORE) {
        for (uint32_t i = threadIdx.x; i < n; i += kT) out[i] = fr[d.payOff + i];
        return;
    }
    if (d.bk == OZL_BK_CONSTANT) {
        uint8_t v = fr[d.auxOff];
        for (uint32_t i = threadIdx.x; i < n; i += kT) out[i] = v;
        return;
    }
    uint32_t nb;
    if (d.bk == OZL_BK_FLATPACK) {
        for (uint32_t i = threadIdx.x; i < d.auxLen; i += kT) alpha[i] = fr[d.auxOff + i];
        nb = d.auxLen <= 1 ? d.auxLen : 1 + (uint32_t)ozl_hb32(d.auxLen - 1);
        if (d.payLen != 1 + nb * n / 8) { if (threadIdx.x == 0) ws.plans[b].status = OZL_ERR_CORRUPTION; return; }
    } else {
        for (uint32_t i = threadIdx.x; i < 256; i += kT) alpha[i] = (uint8_t)i;
        nb = d.aux;
    }
    __syncthreads();
    const uint8_t* pk = fr + d.payOff;
    for (uint32_t i = threadIdx.x; i < n; i += kT) {
        uint64_t bit = (uint64_t)i * nb;
        uint32_t v = (uint32_t)(pk[bit >> 3] | ((bit & 7) + nb > 8 ? (uint32_t)pk[(bit >> 3) + 1] << 8 : 0u));
        v = (v >> (bit & 7)) & ((1u << nb) - 1);
        out[i] = alpha[v & 255];
    }
}

__global__ void __launch_bounds__(64) k_dverify(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                uint32_t B, const uint8_t* __restrict__ dst,
                                                const uint64_t* __restrict__ doff, ozl_dws ws, int* status)
{
    __shared__ uint64_t lds[8 * 128];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_dplan d = ws.plans[b];
    int st = (int)d.status;
    if (st == 0) {
        const uint8_t* fr = frames + foff[b];
        uint64_t hc = ozl_xxh3_wave(fr + d.chunkOff, d.csumOff + 4 - d.chunkOff, lds, 128);
        uint64_t hd = ozl_xxh3_wave(dst + doff[b], d.n, lds, 128);
        if ((uint32_t)hc != ozl_le32(fr + d.csumOff + 4)) st = OZL_ERR_COMPRESSED_CHECKSUM;
        else if ((uint32_t)hd != ozl_le32(fr + d.csumOff)) st = OZL_ERR_CONTENT_CHECKSUM;
    }
    if (threadIdx.x == 0) status[b] = st;
}

// ---------------------------------------------------------------- items of the graph decoder
// One fse_v2 / huffman_v2 decoder of a generic frame: its side stream
// (normalized counts as int16, or Huffman weights as u8, both already
// regenerated in device memory), its bit stream and its output.
}  // namespace

struct ozl_ditem {
    uint64_t side;
    uint64_t bits;
    uint64_t out;
    uint32_t sideN;    // number of side elements
    uint32_t bitsLen;
    uint32_t n;        // symbols to decode
    uint8_t huf;       // 0 FSE, 1 Huffman
    uint8_t nbStates;  // FSE states, or the Huffman 4-stream flag
    uint8_t pad[2];
};

namespace {

__global__ void __launch_bounds__(64) k_prep_items(const ozl_ditem* __restrict__ items, uint32_t B, ozl_dws ws)
{
    __shared__ uint8_t cells[4096];
    __shared__ int16_t norm[256];
    __shared__ uint8_t w[256];
    __shared__ uint8_t len[256];
    __shared__ uint16_t val[256];
    __shared__ int sTl;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_ditem it = items[b];
    uint8_t* tab = ws.tables + (size_t)b * OZL_DTAB_BYTES;
    if (threadIdx.x == 0) {
        ozl_dplan d;
        uint32_t* dw = reinterpret_cast<uint32_t*>(&d);
        for (size_t k = 0; k < sizeof(d) / 4; k++) dw[k] = 0;
        d.n = it.n;
        d.payLen = it.bitsLen;
        d.nbStates = it.nbStates;
        d.srcPtr = it.bits;
        d.outPtr = it.out;
        d.bk = it.huf ? OZL_BK_HUF : OZL_BK_FSE;
        sTl = 0;
        const uint8_t* side = (const uint8_t*)(uintptr_t)it.side;
        if (!it.huf) {
            uint32_t sum = 0;
            for (int s = 0; s < 256; s++) norm[s] = 0;
            for (uint32_t s = 0; s < it.sideN && s < 256; s++) {
                norm[s] = (int16_t)(side[2 * s] | (side[2 * s + 1] << 8));
                sum += norm[s] == -1 ? 1 : (uint32_t)(uint16_t)norm[s];
            }
            unsigned tl = sum ? (unsigned)ozl_hb32(sum) : 0;
            if (it.sideN == 0 || it.sideN > 256 || sum == 0 || (sum & (sum - 1)) || tl < OZL_FSE_MIN_TABLELOG ||
                tl > 11 || (it.nbStates != 2 && it.nbStates != 4))
                d.status = OZL_ERR_CORRUPTION;
            else {
                d.tl = (uint8_t)tl;
                if (ozl_fse_build_dtable(reinterpret_cast<uint32_t*>(tab), cells, norm, it.sideN - 1, tl))
                    d.status = OZL_ERR_CORRUPTION;
            }
        } else {
            for (uint32_t s = 0; s < it.sideN && s < 256; s++) w[s] = side[s];
            int tl = it.sideN <= 256 ? ozl_huf_from_weights(len, val, w, it.sideN) : -1;
            if (tl < 1 || tl > 11 || it.nbStates > 1 || (it.nbStates == 1 && it.n < 12)) d.status = OZL_ERR_CORRUPTION;
            else { d.tl = (uint8_t)tl; sTl = tl; }
        }
        ws.plans[b] = d;
    }
    __syncthreads();
    int tl = sTl;
    if (tl) {
        uint16_t* lut = reinterpret_cast<uint16_t*>(tab);
        for (int s = threadIdx.x; s < 256; s += 64) {
            uint32_t l = len[s];
            if (!l || s >= (int)it.sideN) continue;
            uint32_t first = (uint32_t)val[s] << (tl - l), cnt = 1u << (tl - l);
            for (uint32_t k = 0; k < cnt; k++) lut[first + k] = (uint16_t)(s | (l << 8));
        }
    }
}

__global__ void k_item_status(uint32_t B, ozl_dws ws, int* status)
{
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) status[b] = (int)ws.plans[b].status;
}

}  // namespace

extern "C" {

size_t ozl_entropy_decompress_workspace(uint32_t B)
{
    return (size_t)B * (sizeof(ozl_dplan) + OZL_DTAB_BYTES) + 4096;
}

// Decodes B frames written by ZL_GRAPH_FSE / ZL_GRAPH_HUFFMAN on serial inputs
// below 100000 bytes (the universal decompressor's ZL_decompress on each
// frame, include/openzl/zl_decompress.h:32). Frame b is
// frames[foff[b], foff[b+1]); its content goes to dst[doff[b], doff[b+1]),
// which must be exactly the decoded size. status[b] (device int) receives 0 or
// the ZL error code of that frame (corruption, checksum errors, or
// frameParameter_unsupported-style rejection of other layouts, reported as
// corruption). All pointers are device pointers.
int ozl_entropy_decompress_batch(const void* frames, const uint64_t* foff, uint32_t B, void* dst,
                                 const uint64_t* doff, int* status, void* workspace, size_t wsSize, void* stream)
{
    if (B == 0) return OZL_OK;
    if (wsSize < ozl_entropy_decompress_workspace(B)) return OZL_ERR_INTERNAL_BUFFER;
    hipStream_t st = (hipStream_t)stream;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    ozl_dws ws;
    ws.plans = (ozl_dplan*)w; w += (size_t)B * sizeof(ozl_dplan);
    ws.tables = w;
    const uint8_t* f = (const uint8_t*)frames;
    uint8_t* o = (uint8_t*)dst;
    OZL_LAUNCH("k_dparse", k_dparse, dim3(B), dim3(kT), 0, st, f, foff, B, o, doff, ws);
    unsigned gq = (B + kDecGroups - 1) / kDecGroups;
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3(gq), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3(gq), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_pack_decode", k_pack_decode, dim3(B), dim3(kT), 0, st, f, foff, B, o, doff, ws);
    OZL_LAUNCH("k_dverify", k_dverify, dim3(B), dim3(64), 0, st, f, foff, B, (const uint8_t*)o, doff, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"

size_t ozl_entropy_items_workspace(uint32_t B) { return ozl_entropy_decompress_workspace(B); }

// Decodes B fse_v2 / huffman_v2 items (device array) of a generic frame;
// status (device, B ints) receives 0 or an error code per item.
int ozl_entropy_decode_items(const ozl_ditem* items, uint32_t B, int* status, void* workspace, size_t wsSize,
                             hipStream_t st)
{
    if (B == 0) return OZL_OK;
    if (wsSize < ozl_entropy_decompress_workspace(B)) return OZL_ERR_INTERNAL_BUFFER;
    uint8_t* w = (uint8_t*)(((uintptr_t)workspace + 255) & ~(uintptr_t)255);
    ozl_dws ws;
    ws.plans = (ozl_dplan*)w; w += (size_t)B * sizeof(ozl_dplan);
    ws.tables = w;
    OZL_LAUNCH("k_prep_items", k_prep_items, dim3(B), dim3(64), 0, st, items, B, ws);
    unsigned gq = (B + kDecGroups - 1) / kDecGroups;
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3(gq), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3(gq), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_item_status", k_item_status, dim3((B + 63) / 64), dim3(64), 0, st, B, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
