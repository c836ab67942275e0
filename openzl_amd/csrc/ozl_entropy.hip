// Batched FSE / Huffman dynamic graphs on gfx950: one frame per input block.
//
// Pipeline for B blocks (each < 100000 bytes, so the graphs do not split it,
// src/openzl/codecs/entropy/encode_entropy_binding.c:491-519):
//   k_hist          one workgroup per block: LDS histogram (k_hist_hash in
//                   the batch API: the histogram and the content checksum
//                   in one read of the block)
//   k_plan          one wave per block: one lane decides the backend
//                   (:601-753) and builds the FSE tables or Huffman code, the
//                   NCount, and the nested graph of the Huffman weights
//                   (ozl_core.h); the wave builds the FSE encoding table
//   k_fse_encode    one 4-lane group per block, one lane per interleaved state,
//                   8 blocks per wavefront; a round's bit offsets come from a
//                   quad prefix sum (DPP), tables sit in LDS
//   k_huf_encode    one workgroup per block: code lengths, a block-wide scan of
//                   bit offsets in reverse symbol order, bits OR-ed into LDS
//   k_pack_encode   flatpack / bitpack payloads, one output byte per lane
//   k_frame_size -> k_scan -> k_frame_write_hash (frames and compressed
//                   checksums in one pass)
#include <stddef.h>
#include <type_traits>

#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_profile.h"
#include "ozl_gmem.h"
#include "ozl_entropy_api.h"
#include "ozl_dparse.h"
#include "ozl_wave.h"

namespace {

constexpr int kT = 256;

// 16 copies of the histogram, bin-major (copy c of bin s at 16s + c), so that
// the lanes of a wave that hit the same frequent symbol spread over 16 banks
// and 16 addresses instead of serializing on one.
constexpr int kHistCopies = 16;

__device__ inline void block_hist(const uint8_t* src, uint32_t n, uint32_t* hc /*16*256*/, uint32_t* count)
{
    int tid = threadIdx.x;
    for (int i = tid; i < kHistCopies * 256; i += kT) hc[i] = 0;
    __syncthreads();
    uint32_t* h = hc + (tid & (kHistCopies - 1));
    // aligned 16-byte body, byte head and tail
    uint32_t head = (uint32_t)((16 - ((uintptr_t)src & 15)) & 15);
    if (head > n) head = n;
    for (uint32_t i = tid; i < head; i += kT) atomicAdd(&h[kHistCopies * src[i]], 1u);
    uint32_t nv = (n - head) / 16;
    const uint4* v = reinterpret_cast<const uint4*>(src + head);
    for (uint32_t i = tid; i < nv; i += kT) {
        uint4 q = v[i];
        uint32_t w[4] = { q.x, q.y, q.z, q.w };
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int b = 0; b < 4; b++) atomicAdd(&h[kHistCopies * ((w[k] >> (8 * b)) & 255)], 1u);
    }
    for (uint32_t i = head + nv * 16 + tid; i < n; i += kT) atomicAdd(&h[kHistCopies * src[i]], 1u);
    __syncthreads();
    for (int s = tid; s < 256; s += kT) {
        const uint4* r = reinterpret_cast<const uint4*>(hc + kHistCopies * s);
        uint32_t t = 0;
#pragma unroll
        for (int c = 0; c < kHistCopies / 4; c++) t += r[c].x + r[c].y + r[c].z + r[c].w;
        count[s] = t;
    }
    __syncthreads();
}

struct PlanLds {
    uint32_t count[256];
    int16_t norm[256];
    union {
        ozl_huf_scratch hs;                         // Huffman code
        ozl_wave::FseBuild fb;                      // main FSE table build
    };
    uint8_t wts[256];
    uint32_t wcount[256];
    uint8_t cells[1 << OZL_WEIGHTS_MAX_TABLELOG];   // nested weights graph
    uint16_t wst[1 << OZL_WEIGHTS_MAX_TABLELOG];
    uint32_t wtt[16];       // weights are <= 12: 13 symbols at most
    int16_t wn[16];
    uint8_t hlen[256];
    uint16_t hval[256];
    uint8_t aux[OZL_AUX_BYTES];
    uint16_t cumul[18];   // nested table build workspace (13 weight symbols)
    uint8_t fmap[256];
    uint32_t wWave;       // the weights' FSE bitstream is left to the wave
    uint32_t nPart[4];    // FSE normalisation: lane 0's partial results
    ozl_eplan plan;
};

// The weights' nested FSE graph keeps tableLog 5 (<= 256 weights of value
// <= 12, fse_compress.c:379-396); the wave encoder's transition table covers
// that size.
#define OZL_WEIGHTS_WAVE_TABLELOG 5

// The table area of a block holds its histogram (u32[256]) at this offset
// between k_hist and k_plan; the FSE tables end below it.
#define OZL_TAB_COUNT_OFF 9216

// Nested FSE graph on the Huffman weights (entropyCompressChunk in EBM_fse
// mode on the num->serial converted weights stream).
// L.wcount holds the histogram of the m weights (every other bin zero).
__device__ __forceinline__ void plan_weights(ozl_eplan& p, PlanLds& L, uint8_t* aux, int minSS)
{
    uint32_t m = p.maxSym + 1u;
    if (ozl_small(m, minSS)) { p.wbk = 255; return; }
    // ozl_decide_from_hist over the 13 bins weights can fill
    unsigned wms = 0, wcard = 0;
    for (unsigned v = 0; v <= 12; v++)
        if (L.wcount[v]) { wms = v; wcard++; }
    int wbk = m <= 1 ? OZL_BK_STORE : wcard == 1 ? OZL_BK_CONSTANT
                                               : ozl_decide(ozl_entropy(L.wcount, wms, m), wms, wcard, m, 0);
    p.wbk = (uint8_t)wbk;
    p.wMaxSym = (uint8_t)wms;
    uint8_t* side = aux + OZL_AUX_W_SIDE;
    uint8_t* pay = aux + OZL_AUX_W_PAYLOAD;
    switch (wbk) {
    case OZL_BK_CONSTANT: pay[0] = L.wts[0]; p.wPayload = 1; break;
    case OZL_BK_FLATPACK: {
        size_t ps;
        p.wAux = (uint16_t)ozl_flatpack_encode_seq(side, pay, &ps, L.wts, m, L.wcount, L.fmap);
        p.wPayload = (uint16_t)ps;
        break;
    }
    case OZL_BK_BITPACK: {
        int nb = 1 + (wms ? ozl_hb32(wms) : 0);
        p.wAux8 = (uint8_t)nb;
        p.wPayload = (uint16_t)ozl_bitpack8_seq(pay, L.wts, m, nb);
        break;
    }
    case OZL_BK_FSE: {
        unsigned tl = ozl_fse_optimal_tablelog(OZL_FSE_DEFAULT_TABLELOG, m, wms, 2);
        p.wTl = (uint8_t)tl;
        int16_t* wn = L.wn;
        // at most 256 weights of value <= 12: tableLog <= 5 (fse_compress.c:379-396)
        if (tl > OZL_WEIGHTS_MAX_TABLELOG || wms > 12) { p.status = OZL_ERR_GENERIC; return; }
        if (ozl_fse_normalize(wn, tl, L.wcount, m, wms) != 0) { p.status = OZL_ERR_GENERIC; return; }
        if (ozl_fse_build_ctable(L.wst, L.wtt, L.cells, wn, wms, tl, L.cumul)) { p.status = OZL_ERR_GENERIC; return; }
        if (tl <= OZL_WEIGHTS_WAVE_TABLELOG && m > 2) {
            L.wWave = 1;   // wave_weights_fse writes the payload
        } else {
            size_t bits = ozl_fse_encode_seq(pay, OZL_AUX_BYTES - OZL_AUX_W_PAYLOAD, L.wts, m, L.wst, L.wtt, tl, 2);
            if (bits == 0) { p.status = OZL_ERR_GENERIC; return; }
            p.wPayload = (uint16_t)bits;
        }
        if (ozl_small(2u * (wms + 1u), minSS)) {
            for (unsigned s = 0; s <= wms; s++) { side[2 * s] = (uint8_t)wn[s]; side[2 * s + 1] = (uint8_t)((uint16_t)wn[s] >> 8); }
            p.wAux = (uint16_t)(2 * (wms + 1));
        } else {
            p.wAux = (uint16_t)ozl_fse_write_ncount(side, wn, wms, tl);
        }
        break;
    }
    default: break;
    }
}

// The weights' FSE bitstream by the wave: the bytes ozl_fse_encode_seq
// writes for m > 2 weights with two states (fse_compress.c:568-627 over
// the weights stream). Position p is encoded by the state of positions of its
// parity, from the last position down; every emission (the state's low bits
// before its transition) has a fixed place in the stream: main emissions for
// p = m-3 .. 0, the final states of positions 1 and 0 (tableLog bits each),
// then a 1 bit. Two lanes run the two state chains over a transition table
// of (state, weight) pairs; the wave adds up the bit offsets and places the
// bits. Needs tl <= OZL_WEIGHTS_WAVE_TABLELOG; the table, emissions and
// words reuse the Huffman build's node table, which is dead by then.
__device__ __forceinline__ void wave_weights_fse(PlanLds& L, uint8_t* pay)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t m = L.plan.maxSym + 1u, tl = L.plan.wTl, wms = L.plan.wMaxSym;
    const uint32_t ts = 1u << tl;
    uint16_t* ct = reinterpret_cast<uint16_t*>(L.hs.tbl);   // [state - ts][weight]: next | nbBits << 8
    uint16_t* em = ct + 16 * (1u << OZL_WEIGHTS_WAVE_TABLELOG);   // [m + 1]: value | nbBits << 8
    uint32_t* words = reinterpret_cast<uint32_t*>(em + 264);      // the stream, 32 bits a word
    static_assert(sizeof(uint16_t) * (16 * 32 + 264) + 4 * 48 <= sizeof(ozl_huf_scratch::tbl), "weights area");
    for (uint32_t i = lane; i < 16 * ts; i += 64) {
        const uint32_t u = i >> 4, w = i & 15u;
        uint16_t e = 0;
        if (w <= wms && L.wn[w] != 0) {
            const uint32_t t = L.wtt[w], x = ts + u;
            const uint32_t nb = (x + ozl_tt_dnb(t)) >> 16;
            e = (uint16_t)(L.wst[(int)(x >> nb) + ozl_tt_dfs(t)] | (nb << 8));
        }
        ct[i] = e;
    }
    for (uint32_t i = lane; i < 48; i += 64) words[i] = 0;
    __syncthreads();
    if (lane < 2) {
        // lane c: the state started on position m-1-c
        int p = (int)m - 1 - (int)lane;
        const uint32_t t = L.wtt[L.wts[p]];
        const uint32_t dnb = ozl_tt_dnb(t);
        const uint32_t nb0 = (dnb + (1u << 15)) >> 16;
        const uint32_t v = (nb0 << 16) - dnb;
        uint32_t x = L.wst[(int)(v >> nb0) + ozl_tt_dfs(t)];
        p -= 2;
        uint32_t w = p >= 0 ? L.wts[p] : 0u;
        for (; p >= 0; p -= 2) {
            const uint32_t wNext = p >= 2 ? L.wts[p - 2] : 0u;
            const uint32_t e = ct[((x - ts) << 4) + w];
            const uint32_t nb = e >> 8;
            em[m - 3 - (uint32_t)p] = (uint16_t)(__builtin_amdgcn_ubfe(x, 0, nb) | (nb << 8));
            x = e & 255u;
            w = wNext;
        }
        const uint32_t odd = (m - 1u - lane) & 1u;
        em[odd ? m - 2 : m - 1] = (uint16_t)((x & (ts - 1u)) | (tl << 8));
        if (lane == 0) em[m] = (uint16_t)(1u | (1u << 8));
    }
    __syncthreads();
    // five emissions a lane (m + 1 <= 257), offsets by a wave scan
    uint32_t loc[5], nbs = 0;
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const uint32_t e = lane * 5u + (uint32_t)k;
        loc[k] = e <= m ? em[e] : 0u;
        nbs += loc[k] >> 8;
    }
    uint32_t inc = nbs;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += y;
    }
    uint32_t P = inc - nbs;
    const uint32_t totalBits = __shfl(inc, 63, 64);
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const uint32_t nb = loc[k] >> 8, val = loc[k] & 255u, sh = P & 31u;
        if (nb) {
            atomicOr(&words[P >> 5], val << sh);
            if (sh + nb > 32) atomicOr(&words[(P >> 5) + 1], val >> (32 - sh));
        }
        P += nb;
    }
    __syncthreads();
    const uint32_t bytes = (totalBits + 7) >> 3;
    if (bytes + 8 > OZL_AUX_BYTES - OZL_AUX_W_PAYLOAD) {
        if (lane == 0) L.plan.status = OZL_ERR_GENERIC;
        return;
    }
    for (uint32_t i = lane; i < bytes; i += 64) pay[i] = (uint8_t)(words[i >> 2] >> (8 * (i & 3)));
    if (lane == 0) L.plan.wPayload = (uint16_t)bytes;
}

// ozl_fse_normalize (FSE_normalizeCount with low-probability counts,
// fse_compress.c:487-549) by the wave: every symbol's count is its own
// (lane s mod 64 takes symbols s, s+64, ...); the counts' sum, the constant
// test and the largest count (first symbol among equal ones, as the
// sequential strict comparison keeps) are wave reductions. Lane 0 then
// applies the correction to the largest, or runs ozl_fse_normalize_m2 when
// the reference does. Returns ozl_fse_normalize's result in every lane.
__device__ int wave_fse_normalize(PlanLds& L, uint32_t total, uint32_t maxSym, uint32_t tl)
{
    const uint32_t lane = threadIdx.x;
    if (tl < OZL_FSE_MIN_TABLELOG || tl > OZL_FSE_MAX_TABLELOG || tl < ozl_fse_min_tablelog(total, maxSym)) return -1;
    const uint64_t scale = 62 - tl;
    const uint64_t step = (1ull << 62) / total;
    const uint64_t vStep = 1ull << (scale - 20);
    const uint32_t lowThreshold = total >> tl;
    int used = 0;             // counts given out (a -1 takes one)
    uint32_t constant = 0, key = 255;   // key: largest count << 8 | (255 - symbol)
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t sy = lane + 64u * k;
        if (sy > maxSym) continue;
        const uint32_t c = L.count[sy];
        int16_t v;
        if (c == total) constant = 1;
        if (c == 0) v = 0;
        else if (c <= lowThreshold) v = -1;
        else {
            const uint64_t cs = (uint64_t)c * step;
            int p = (int)(cs >> scale);
            if (p < 8) p += cs - ((uint64_t)p << scale) > vStep * ozl_fse_rtb(p) ? 1 : 0;
            v = (int16_t)p;
            const uint32_t kk = ((uint32_t)p << 8) | (255u - sy);
            key = kk > key ? kk : key;
        }
        used += v < 0 ? 1 : v;
        L.norm[sy] = v;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        used += __shfl_xor(used, d, 64);
        constant |= __shfl_xor(constant, d, 64);
        const uint32_t o = __shfl_xor(key, d, 64);
        key = o > key ? o : key;
    }
    __syncthreads();
    if (lane == 0) {
        int rc = 0;
        if (constant) rc = 1;
        else {
            const int still = (1 << tl) - used;
            const uint32_t largest = 255u - (key & 255u);
            if (-still >= (L.norm[largest] >> 1)) rc = ozl_fse_normalize_m2(L.norm, tl, L.count, total, maxSym);
            else L.norm[largest] = (int16_t)(L.norm[largest] + still);
        }
        L.nPart[0] = (uint32_t)rc;
    }
    __syncthreads();
    return (int)L.nPart[0];
}

// ozl_fse_write_ncount (FSE_writeNCount_generic, fse_compress.c:246-357) by
// the wave. The header is the LSB-first concatenation of: tableLog - 5 in 4
// bits; for each symbol up to the one that brings `remaining` to 1, its
// field, except zeros that follow a zero, whose run is a repeat field written
// before the next nonzero symbol (16 one-bits per 24 zeros, 2 one-bits per 3
// more, then the rest in 2 bits). A symbol's field depends only on
// remaining before it (2^tl + 1 minus the counts before it), which a wave
// scan gives: threshold = the largest power of two <= remaining (at most
// 2^tl), max = 2 threshold - 1 - remaining, value count + 1 (+ max when it
// reaches threshold), tl + 1 - log2(2^tl / threshold) bits less one when
// the value is below max. Lane l takes symbols 4l..4l+3; the fields' offsets
// come from a wave scan of their widths and the bits are OR-ed into `words`
// (LDS, OZL_NCOUNT_WORDS, cleared here). Returns the header's bytes (written
// to dst), or 0 where the reference writer returns 0.
#define OZL_NCOUNT_WORDS 128
__device__ uint32_t wave_fse_ncount(PlanLds& L, uint8_t* dst, uint32_t maxSym, uint32_t tl, uint32_t* words)
{
    const uint32_t lane = threadIdx.x;
    const int ts = 1 << tl;
    for (uint32_t i = lane; i < OZL_NCOUNT_WORDS; i += 64) words[i] = 0;
    // the lane's counts, their weights (|count|), zero flags and the last
    // nonzero symbol among them
    int v[4];
    int aLoc = 0, lnz = -1;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t sy = 4 * lane + (uint32_t)k;
        v[k] = sy <= maxSym ? L.norm[sy] : 0;
        aLoc += v[k] < 0 ? -v[k] : v[k];
        if (sy <= maxSym && v[k] != 0) lnz = (int)sy;
    }
    // exclusive scans: weights before the lane, last nonzero symbol before it
    int aInc = aLoc, zInc = lnz;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int ya = __shfl_up(aInc, d, 64), yz = __shfl_up(zInc, d, 64);
        if ((int)lane >= d) {
            aInc += ya;
            zInc = yz > zInc ? yz : zInc;
        }
    }
    int A = aInc - aLoc;                          // weights before symbol 4l
    int lastNz = __shfl_up(zInc, 1, 64);          // last nonzero symbol before 4l
    if (lane == 0) lastNz = -1;
    // previous symbol's count (symbol 4l - 1)
    int prevV = __shfl_up(v[3], 1, 64);
    if (lane == 0) prevV = 1;                     // symbol 0: no zero before it
    // fields
    uint32_t fw[4], fv[4], rw[4], rr[4];
    uint32_t wLoc = 0;
    bool bad = false, reached = false;            // remaining went below 1 / reached 1
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t sy = 4 * lane + (uint32_t)k;
        const int R = ts + 1 - A;                 // remaining before sy
        const bool live = sy <= maxSym && R > 1;  // the writer's loop reaches sy
        const int pv = k == 0 ? prevV : v[k - 1];
        const bool cont = live && v[k] == 0 && sy > 0 && pv == 0;   // inside a zero run
        fw[k] = fv[k] = rw[k] = rr[k] = 0;
        if (live && !cont) {
            if (v[k] != 0 && sy > 0 && pv == 0) {
                // repeat field of the zeros since the run's first one
                const uint32_t r = sy - 2 - (uint32_t)lastNz;
                rw[k] = 16 * (r / 24) + 2 * ((r % 24) / 3) + 2;
                rr[k] = r % 3;
            }
            const int thr = R > ts ? ts : 1 << ozl_hb32((uint32_t)R);
            const int nb = ozl_hb32((uint32_t)thr) + 1;
            const int mx = 2 * thr - 1 - R;
            int c = v[k] + 1;
            if (c >= thr) c += mx;
            fv[k] = (uint32_t)c;
            fw[k] = (uint32_t)(nb - (c < mx ? 1 : 0));
            const int R2 = R - (v[k] < 0 ? -v[k] : v[k]);
            if (R2 < 1) bad = true;
            if (R2 == 1) reached = true;
        }
        if (sy <= maxSym && v[k] != 0) lastNz = (int)sy;
        A += v[k] < 0 ? -v[k] : v[k];
        wLoc += rw[k] + fw[k];
    }
    // the writer stops where remaining first drops to 1 or below: below 1,
    // or never reaching 1 within the alphabet, it returns 0
    const bool ok = __ballot(reached) != 0 && __ballot(bad) == 0;
    uint32_t wInc = wLoc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(wInc, d, 64);
        if ((int)lane >= d) wInc += y;
    }
    const uint32_t totalBits = 4 + __shfl(wInc, 63, 64);
    if (!ok || totalBits > 32 * OZL_NCOUNT_WORDS - 32) return 0;
    __syncthreads();
    auto put = [&](uint32_t P, uint32_t val, uint32_t w) {   // w <= 32, val < 2^w
        if (!w) return;
        const uint32_t sh = P & 31;
        atomicOr(&words[P >> 5], val << sh);
        if (sh + w > 32) atomicOr(&words[(P >> 5) + 1], val >> (32 - sh));
    };
    if (lane == 0) put(0, tl - OZL_FSE_MIN_TABLELOG, 4);
    uint32_t P = 4 + wInc - wLoc;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (rw[k]) {
            uint32_t ones = rw[k] - 2;
            while (ones) {
                const uint32_t c = ones < 32 ? ones : 32;
                put(P, c == 32 ? 0xFFFFFFFFu : (1u << c) - 1u, c);
                P += c;
                ones -= c;
            }
            put(P, rr[k], 2);
            P += 2;
        }
        put(P, fv[k], fw[k]);
        P += fw[k];
    }
    __syncthreads();
    const uint32_t bytes = (totalBits + 7) / 8;
    for (uint32_t i = lane; i < bytes; i += 64) dst[i] = (uint8_t)(words[i >> 2] >> (8 * (i & 3)));
    return bytes;
}

// Histogram of every block (one 256-thread workgroup per block, LDS
// atomics on 16 interleaved copies), parked in the block's table area.
__global__ void __launch_bounds__(kT) k_hist(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                             const uint64_t* __restrict__ offEnd, uint32_t B, ozl_ews ws)
{
    __shared__ __attribute__((aligned(16))) uint32_t hc[kHistCopies * 256];
    __shared__ uint32_t count[256];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    block_hist(src + off[b], (uint32_t)(offEnd[b] - off[b]), hc, count);
    uint32_t* cg = reinterpret_cast<uint32_t*>(ws.tables + (size_t)b * OZL_TAB_BYTES + OZL_TAB_COUNT_OFF);
    for (int s = threadIdx.x; s < 256; s += kT) cg[s] = count[s];
}

// Histogram and content checksum of every block in one read of the input
// (the batch API; the engine hashes whole inputs): one 256-thread workgroup
// per block. Each thread takes 16-byte pieces t, t + 256, ...: 16 LDS atomics
// for the histogram and, when the block starts 16-byte aligned, the
// piece's XXH3 terms, so that wave w covers 1 KiB blocks w, w + 4, ... and
// reduces each block's stripe sums (ozl_xxh3_sums_of) into LDS; wave 0 then
// runs the scramble chain over them and the tail stripes. Unaligned blocks
// hash with ozl_xxh3_wave after the histogram. The hash is parked in
// hashOut[b] (the status array, which the frame checksum kernel overwrites
// after the frame writer has read it).
__global__ void __launch_bounds__(kT) k_hist_hash(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                                  uint32_t B, ozl_ews ws, uint32_t* __restrict__ hashOut)
{
    __shared__ __attribute__((aligned(16))) uint32_t hc[kHistCopies * 256];
    __shared__ uint32_t count[256];
    __shared__ uint64_t bsum[100][8];   // blocks are below 100000 bytes: < 98 full KiB
    const uint32_t b = blockIdx.x;
    if (b >= B) return;
    const uint8_t* in = src + off[b];
    const uint32_t n = (uint32_t)(off[b + 1] - off[b]);
    const int tid = threadIdx.x, lane = tid & 63;
    const bool aligned = ((uintptr_t)in & 15) == 0;
    const uint32_t nbBlocks = n > 240 ? (n - 1) / 1024 : 0u;
    for (int i = tid; i < kHistCopies * 256; i += kT) hc[i] = 0;
    __syncthreads();
    uint32_t* h = hc + (tid & (kHistCopies - 1));
    if (aligned) {
        const ozl_xxh3_lane XL = ozl_xxh3_lane_init();
        const uint32_t nv = n / 16;
        // whole waves of pieces: a wave's 64 pieces are one 1 KiB block; four
        // pieces per thread are loaded before any is used
        constexpr int kU = 4;
        for (uint32_t c0 = (uint32_t)(tid - lane); c0 < nv; c0 += kU * kT) {
            uint4 q[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t c = c0 + u * kT + lane;
                q[u] = c < nv ? *reinterpret_cast<const uint4*>(in + 16 * (size_t)c) : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const uint32_t cu = c0 + u * kT;
                if (cu + lane < nv) {
                    const uint32_t w[4] = { q[u].x, q[u].y, q[u].z, q[u].w };
#pragma unroll
                    for (int k = 0; k < 4; k++)
#pragma unroll
                        for (int bb = 0; bb < 4; bb++) atomicAdd(&h[kHistCopies * ((w[k] >> (8 * bb)) & 255)], 1u);
                }
                const uint32_t blk = cu / 64;
                if (blk < nbBlocks) {
                    uint64_t sa, sb;
                    ozl_xxh3_sums_of(q[u].x | ((uint64_t)q[u].y << 32), q[u].z | ((uint64_t)q[u].w << 32), XL, sa, sb);
                    if (lane < 4) {
                        bsum[blk][2 * lane] = sa;
                        bsum[blk][2 * lane + 1] = sb;
                    }
                }
            }
        }
        for (uint32_t i = 16 * nv + tid; i < n; i += kT) atomicAdd(&h[kHistCopies * in[i]], 1u);
    } else {
        uint32_t head = (uint32_t)((16 - ((uintptr_t)in & 15)) & 15);
        if (head > n) head = n;
        for (uint32_t i = tid; i < head; i += kT) atomicAdd(&h[kHistCopies * in[i]], 1u);
        const uint32_t nv = (n - head) / 16;
        const uint4* v = reinterpret_cast<const uint4*>(in + head);
        for (uint32_t i = tid; i < nv; i += kT) {
            const uint4 q = v[i];
            const uint32_t w[4] = { q.x, q.y, q.z, q.w };
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int bb = 0; bb < 4; bb++) atomicAdd(&h[kHistCopies * ((w[k] >> (8 * bb)) & 255)], 1u);
        }
        for (uint32_t i = head + nv * 16 + tid; i < n; i += kT) atomicAdd(&h[kHistCopies * in[i]], 1u);
    }
    __syncthreads();
    for (int sy = tid; sy < 256; sy += kT) {
        const uint4* r = reinterpret_cast<const uint4*>(hc + kHistCopies * sy);
        uint32_t t = 0;
#pragma unroll
        for (int c = 0; c < kHistCopies / 4; c++) t += r[c].x + r[c].y + r[c].z + r[c].w;
        count[sy] = t;
    }
    __syncthreads();
    uint32_t* cg = reinterpret_cast<uint32_t*>(ws.tables + (size_t)b * OZL_TAB_BYTES + OZL_TAB_COUNT_OFF);
    for (int sy = tid; sy < 256; sy += kT) cg[sy] = count[sy];
    if (tid >= 64) return;
    uint64_t hv;
    if (n <= 240) {
        hv = ozl_xxh3_short(in, n);
    } else if (!aligned) {
        hv = ozl_xxh3_wave(in, n, nullptr, 0);
    } else {
        // the chain over the block sums: lane j keeps accumulators 2j, 2j+1
        const int j = lane & 3;
        const ozl_xxh3_lane XL = ozl_xxh3_lane_init();
        uint64_t A = ozl_xxh3_init_acc(2 * j), Bc = ozl_xxh3_init_acc(2 * j + 1);
        for (uint32_t k = 0; k < nbBlocks; k++) {
            A = ozl_xxh3_scr(A + bsum[k][2 * j], XL.c0);
            Bc = ozl_xxh3_scr(Bc + bsum[k][2 * j + 1], XL.c1);
        }
        uint64_t acc = 0;
        for (int qq = 0; qq < 4; qq++) {
            const uint64_t a = __shfl(A, qq, 64), bq = __shfl(Bc, qq, 64);
            if (lane == 2 * qq) acc = a;
            if (lane == 2 * qq + 1) acc = bq;
        }
        hv = ozl_xxh3_finish(in, n, nbBlocks, acc);
    }
    if (lane == 0) hashOut[b] = (uint32_t)hv;
}

// The placement of ozl_huf_place by the wave: the same node order (buckets
// of descending rank, symbol order within a bucket) and the same W.base /
// W.curr. W.tbl and W.base start at zero, and so do the 192 words of `hist`.
// Within a round of 64 symbols, a symbol's rank in its bucket is the number
// of lower lanes whose bucket matches all 8 bits of its own.
__device__ inline void wave_huf_place(const uint32_t* count, uint32_t maxSym, ozl_huf_scratch& W, uint32_t* hist)
{
    const uint32_t lane = threadIdx.x;
    for (uint32_t s = lane; s <= maxSym; s += 64) atomicAdd(&hist[ozl_huf_bucket(count[s])], 1u);
    __syncthreads();
    {
        // base[r] = symbols of bucket >= r; lane i holds buckets 3i .. 3i+2
        const uint32_t h0 = hist[3 * lane], h1 = hist[3 * lane + 1], h2 = hist[3 * lane + 2];
        uint32_t t = h0 + h1 + h2;
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_down(t, d, 64);
            if (lane + d < 64) t += o;
        }
        const uint16_t b0 = (uint16_t)t, b1 = (uint16_t)(t - h0), b2 = (uint16_t)(t - h0 - h1);
        W.base[3 * lane] = b0;
        W.base[3 * lane + 1] = b1;
        W.base[3 * lane + 2] = b2;
        W.curr[3 * lane] = b0;
        W.curr[3 * lane + 1] = b1;
        W.curr[3 * lane + 2] = b2;
    }
    __syncthreads();
    ozl_hnode* node = W.tbl + 1;
    const uint64_t below = (1ull << lane) - 1ull;
    for (uint32_t s0 = 0; s0 <= maxSym; s0 += 64) {
        const uint32_t s = s0 + lane;
        const bool act = s <= maxSym;
        const uint32_t c = act ? count[s] : 0u;
        const uint32_t r = ozl_huf_bucket(c) + 1;
        const uint64_t m = ozl_wave::match8(r, act);
        if (act) {
            const uint32_t pos = W.curr[r] + (uint32_t)__popcll(m & below);
            node[pos].count = c;
            node[pos].byte = (uint8_t)s;
        }
        __syncthreads();
        if (act && (m >> lane) == 1ull) W.curr[r] = (uint16_t)(W.curr[r] + __popcll(m));
        __syncthreads();
    }
}

// The depths of ozl_huf_depths by the wave: pointer jumping over the
// internal nodes (word j of `P` = depth << 16 | the internal ancestor it
// has reached, for node 256 + j; the root points at itself with depth 0),
// then every leaf one below its parent. 256 words of `P`.
__device__ inline void wave_huf_depths(ozl_huf_scratch& W, uint32_t* P)
{
    ozl_hnode* node = W.tbl + 1;
    const int nonNull = W.nonNull;
    const uint32_t root = (uint32_t)nonNull - 1u, lane = threadIdx.x;
    uint32_t v[4];
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t j = lane + 64 * i;
        v[i] = j < root ? (1u << 16) | (uint32_t)(node[256 + j].parent - 256) : root;
    }
    // depths of internal nodes are below 255: at most 8 jumps
    for (;;) {
        bool more = false;
        for (uint32_t i = 0; i < 4; i++) {
            P[lane + 64 * i] = v[i];
            more |= (v[i] & 0xFFFFu) != root;
        }
        __syncthreads();
        if (!__any(more)) break;
        for (uint32_t i = 0; i < 4; i++) {
            const uint32_t w = P[v[i] & 0xFFFFu];
            v[i] = ((v[i] & 0xFFFF0000u) + (w & 0xFFFF0000u)) | (w & 0xFFFFu);
        }
        __syncthreads();
    }
    for (uint32_t i = 0; i < 4; i++) {
        const uint32_t j = lane + 64 * i;
        if (j <= root) node[256 + j].nbBits = (uint8_t)(v[i] >> 16);
        if ((int)j <= nonNull) node[j].nbBits = (uint8_t)((P[node[j].parent - 256] >> 16) + 1);
    }
    __syncthreads();
}

// One wave per block: decision / normalisation / Huffman code / nested
// weights graph (lane 0, ozl_core.h), FSE encoding table (the wave,
// ozl_wave::wave_fse_ctable, straight into the block's table area). Small LDS
// (12 KB) lets a CU run a dozen of these serial builds at once.
template <int HUF>
__global__ void __launch_bounds__(64) k_plan(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                             const uint64_t* __restrict__ offEnd, uint32_t B, ozl_ews ws)
{
    __shared__ PlanLds L;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    uint32_t n = (uint32_t)(offEnd[b] - off[b]);
    uint8_t* tab = ws.tables + (size_t)b * OZL_TAB_BYTES;
    {
        const uint32_t* cg = reinterpret_cast<const uint32_t*>(tab + OZL_TAB_COUNT_OFF);
        for (int s = threadIdx.x; s < 256; s += 64) L.count[s] = cg[s];
    }
    __syncthreads();
    // the histogram's statistics by the wave: four symbols per lane, exact
    // integer reductions (the entropy terms' sum does not depend on order)
    int64_t tsum = 0;
    uint32_t ms = 0, cd = 0;
    {
        const int64_t norm = n ? ((int64_t)1 << 62) / (int64_t)n : 0;
        for (int k = 0; k < 4; k++) {
            const uint32_t s = threadIdx.x + 64 * k, c = L.count[s];
            if (c) {
                tsum += ozl_entropy_term(c, norm);
                ms = s;
                cd++;
            }
        }
        for (int d = 32; d >= 1; d >>= 1) {
            tsum += __shfl_xor(tsum, d, 64);
            cd += __shfl_xor(cd, d, 64);
            const uint32_t o = __shfl_xor(ms, d, 64);
            ms = o > ms ? o : ms;
        }
    }
    if (HUF) {
        // the Huffman build's node table and buckets and the weights'
        // histogram start at zero: cleared by the wave
        uint32_t* tb = reinterpret_cast<uint32_t*>(L.hs.tbl);
        for (int i = threadIdx.x; i < 2 * (2 * 256 + 2); i += 64) tb[i] = 0;
        for (int i = threadIdx.x; i < 192; i += 64) L.hs.base[i] = 0;
        for (int i = threadIdx.x; i < 256; i += 64) L.wcount[i] = 0;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // the plan is built in place in LDS (a local copy handed to
        // plan_weights by reference lived in scratch memory)
        ozl_eplan& p = L.plan;
        p = ozl_eplan{};
        p.n = n;
        uint8_t* aux = L.aux;
        const unsigned maxSym = ms, card = cd;
        int bk = ozl_decide_from_stats(tsum, n, maxSym, card, HUF);
        if (ozl_small(n, ws.minSS)) bk = OZL_BK_STORE;
        p.bk = (uint8_t)bk;
        p.maxSym = (uint8_t)maxSym;
        p.card = (uint16_t)card;
        switch (bk) {
        case OZL_BK_CONSTANT: aux[0] = (uint8_t)maxSym; p.aux = (uint16_t)maxSym; break;
        case OZL_BK_FLATPACK: {
            unsigned c = 0;
            for (int s = 0; s < 256; s++)
                if (L.count[s]) aux[c++] = (uint8_t)s;
            p.aux = (uint16_t)c;
            unsigned nb = c <= 1 ? c : 1 + (unsigned)ozl_hb32(c - 1);
            p.payload = 1 + nb * n / 8;
            break;
        }
        case OZL_BK_BITPACK: {
            int nb = 1 + (maxSym ? ozl_hb32(maxSym) : 0);
            p.aux = (uint16_t)nb;
            p.payload = (uint32_t)(((size_t)n * nb + 7) / 8);
            break;
        }
        case OZL_BK_FSE: {
            unsigned tl = ozl_fse_optimal_tablelog(OZL_FSE_DEFAULT_TABLELOG, n, maxSym, 2);
            p.tl = (uint8_t)tl;
            p.nbStates = n < 1000 ? 2 : 4;
            if (tl > 11) p.status = OZL_ERR_GENERIC;
            break;   // normalisation and NCount by the wave, below
        }
        case OZL_BK_HUF:
            // the code length limit; the code itself follows
            p.tl = (uint8_t)ozl_fse_optimal_tablelog(OZL_HUF_DEFAULT_TABLELOG, n, maxSym, 1);
            p.nbStates = n > 1000 ? 1 : 0;
            break;
        default: break;
        }
    }
    __syncthreads();
    if (!HUF && L.plan.bk == OZL_BK_FSE && L.plan.status == 0) {
        const uint32_t maxSym = L.plan.maxSym, tl = L.plan.tl;
        const int rc = wave_fse_normalize(L, n, maxSym, tl);
        if (rc != 0) {
            if (threadIdx.x == 0) L.plan.status = OZL_ERR_GENERIC;
        } else if (ozl_small(2u * (maxSym + 1u), ws.minSS)) {
            // raw int16 counts (stored after a num->serial conversion)
            for (uint32_t s = threadIdx.x; s <= maxSym; s += 64) {
                L.aux[2 * s] = (uint8_t)L.norm[s];
                L.aux[2 * s + 1] = (uint8_t)((uint16_t)L.norm[s] >> 8);
            }
            if (threadIdx.x == 0) L.plan.aux = (uint16_t)(2 * (maxSym + 1));
        } else {
            // the words live in the table build's area, which is free until
            // wave_fse_ctable below
            static_assert(OZL_NCOUNT_WORDS * 4 <= sizeof(ozl_wave::FseBuild), "ncount words");
            const uint32_t bytes = wave_fse_ncount(L, L.aux, maxSym, tl, reinterpret_cast<uint32_t*>(&L.fb));
            if (threadIdx.x == 0) L.plan.aux = (uint16_t)bytes;
        }
        __syncthreads();
    }
    if (HUF && L.plan.bk == OZL_BK_HUF) {
        // code lengths of the sorted leaves (huf_compress.c:731-764): the
        // sort's placement and the depths by the wave, the sort of the large
        // buckets, the tree and the height limit by lane 0
        wave_huf_place(L.count, L.plan.maxSym, L.hs, L.wcount);
        if (threadIdx.x == 0) ozl_huf_sort_tree(L.plan.maxSym, &L.hs);
        __syncthreads();
        wave_huf_depths(L.hs, L.wcount);
        for (int i = threadIdx.x; i < 256; i += 64) L.wcount[i] = 0;
        if (threadIdx.x == 0) {
            const int mx = ozl_huf_limit(L.plan.tl, &L.hs);
            if (mx < 0) L.plan.status = OZL_ERR_GENERIC;
            else L.plan.tl = (uint8_t)mx;
        }
        __syncthreads();
    }
    if (HUF && L.plan.bk == OZL_BK_HUF && L.plan.status == 0) {
        // canonical codes (huf_compress.c:706-729) by the wave: leaves per
        // length, the first value of each length, every symbol's length, and
        // its value = the first value of its length + its rank among the
        // symbols of that length in symbol order
        const ozl_hnode* leaf = L.hs.tbl + 1;
        const int nonNull = L.hs.nonNull;
        const uint32_t largest = L.plan.tl, m = L.plan.maxSym + 1u;
        const uint32_t lane = threadIdx.x;
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t perRank[13];
        for (uint32_t r = 0; r <= 12; r++) perRank[r] = 0;
        for (int k0 = 0; k0 <= nonNull; k0 += 64) {
            const int k = k0 + (int)lane;
            const uint32_t len = k <= nonNull ? leaf[k].nbBits : 0u;
            for (uint32_t r = 1; r <= largest; r++) perRank[r] += (uint32_t)__popcll(__ballot(len == r));
        }
        uint32_t first[13];
        {
            uint16_t mn = 0;
            for (uint32_t r = 0; r <= 12; r++) first[r] = 0;
            for (int r = (int)largest; r > 0; r--) {
                first[r] = mn;
                mn = (uint16_t)(mn + (uint16_t)perRank[r]);
                mn >>= 1;
            }
        }
        for (uint32_t k = lane; k < m; k += 64) L.hlen[leaf[k].byte] = leaf[k].nbBits;
        __syncthreads();
        uint32_t runs[13];
        for (uint32_t r = 0; r <= 12; r++) runs[r] = 0;
        for (uint32_t s0 = 0; s0 < m; s0 += 64) {
            const uint32_t sy = s0 + lane;
            const uint32_t len = sy < m ? L.hlen[sy] : 0u;
            uint32_t v = 0;
            for (uint32_t r = 1; r <= largest; r++) {
                const uint64_t mk = __ballot(len == r);
                if (len == r) v = (uint16_t)(first[r] + runs[r] + (uint32_t)__popcll(mk & below));
                runs[r] += (uint32_t)__popcll(mk);
            }
            if (sy < m) {
                L.hval[sy] = (uint16_t)v;
                const uint8_t w = len ? (uint8_t)(largest + 1 - len) : 0;
                L.wts[sy] = w;
                L.aux[sy] = w;
            }
        }
        __syncthreads();
        // histogram of the weights (values 0..12: code lengths are at most
        // 11) by ballots of the wave
        uint32_t seen = 0;
        for (uint32_t v = 0; v <= 12; v++) {
            uint32_t c = 0;
            for (uint32_t i0 = 0; i0 < m; i0 += 64) {
                const uint32_t i = i0 + threadIdx.x;
                c += (uint32_t)__popcll(__ballot(i < m && L.wts[i] == v));
            }
            seen += c;
            if (threadIdx.x == 0) L.wcount[v] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            L.wWave = 0;
            if (seen != m) L.plan.status = OZL_ERR_GENERIC;
            else plan_weights(L.plan, L, L.aux, ws.minSS);
        }
        __syncthreads();
        if (L.wWave && L.plan.status == 0) {
            wave_weights_fse(L, L.aux + OZL_AUX_W_PAYLOAD);
            __syncthreads();
        }
    }
    const ozl_eplan& p = L.plan;
    if (p.bk == OZL_BK_FSE && p.status == 0) {
        ozl_wave::wave_fse_ctable(reinterpret_cast<uint16_t*>(tab), reinterpret_cast<uint32_t*>(tab + 8192), L.norm,
                                  p.maxSym, p.tl, L.fb);
    } else if (p.bk == OZL_BK_HUF && p.status == 0) {
        uint16_t* val = reinterpret_cast<uint16_t*>(tab + 256);
        for (unsigned s = threadIdx.x; s < 256; s += 64) {
            tab[s] = s <= p.maxSym ? L.hlen[s] : 0;
            val[s] = s <= p.maxSym ? L.hval[s] : 0;
        }
    }
    uint8_t* aux = ws.aux + (size_t)b * OZL_AUX_BYTES;
    for (unsigned i = threadIdx.x; i < OZL_AUX_BYTES / 4; i += 64)
        reinterpret_cast<uint32_t*>(aux)[i] = reinterpret_cast<const uint32_t*>(L.aux)[i];
    if (threadIdx.x == 0) ws.plans[b] = p;
}

// ------------------------------------------------------------------ FSE encode
// 8 blocks per one-wave workgroup, a quad of lanes per block; lane q owns
// encoder state q (fse_compress.c:607-660). Emission e (e = 0.. n-S-1)
// encodes position p = n-S-1-e with state (e - n) mod S; a round gives each
// lane one emission, in slot (q + n) mod S of the round. The quad turns its
// bit counts into offsets in slot order (each lane's count copied into bytes
// slot..3 of a word, the words summed with two DPP exchanges, so that byte k
// holds the bits of slots 0..k), ORs the shifted
// values into one 64-bit round word, and appends that to a 32-bit word writer
// whose completed words lanes 0 and 1 store. Each lane's input symbols are
// loaded 16 rounds ahead of use (an unrolled register queue), so the global
// load latency stays off the state chain. Tables take 5 KB per block (40 KB
// per workgroup, 4 workgroups per CU).
//
// In LDS a symbol's transform is repacked so that a round's state update is
// four instructions: the high half holds c = (maxBitsOut - 1) << (tl + 1) +
// 2T - minStatePlus, so that ((state << 16) + tt) >> (tl + 17) is
// maxBitsOut - (state < minStatePlus), the number of bits out (state lies in
// [T, 2T) and minStatePlus in [T, 2T]); the low half holds the LDS byte offset
// of the block's state table entry deltaFindState (positive: the tables follow
// the 8 KB of transforms and deltaFindState >= -T), to which 2 * (state >>
// nbBitsOut) is added.
constexpr int kFseGroups = 8;
constexpr int kFsePipe = 16;

struct FseEncLds {
    uint32_t tt[kFseGroups][256];
    uint16_t st[kFseGroups][2048];
};

typedef __attribute__((address_space(1))) const uint8_t gc_uint8;
typedef __attribute__((address_space(1))) uint32_t g_uint32;

__device__ inline uint32_t sel4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t i)
{
    return i == 0 ? a0 : i == 1 ? a1 : i == 2 ? a2 : a3;
}

// the LDS form of a symbol transform (see above) for block slot f
__device__ inline uint32_t tt_lds(uint32_t p, uint32_t tl, uint32_t f)
{
    const uint32_t T = 1u << tl;
    const uint32_t c = (((p & 15u) - 1u) << (tl + 1)) + 2 * T - ((p >> 4) & 0x3FFFu);
    const uint32_t a = (uint32_t)offsetof(FseEncLds, st) + 4096u * f + 2u * ((p >> 18) - 4096u);
    return (c << 16) | (a & 0xFFFFu);
}

__global__ void __launch_bounds__(64) k_fse_encode(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                                   uint32_t B, ozl_ews ws)
{
    __shared__ FseEncLds L;
    const int lane = threadIdx.x & 63, g = lane >> 2, q = lane & 3;
    const uint32_t b = blockIdx.x * kFseGroups + g;
    bool active = g < kFseGroups && b < B;
    ozl_eplan p = {};
    if (active) {
        p = ws.plans[b];
        active = p.bk == OZL_BK_FSE && p.status == 0 && p.tl <= 11 && p.n > p.nbStates;
    }
    // the wave copies each block's state table and symbol transforms to LDS
    for (int f = 0; f < kFseGroups; f++) {
        const int actf = __shfl((int)active, 4 * f, 64);
        if (!actf) continue;
        const uint32_t tlf = (uint32_t)__shfl((int)p.tl, 4 * f, 64), msf = (uint32_t)__shfl((int)p.maxSym, 4 * f, 64);
        const uint4* tab = reinterpret_cast<const uint4*>(ws.tables + (size_t)(blockIdx.x * kFseGroups + f) * OZL_TAB_BYTES);
        const uint32_t nS = (1u << tlf) / 8, nT = (msf + 4) / 4;
        for (uint32_t i = lane; i < nS; i += 64) reinterpret_cast<uint4*>(L.st[f])[i] = tab[i];
        for (uint32_t i = lane; i < nT; i += 64) {
            const uint4 x = tab[512 + i];
            reinterpret_cast<uint4*>(L.tt[f])[i] = make_uint4(tt_lds(x.x, tlf, f), tt_lds(x.y, tlf, f),
                                                             tt_lds(x.z, tlf, f), tt_lds(x.w, tlf, f));
        }
    }
    ozl_wave::wave_sync();
    if (!active) return;
    gc_uint8* in = (gc_uint8*)(src + off[b]);
    const uint32_t n = p.n, S = p.nbStates, tl = p.tl;
    const uint32_t* TT = L.tt[g];
    // payload words of the workgroup's blocks from one base (32-bit offsets):
    // this block's words start at gw; word capWords + 1 is a sink the payload
    // never reaches (its tail ends in word capWords at the latest)
    g_uint32* outw = (g_uint32*)(ws.payload + (size_t)blockIdx.x * kFseGroups * ws.payloadCap);
    const uint32_t gw = (uint32_t)g * (uint32_t)(ws.payloadCap / 4);
    const uint32_t capWords = (uint32_t)(ws.payloadCap / 4) - 2;
    const bool mine = (uint32_t)q < S;
    // init: the one position among the last S whose state index is q
    // (FSE_initCState2), from the block's transforms in their packed form
    uint32_t state = 0;
    if (mine) {
        uint32_t pinit = 0;
        for (uint32_t k = 0; k < S; k++) {
            uint32_t pp = n - 1 - k;
            if ((S - 1 - (pp % S)) % S == (uint32_t)q) pinit = pp;
        }
        const uint32_t* ttg = reinterpret_cast<const uint32_t*>(ws.tables + (size_t)b * OZL_TAB_BYTES + 8192);
        uint32_t t = ttg[in[pinit]];
        uint32_t dnb = ozl_tt_dnb(t);
        uint32_t nbOut = (dnb + (1u << 15)) >> 16;
        uint32_t v = (nbOut << 16) - dnb;
        state = L.st[g][(int)(v >> nbOut) + ozl_tt_dfs(t)];
    }
    const uint32_t nEmit = n - S;
    const uint32_t kq = mine ? (uint32_t)(q + n) % S : (uint32_t)q;  // this lane's slot in a round
    const uint32_t sq = 8 * kq;
    // v_perm selector of the bit count in bytes kq..3 (0: byte 0, 0x0c: zero)
    const uint32_t selq = 0x0c0c0c0cu & ~(0xFFFFFFFFu << sq);
    const uint32_t rounds = (nEmit + S - 1) / S;
    uint32_t qbuf[kFsePipe];
#pragma unroll
    for (int k = 0; k < kFsePipe; k++) {
        uint32_t e = (uint32_t)k * S + kq;
        qbuf[k] = in[(mine && e < nEmit) ? nEmit - 1 - e : 0u];
    }
    uint32_t tn = TT[qbuf[0]];          // transform of the next round's symbol
    uint32_t a0 = 0, ab = 0, wpos = 0;  // pending bits (ab < 32) and words written
    // one emission per lane: the state update, then the quad appends the
    // round's bits in slot order. FAST rounds store every lane's word at
    // wpos + q, complete or not (a word keeps being rewritten until it is
    // complete, and a batch of them checks the capacity once); the others
    // store completed words only, the rest to the sink.
    auto emit = [&](uint32_t t, bool doit, auto fastTag) {
        constexpr bool FAST = decltype(fastTag)::value;
        const uint32_t nbx = ((state << 16) + t) >> (tl + 17);
        const uint32_t nb = FAST || doit ? nbx : 0u;
        const uint32_t val = __builtin_amdgcn_ubfe(state, 0, nb);
        // (a lane without an emission reads the first table entry instead)
        const uint32_t at = FAST || doit ? ((state >> nb) << 1) + (t & 0xFFFFu) : (uint32_t)offsetof(FseEncLds, st);
        const uint32_t ns = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const uint8_t*>(&L) + at);
        state = FAST || doit ? ns : state;
        // nb in bytes kq..3 (a byte broadcast by v_perm), summed over the quad
        uint32_t v = __builtin_amdgcn_perm(0u, nb, selq);
        v += ozl_wave::quad_xor1(v);
        const uint32_t M = v + ozl_wave::quad_xor2(v);  // byte k: bits of slots 0..k
        const uint32_t excl = __builtin_amdgcn_ubfe(M, sq, 8) - nb;
        const uint64_t R = (uint64_t)val << excl;
        uint32_t rl = (uint32_t)R, rh = (uint32_t)(R >> 32);
        rl |= ozl_wave::quad_xor1(rl);
        rh |= ozl_wave::quad_xor1(rh);
        rl |= ozl_wave::quad_xor2(rl);
        rh |= ozl_wave::quad_xor2(rh);
        const uint64_t T = (((uint64_t)rh << 32) | rl) << ab;
        const uint32_t w0 = a0 | (uint32_t)T, w1 = (uint32_t)(T >> 32), w2 = __builtin_amdgcn_ubfe(rh, 32 - ab, ab);
        const uint32_t nAcc = ab + (M >> 24), ne = nAcc >> 5;
        // (a 32-bit byte offset from the workgroup's base: one address VGPR)
        const uint32_t wi = FAST ? wpos + (uint32_t)q : (uint32_t)q < ne && wpos + q < capWords ? wpos + q : capWords + 1;
        *(g_uint32*)((__attribute__((address_space(1))) uint8_t*)outw + 4u * (gw + wi)) = q == 0 ? w0 : w1;
        a0 = ne == 0 ? w0 : ne == 1 ? w1 : w2;
        ab = nAcc & 31;
        wpos += ne;
    };
    uint32_t r0 = 0;
    if (S == 4) {
        // rounds whose 4 emissions all exist, while the queue's next 16 do too:
        // no per-lane conditions, queue loads at immediate offsets
        // (at most 2 words complete per round: a batch needs 2 * kFsePipe + 4
        // words of room for its stores)
        const uint32_t RF = nEmit / 4;
        for (; r0 + 2 * kFsePipe <= RF && wpos + 2 * kFsePipe + 4 <= capWords; r0 += kFsePipe) {
            gc_uint8* pb = in + (nEmit - 1 - 4 * (r0 + kFsePipe) - kq);
#pragma unroll
            for (int k = 0; k < kFsePipe; k++) {
                const uint32_t t = tn;
                tn = TT[qbuf[(k + 1) % kFsePipe]];
                qbuf[k] = pb[-4 * k];
                emit(t, true, std::true_type{});
            }
        }
    }
    for (; r0 < rounds; r0 += kFsePipe) {
#pragma unroll
        for (int k = 0; k < kFsePipe; k++) {
            const uint32_t e = (r0 + (uint32_t)k) * S + kq;
            const uint32_t t = tn;
            tn = TT[qbuf[(k + 1) % kFsePipe]];
            const uint32_t eNext = e + kFsePipe * S;
            qbuf[k] = in[(mine && eNext < nEmit) ? nEmit - 1 - eNext : 0u];
            emit(t, mine && e < nEmit, std::false_type{});
        }
    }
    // the states in index order, then the end mark
    const uint32_t f0 = ozl_wave::quad_lane<0>(state), f1 = ozl_wave::quad_lane<1>(state);
    const uint32_t f2 = ozl_wave::quad_lane<2>(state), f3 = ozl_wave::quad_lane<3>(state);
    if (q != 0) return;
    uint64_t acc = a0;
    uint32_t accBits = ab;
    for (uint32_t i = 0; i < S; i++) {
        uint32_t sv = sel4(f0, f1, f2, f3, i) & ((1u << tl) - 1);
        acc |= (uint64_t)sv << accBits;
        accBits += tl;
        if (accBits >= 32) {
            if (wpos < capWords) outw[gw + wpos] = (uint32_t)acc;
            wpos++;
            acc >>= 32;
            accBits -= 32;
        }
    }
    acc |= 1ull << accBits;
    accBits += 1;
    uint32_t tailBytes = (accBits + 7) / 8;
    if (wpos >= capWords) {
        ws.plans[b].status = OZL_ERR_INTERNAL_BUFFER;
    } else {
        uint8_t* ob = reinterpret_cast<uint8_t*>(ws.payload + (size_t)b * ws.payloadCap + 4 * (size_t)wpos);
        for (uint32_t k = 0; k < tailBytes; k++) ob[k] = (uint8_t)(acc >> (8 * k));
    }
    ws.plans[b].payload = wpos * 4 + tailBytes;
}

// ------------------------------------------------------------------ Huffman encode
// One workgroup per block (huf_compress.c:1139-1196 for 4 streams, :1022-1086
// per stream). Per stream the symbols are coded from the last to the first:
// lane t takes the t-th run of `per` symbols from the end (per a multiple of
// 4, read as dwords), sums their code lengths, a block scan gives its
// starting bit, then it packs its codes in a 64-bit register and writes each
// completed word to an LDS bit buffer (an atomic OR only for its first and
// last words, which it shares with its neighbours); the buffer leaves in
// dwords.
constexpr int kHufMaxSeg = 25008;  // (100000+3)/4 symbols per stream
// dwords a thread of k_huf_encode holds: a stream of <= kHufMaxSeg symbols
// over 256 threads gives each at most 100 symbols
constexpr int kHufRegDwords = 25;
static_assert(((kHufMaxSeg + 255) / 256 + 3) / 4 <= kHufRegDwords, "register dwords per thread");

template <typename T>
__device__ T block_scan_excl(T v, T* lds, T& total)
{
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    for (int d = 1; d < 64; d <<= 1) {
        T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T off = 0, all = 0;
    for (int w = 0; w < kT / 64; w++) {
        T s = lds[w];
        if (w < wave) off += s;
        all += s;
    }
    __syncthreads();
    total = all;
    return off + x - v;
}

__global__ void __launch_bounds__(kT) k_huf_encode(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                                   uint32_t B, ozl_ews ws)
{
    __shared__ uint32_t sCode[256];  // code value | length << 16
    constexpr uint32_t kBitWords = (kHufMaxSeg * 11) / 32 + 64;
    __shared__ uint32_t sBits[kBitWords + kT];   // the stream's words, then one sink word per thread
    __shared__ uint32_t sScan[4];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_eplan p = ws.plans[b];
    if (p.bk != OZL_BK_HUF || p.status != 0) return;
    const uint8_t* tab = ws.tables + (size_t)b * OZL_TAB_BYTES;
    for (int s = threadIdx.x; s < 256; s += kT)
        sCode[s] = s <= p.maxSym ? (uint32_t)reinterpret_cast<const uint16_t*>(tab + 256)[s] | ((uint32_t)tab[s] << 16) : 0u;
    const uint8_t* in = src + off[b];
    uint8_t* out = ws.payload + (size_t)b * ws.payloadCap;
    const uint32_t n = p.n;
    const int x4 = p.nbStates;
    const uint32_t seg = x4 ? (n + 3) / 4 : n;
    const int nbSeg = x4 ? 4 : 1;
    uint32_t opos = x4 ? 6 : 0;
    // this lane's symbols of segment k: positions [lo, hi) of the stream,
    // coded downward; its dwords (positions hi-4(i+1), coded from the top)
    // stay in registers between the two passes, and the < 4 bytes below them
    // ([lo, low)) are read again
    struct SegGeom {
        uint32_t len, hi, lo, nq, low;
        const uint8_t* base;
    };
    auto geom = [&](int k) {
        SegGeom g;
        const uint32_t s0 = k * seg;
        g.len = (k < nbSeg - 1) ? seg : n - s0;
        const uint32_t per = ((g.len + kT - 1) / kT + 3) & ~3u;
        const uint32_t hiR = threadIdx.x * per, loR = hiR + per;
        g.hi = hiR < g.len ? g.len - hiR : 0u;
        g.lo = loR < g.len ? g.len - loR : 0u;
        g.base = in + s0;
        g.nq = (g.hi - g.lo) / 4;
        g.low = g.lo + (g.hi - g.lo) % 4;
        return g;
    };
    auto load = [&](const SegGeom& g, uint32_t* d) {
#pragma unroll
        for (int i = 0; i < kHufRegDwords; i++)
            d[i] = (uint32_t)i < g.nq ? *reinterpret_cast<const uint32_t*>(g.base + g.hi - 4 * (i + 1)) : 0u;
    };
    // the next segment's dwords are loaded while this one is packed and
    // stored
    uint32_t dq[kHufRegDwords], dn[kHufRegDwords];
    load(geom(0), dq);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (k >= nbSeg) break;
        const SegGeom g = geom(k);
        const uint32_t len = g.len, hi = g.hi, lo = g.lo, nq = g.nq, low = g.low;
        const uint8_t* base = g.base;
        (void)hi;
        uint32_t mine = 0;
#pragma unroll
        for (int i = 0; i < kHufRegDwords; i++) {
            const uint32_t d = dq[i];
            if ((uint32_t)i < nq)
                mine += (sCode[d >> 24] >> 16) + (sCode[(d >> 16) & 255] >> 16) + (sCode[(d >> 8) & 255] >> 16) +
                        (sCode[d & 255] >> 16);
        }
        for (uint32_t q = low; q > lo; q--) mine += sCode[base[q - 1]] >> 16;
        if (k + 1 < nbSeg) load(geom(k + 1), dn);
        uint32_t total;
        const uint32_t bit = block_scan_excl<uint32_t>(mine, sScan, total);
        const uint32_t words = (total + 1 + 31) / 32;
        for (uint32_t w = threadIdx.x; w < words; w += kT) sBits[w] = 0;
        __syncthreads();
        if (mine) {
            // the thread's first word is shared with the thread before (kept
            // in `first`, OR-ed at the end), its last with the one after;
            // the words between are its own. Every step stores: a completed
            // word to its place, else to the thread's sink word (no branch).
            const uint32_t w0 = bit >> 5, sink = kBitWords + threadIdx.x;
            uint64_t acc = 0;
            uint32_t nacc = bit & 31, wpos = w0, first = 0;
            auto flush = [&]() {   // nacc < 54 on entry, < 32 after
                const bool full = nacc >= 32;
                const uint32_t lo32 = (uint32_t)acc;
                sBits[full && wpos != w0 ? wpos : sink] = lo32;
                first = full && wpos == w0 ? lo32 : first;
                wpos += full ? 1u : 0u;
                acc = full ? acc >> 32 : acc;
                nacc -= full ? 32u : 0u;
            };
            auto put2 = [&](uint32_t c0, uint32_t c1) {   // two codes of <= 11 bits
                acc |= (uint64_t)(c0 & 0xFFFFu) << nacc;
                nacc += c0 >> 16;
                acc |= (uint64_t)(c1 & 0xFFFFu) << nacc;
                nacc += c1 >> 16;
                flush();
            };
#pragma unroll
            for (int i = 0; i < kHufRegDwords; i++) {
                if ((uint32_t)i < nq) {
                    const uint32_t d = dq[i];
                    put2(sCode[d >> 24], sCode[(d >> 16) & 255]);
                    put2(sCode[(d >> 8) & 255], sCode[d & 255]);
                }
            }
            for (uint32_t q = low; q > lo; q--) put2(sCode[base[q - 1]], 0u);
            if (wpos != w0) atomicOr(&sBits[w0], first);
            if (nacc) atomicOr(&sBits[wpos], (uint32_t)acc);
        }
        if (threadIdx.x == 0) atomicOr(&sBits[total >> 5], 1u << (total & 31));  // end mark
        __syncthreads();
        const uint32_t bytes = (total + 1 + 7) / 8;
        if (opos + bytes > ws.payloadCap) {  // cannot happen for codes of <= 11 bits chosen by the estimate
            if (threadIdx.x == 0) ws.plans[b].status = OZL_ERR_INTERNAL_BUFFER;
            return;
        }
        // out + opos need not be 4-aligned: lane i writes the i-th dword of
        // the aligned span, assembled from the two buffer words it covers
        uint8_t* dst = out + opos;
        const uint32_t mis = (uint32_t)((uintptr_t)dst & 3);
        uint32_t* dw = reinterpret_cast<uint32_t*>(dst - mis);
        const uint32_t nw = (mis + bytes + 3) / 4;
        for (uint32_t i = threadIdx.x; i < nw; i += kT) {
            // dword i covers stream bytes [4i - mis, 4i - mis + 4)
            const int sb = 4 * (int)i - (int)mis;
            uint32_t v;
            if (mis == 0) {
                v = sBits[i];
            } else {
                const uint32_t lo32 = i > 0 ? sBits[i - 1] : 0u, hi32 = sBits[i];
                v = __builtin_amdgcn_alignbit(hi32, lo32, 32 - 8 * mis);
            }
            if (sb >= 0 && sb + 4 <= (int)bytes) {
                dw[i] = v;
            } else {
                for (int q = 0; q < 4; q++)
                    if (sb + q >= 0 && sb + q < (int)bytes) dst[sb + q] = (uint8_t)(v >> (8 * q));
            }
        }
        if (threadIdx.x == 0 && x4 && k < 3) {
            out[2 * k] = (uint8_t)bytes;
            out[2 * k + 1] = (uint8_t)(bytes >> 8);
        }
        opos += bytes;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kHufRegDwords; i++) dq[i] = dn[i];
    }
    if (threadIdx.x == 0) ws.plans[b].payload = opos;
}

// ------------------------------------------------------------------ flatpack / bitpack payloads
__global__ void __launch_bounds__(kT) k_pack_encode(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                                    uint32_t B, ozl_ews ws)
{
    __shared__ uint8_t sMap[256];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_eplan p = ws.plans[b];
    if ((p.bk != OZL_BK_FLATPACK && p.bk != OZL_BK_BITPACK) || p.status != 0) return;
    const uint8_t* in = src + off[b];
    uint8_t* out = ws.payload + (size_t)b * ws.payloadCap;
    const uint8_t* aux = ws.aux + (size_t)b * OZL_AUX_BYTES;
    uint32_t n = p.n, nb;
    bool flat = p.bk == OZL_BK_FLATPACK;
    if (flat) {
        for (int s = threadIdx.x; s < 256; s += kT) sMap[s] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < p.aux; i += kT) sMap[aux[i]] = (uint8_t)i;
        nb = p.aux <= 1 ? p.aux : 1 + (uint32_t)ozl_hb32(p.aux - 1u);
    } else {
        for (int s = threadIdx.x; s < 256; s += kT) sMap[s] = (uint8_t)s;
        nb = p.aux;
    }
    __syncthreads();
    uint64_t totalBits = (uint64_t)n * nb;
    uint32_t bytes = p.payload;
    for (uint32_t o = threadIdx.x; o < bytes; o += kT) {
        uint64_t bit0 = (uint64_t)o * 8, bit1 = bit0 + 8;
        uint32_t v = 0;
        if (nb && bit0 < totalBits) {
            uint64_t e0 = bit0 / nb, e1 = ((bit1 < totalBits ? bit1 : totalBits) - 1) / nb;
            for (uint64_t e = e0; e <= e1; e++) {
                long long sh = (long long)(e * nb) - (long long)bit0;
                uint32_t x = sMap[in[e]];
                v |= sh >= 0 ? (x << sh) : (x >> (-sh));
            }
        }
        if (flat && totalBits >= bit0 && totalBits < bit1) v |= 1u << (totalBits - bit0);  // terminating bit
        out[o] = (uint8_t)v;
    }
}

// ------------------------------------------------------------------ checksums and frames
__device__ inline size_t stored_bytes(const ozl_fstream* st, const int* order, int nbStored)
{
    size_t t = 0;
    for (int k = 0; k < nbStored; k++) t += st[order[k]].size;
    return t;
}

__global__ void k_frame_size(uint32_t B, ozl_ews ws)
{
    uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    ozl_eplan p = ws.plans[b];
    if (p.status != 0) { ws.frameOff[b] = 0; return; }
    ozl_fnode nodes[8];
    ozl_fstream st[12];
    int ns, order[12], nbStored;
    int nn = ozl_entropy_layout(&p, nodes, st, &ns);
    uint8_t hdr[512];
    size_t fh = ozl_write_frame_header(hdr, OZL_T_SERIAL, p.n, 1, 1);
    size_t ch = ozl_write_chunk_header(hdr, nodes, nn, st, ns, order, &nbStored);
    ws.frameOff[b] = fh + ch + stored_bytes(st, order, nbStored) + 8 + 1;
}

// exclusive scan of B+1 entries (last = total) in one workgroup
__global__ void __launch_bounds__(kT) k_scan_u64(uint64_t* v, uint32_t B)
{
    __shared__ uint64_t lds[4];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < B; base += kT) {
        uint32_t i = base + threadIdx.x;
        uint64_t x = i < B ? v[i] : 0;
        uint64_t total;
        uint64_t ex = block_scan_excl<uint64_t>(x, lds, total);
        if (i < B) v[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) v[B] = carry;
}

// Frame writer and compressed checksum in one pass over each frame (one
// 256-thread workgroup per frame). The checksum region (chunk header, stored
// streams, content checksum) is cut into 16-byte pieces: wave w takes the
// region's 1 KiB blocks w, w + 4, ..., lane l piece l of the block, gathers
// it from its source (the header and the content checksum in LDS, a stream
// in the input, payload or auxiliary area: two aligned 16-byte loads and a
// funnel shift when the piece lies inside one workspace stream, bytes
// otherwise), stores it into the frame and adds its XXH3 terms to the
// block's stripe sums (ozl_xxh3_sums_of). After a barrier wave 0 runs the
// scramble chain, hashes the tail stripes from the written frame, and writes
// the checksum and the end byte. The content checksums arrive in `status`
// (k_hist_hash), which ends up holding each block's status.
constexpr int kMaxSeg = 14;

struct FrameWriteLds {
    uint8_t hdr[512];          // frame header + chunk header
    uint32_t segStart[kMaxSeg];  // region offsets of the streams, then the content checksum
    uint64_t segSrc[kMaxSeg];
    uint32_t segSafe[kMaxSeg];   // bytes of the stream a 32-byte over-read may start below
    uint32_t nSeg, fh, ch, clen, hash;
    uint64_t bsum[100][8];
};

// byte x of the checksum region
__device__ inline uint32_t region_byte(const FrameWriteLds& F, uint32_t x)
{
    if (x < F.ch) return F.hdr[F.fh + x];
    if (x >= F.clen - 4) return (F.hash >> (8 * (x - (F.clen - 4)))) & 255;
    uint32_t k = 0;
    for (uint32_t i = 1; i < F.nSeg; i++) k = x >= F.segStart[i] ? i : k;
    return ozl_gld1(reinterpret_cast<const uint8_t*>((uintptr_t)F.segSrc[k]) + (x - F.segStart[k]));
}

// the 16 region bytes at x (x + 16 <= clen) as two little-endian words
__device__ inline void region_piece(const FrameWriteLds& F, uint32_t x, uint64_t& v0, uint64_t& v1)
{
    uint32_t k = 0;
    for (uint32_t i = 1; i < F.nSeg; i++) k = x >= F.segStart[i] ? i : k;
    const bool inside = x >= F.ch && x < F.clen - 4 && x + 16 <= (k + 1 < F.nSeg ? F.segStart[k + 1] : F.clen - 4) &&
                        x - F.segStart[k] < F.segSafe[k];
    uint32_t d0, d1, d2, d3;
    if (inside) {
        const uintptr_t ad = (uintptr_t)F.segSrc[k] + (x - F.segStart[k]);
        const uint8_t* base = reinterpret_cast<const uint8_t*>(ad & ~(uintptr_t)15);
        const uint32_t sh = (uint32_t)(ad & 15), bs = 8 * (sh & 3);
        const ozl_u32x4 lo = ozl_gld16(base), hi = ozl_gld16(base + 16);
        const uint32_t d[8] = { lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w };
        switch (sh >> 2) {
        case 0: ozl_xxh3_pick<0>(d, bs, d0, d1, d2, d3); break;
        case 1: ozl_xxh3_pick<1>(d, bs, d0, d1, d2, d3); break;
        case 2: ozl_xxh3_pick<2>(d, bs, d0, d1, d2, d3); break;
        default: ozl_xxh3_pick<3>(d, bs, d0, d1, d2, d3); break;
        }
    } else {
        uint32_t w[4] = { 0, 0, 0, 0 };
        for (uint32_t i = 0; i < 16; i++) w[i >> 2] |= region_byte(F, x + i) << (8 * (i & 3));
        d0 = w[0]; d1 = w[1]; d2 = w[2]; d3 = w[3];
    }
    v0 = d0 | ((uint64_t)d1 << 32);
    v1 = d2 | ((uint64_t)d3 << 32);
}

__global__ void __launch_bounds__(kT) k_frame_write_hash(const uint8_t* __restrict__ src, const uint64_t* __restrict__ off,
                                                         uint32_t B, ozl_ews ws, uint8_t* __restrict__ dst,
                                                         int* __restrict__ status)
{
    __shared__ FrameWriteLds F;
    __shared__ ozl_fstream st[12];
    __shared__ int order[12];
    const uint32_t b = blockIdx.x;
    if (b >= B) return;
    const ozl_eplan p = ws.plans[b];
    const int tid = threadIdx.x, lane = tid & 63;
    if (p.status != 0) {
        if (tid == 0) status[b] = (int)p.status;
        return;
    }
    if (tid == 0) {
        ozl_fnode nodes[8];
        int ns, nbStored;
        const int nn = ozl_entropy_layout(&p, nodes, st, &ns);
        const size_t fh = ozl_write_frame_header(F.hdr, OZL_T_SERIAL, p.n, 1, 1);
        const size_t ch = ozl_write_chunk_header(F.hdr + fh, nodes, nn, st, ns, order, &nbStored);
        F.fh = (uint32_t)fh;
        F.ch = (uint32_t)ch;
        const uint8_t* in = src + off[b];
        const uint8_t* aux = ws.aux + (size_t)b * OZL_AUX_BYTES;
        const uint8_t* pay = ws.payload + (size_t)b * ws.payloadCap;
        uint32_t pos = (uint32_t)ch, k = 0;
        for (int i = 0; i < nbStored; i++) {
            const ozl_fstream s = st[order[i]];
            if (s.size == 0) continue;
            const uint8_t* from = s.src == OZL_SRC_INPUT ? in : s.src == OZL_SRC_PAYLOAD ? pay : aux + s.srcOff;
            F.segStart[k] = pos;
            F.segSrc[k] = (uint64_t)(uintptr_t)from;
            // workspace streams may be read up to 31 bytes past their end
            // (the workspace has room); the caller's input may not
            F.segSafe[k] = s.src == OZL_SRC_INPUT ? (s.size > 32 ? s.size - 32 : 0u) : s.size;
            pos += s.size;
            k++;
        }
        if (k == 0) {   // no stored stream: the region is the header and the checksum
            F.segStart[0] = pos;
            F.segSrc[0] = 0;
            F.segSafe[0] = 0;
            k = 1;
        }
        F.nSeg = k;
        F.clen = pos + 4;
        F.hash = (uint32_t)status[b];
    }
    __syncthreads();
    uint8_t* fr = dst + ws.frameOff[b];
    const uint32_t fh = F.fh, clen = F.clen;
    for (uint32_t i = tid; i < fh; i += kT) fr[i] = F.hdr[i];
    uint8_t* rg = fr + fh;
    const uint32_t nbBlocks = clen > 240 ? (clen - 1) / 1024 : 0u;
    const ozl_xxh3_lane XL = ozl_xxh3_lane_init();
    for (uint32_t blk = (uint32_t)(tid >> 6); 1024 * blk < clen; blk += kT / 64) {
        const uint32_t x = 1024 * blk + 16 * (uint32_t)lane;
        uint64_t v0 = 0, v1 = 0;
        if (x + 16 <= clen) {
            region_piece(F, x, v0, v1);
            typedef __attribute__((address_space(1))) ozl_u32x4 g_u32x4;
            ozl_u32x4 v;
            v.x = (uint32_t)v0; v.y = (uint32_t)(v0 >> 32); v.z = (uint32_t)v1; v.w = (uint32_t)(v1 >> 32);
            *(g_u32x4*)(uintptr_t)(rg + x) = v;
        } else {
            for (uint32_t i = x; i < clen && i < x + 16; i++) ozl_gst1(rg + i, region_byte(F, i));
        }
        if (blk < nbBlocks) {
            uint64_t sa, sb;
            ozl_xxh3_sums_of(v0, v1, XL, sa, sb);
            if (lane < 4) {
                F.bsum[blk][2 * lane] = sa;
                F.bsum[blk][2 * lane + 1] = sb;
            }
        }
    }
    __syncthreads();
    if (tid >= 64) return;
    uint64_t hv;
    if (clen <= 240) {
        hv = ozl_xxh3_short(rg, clen);
    } else {
        const int j = lane & 3;
        uint64_t A = ozl_xxh3_init_acc(2 * j), Bc = ozl_xxh3_init_acc(2 * j + 1);
        for (uint32_t k = 0; k < nbBlocks; k++) {
            A = ozl_xxh3_scr(A + F.bsum[k][2 * j], XL.c0);
            Bc = ozl_xxh3_scr(Bc + F.bsum[k][2 * j + 1], XL.c1);
        }
        uint64_t acc = 0;
        for (int qq = 0; qq < 4; qq++) {
            const uint64_t a = __shfl(A, qq, 64), bq = __shfl(Bc, qq, 64);
            if (lane == 2 * qq) acc = a;
            if (lane == 2 * qq + 1) acc = bq;
        }
        hv = ozl_xxh3_finish(rg, clen, nbBlocks, acc);
    }
    if (lane == 0) {
        for (int k = 0; k < 4; k++) rg[clen + k] = (uint8_t)(hv >> (8 * k));
        rg[clen + 4] = 0;
        status[b] = 0;
    }
}

// plan, coder and packed payloads of blocks whose histograms are in place
int entropy_encode_after_hist(int hufMode, const uint8_t* s, const uint64_t* off, const uint64_t* offEnd, uint32_t B,
                              const ozl_ews& ws, hipStream_t st)
{
    if (hufMode) OZL_LAUNCH("k_plan", k_plan<1>, dim3(B), dim3(64), 0, st, s, off, offEnd, B, ws);
    else OZL_LAUNCH("k_plan", k_plan<0>, dim3(B), dim3(64), 0, st, s, off, offEnd, B, ws);
    if (hufMode) OZL_LAUNCH("k_huf_encode", k_huf_encode, dim3(B), dim3(kT), 0, st, s, off, B, ws);
    else OZL_LAUNCH("k_fse_encode", k_fse_encode, dim3((B + kFseGroups - 1) / kFseGroups), dim3(64), 0, st, s, off, B, ws);
    OZL_LAUNCH("k_pack_encode", k_pack_encode, dim3(B), dim3(kT), 0, st, s, off, B, ws);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
int entropy_encode_after_hist(int hufMode, const uint8_t* s, const uint64_t* off, uint32_t B, const ozl_ews& ws,
                              hipStream_t st)
{
    return entropy_encode_after_hist(hufMode, s, off, off + 1, B, ws, st);
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

// Workspace bytes for ozl_entropy_compress_batch on B blocks of at most
// maxBlock bytes.
size_t ozl_entropy_compress_workspace(uint32_t B, size_t maxBlock)
{
    return (size_t)B * (sizeof(ozl_eplan) + OZL_TAB_BYTES + OZL_AUX_BYTES + ozl_payload_cap(maxBlock)) + 4096;
}

// Upper bound of one frame for a block of n bytes.
size_t ozl_entropy_frame_bound(size_t n) { return n + n / 128 + 1024; }

// Compresses B serial blocks (block b = src[off[b], off[b+1]), each below
// 100000 bytes) into B frames with ZL_GRAPH_FSE (hufMode 0) or
// ZL_GRAPH_HUFFMAN (hufMode 1), format version 23, both checksums on: the
// frames the reference produces with ZL_CCtx_compressTypedRef on each block
// (include/openzl/zl_compress.h:274). Frames are written back to back into
// dst; frameOff (device, B+1 entries) receives their offsets and status
// (device, B ints) 0 or the error code of each block (a failed block gets an
// empty frame). All pointers are device pointers.
int ozl_entropy_compress_batch(int hufMode, const void* src, const uint64_t* off, uint32_t B, size_t maxBlock,
                               void* dst, size_t dstCapacity, uint64_t* frameOff, int* status, void* workspace,
                               size_t wsSize, void* stream)
{
    if (B == 0) return OZL_OK;
    if (maxBlock >= 100000) return OZL_ERR_PARAMETER_INVALID;
    if (wsSize < ozl_entropy_compress_workspace(B, maxBlock)) return OZL_ERR_INTERNAL_BUFFER;
    if (dstCapacity < (size_t)B * ozl_entropy_frame_bound(maxBlock)) return OZL_ERR_DSTCAPACITY_TOOSMALL;
    hipStream_t st = (hipStream_t)stream;
    ozl_ews ws = ozl_ews_carve(workspace, B, maxBlock);
    ws.frameOff = frameOff;
    const uint8_t* s = (const uint8_t*)src;
    // the content checksums wait in status until the frame writer has read them
    OZL_LAUNCH("k_hist_hash", k_hist_hash, dim3(B), dim3(kT), 0, st, s, off, B, ws, (uint32_t*)status);
    int rc = entropy_encode_after_hist(hufMode, s, off, B, ws, st);
    if (rc) return rc;
    OZL_LAUNCH("k_frame_size", k_frame_size, dim3((B + 63) / 64), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_scan_u64", k_scan_u64, dim3(1), dim3(kT), 0, st, frameOff, B);
    OZL_LAUNCH("k_frame_write_hash", k_frame_write_hash, dim3(B), dim3(kT), 0, st, s, off, B, ws, (uint8_t*)dst,
               status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"

int ozl_entropy_encode_chunks(int hufMode, const uint8_t* s, const uint64_t* off, const uint64_t* offEnd, uint32_t B,
                              size_t maxBlock, const ozl_ews& ws, hipStream_t st)
{
    if (B == 0) return OZL_OK;
    if (maxBlock >= 100000) return OZL_ERR_PARAMETER_INVALID;
    OZL_LAUNCH("k_hist", k_hist, dim3(B), dim3(kT), 0, st, s, off, offEnd, B, ws);
    return entropy_encode_after_hist(hufMode, s, off, offEnd, B, ws, st);
}
