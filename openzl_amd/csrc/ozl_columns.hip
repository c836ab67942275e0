// Batched numeric columns on gfx950: ZL_NODE_DELTA_INT -> ZL_NODE_ZIGZAG ->
// ZL_GRAPH_BITPACK, one frame per column, thousands of columns of uneven
// length per call.
//
// The graph per column follows the engine path (and the reference):
// delta_int (encode_delta_binding.c:11-45, header = first element),
// zigzag (encode_zigzag_binding.c:8-40), then the bitpack selector to
// bitpack_int (encode_bitpack_binding.c:92-147: nbBits = 1 + highbit(max),
// header byte0 = log2(width) << 6 | (nbBits - 1), byte1 = padding count when
// the packed bytes could hold more elements) and a serial store. The frame
// holds the three nodes, one stored stream, the content checksum (XXH3 of the
// column) and the compressed checksum (XXH3 of the chunk and the content
// checksum), as the engine writes it (ozl_frame.h).
//
// Work is cut into items of `tile` deltas (a multiple of 32, capped at 32 KiB
// of input per item); each pass over items is one persistent launch whose
// workgroups take items from a device counter until none is left:
//   encode  k_col_setup -> k_col_scan_items -> k_col_map -> hash pass over the
//           columns (content checksums, and the largest zigzag delta of each
//           column from the same reads) -> k_col_plan (nbBits, headers) -> k_col_scan_frames ->
//           k_col_pack (bits straight into the frames) -> k_col_finish1 ->
//           hash pass over the chunks (compressed checksums) -> k_col_finish2
//   decode  k_cold_parse -> k_col_scan_items -> k_col_map -> k_cold_sum (tile
//           sums of the deltas) -> k_cold_carry (per-column scan of the tile
//           sums) -> k_cold_apply (unpack, prefix sum, write) -> hash pass over
//           the columns and the chunks -> k_cold_verify
// A hash pass computes the XXH3 of many regions with many waves per region:
// every 1 KiB block's stripe sums are independent, only the scramble chain
// over a region's blocks is sequential (ozl_xxh3.h).
// Every launch clears what it needs with kernels, never with memsets, so a
// call can be captured in a hipGraph and replayed.
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_profile.h"

#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace {

constexpr int kCT = 256;          // threads per workgroup
constexpr int kHdrSlot = 96;      // per-column header bytes (frame + chunk header)
// Items hold at most 2048 deltas: 8 per thread keeps the item kernels under
// 64 VGPRs and 17 KB of LDS, so eight 256-thread workgroups (the wave limit)
// fit a CU; measured, that beats 4096-delta items at four workgroups per CU.
// Narrow columns (width <= 4) run in 32-bit arithmetic, wide ones in 64.
constexpr uint32_t kTileMax = 2048;
constexpr uint32_t kTileWide = 2048;
template <bool NARROW> constexpr int per_of() { return (int)(NARROW ? kTileMax : kTileWide) / kCT; }

struct ColPlan {                  // 64 bytes, one per column
    uint64_t src;                 // column address (encode: input, decode: packed stream)
    uint64_t m;                   // deltas (elements - 1)
    uint64_t packed;              // packed stream bytes
    unsigned long long maxZ;      // encode: largest zigzag delta; decode: first element
    uint32_t itemBase;            // first work item of the column
    uint32_t nItems;
    uint32_t tile;                // deltas per item
    uint32_t status;
    uint32_t hashes[2];           // content, compressed checksum (decode: stored values)
    uint8_t w, nb, fhLen, hdrLen; // width, bits, frame header bytes, frame + chunk header bytes
    uint8_t kind;                 // layout, see below
    uint8_t pad[3];
};
// Layouts (cctx.c:1127-1140: a successor graph's input below minStreamSize, 10
// bytes, is stored after a num->serial conversion, header [log2 width]):
//   kBitpack  n*w >= 10 and (n-1)*w >= 10: delta, zigzag, bitpack_int
//   kTinyIn   n*w < 10: the input itself is converted and stored
//   kTinyDz   (n-1)*w < 10 <= n*w: delta, then the deltas stored (the zigzag
//             node starts the successor graph, so it is not run either)
enum { kBitpack = 0, kTinyIn = 1, kTinyDz = 2 };
constexpr uint64_t kMinStream = 10;
static_assert(sizeof(ColPlan) == 64, "ColPlan layout");

struct ColCtr {
    uint32_t queue[8];            // work counters: column items (narrow 0, 1; wide 4, 5), hash items (2, 3)
    uint32_t totalItems;          // column items
    uint32_t error;               // bit 0: column items do not fit, bit 1: hash blocks do not fit
    uint32_t hTotalItems;         // hash items of the current hash pass
    uint32_t hTotalBlocks;
    uint32_t nWide;               // columns of width 8 (the wide item kernels exit at once when 0)
    uint32_t pad[3];
};

// one XXH3 input of a hash pass: its 1 KiB blocks are summed by many waves
// (items of kHG blocks), then one wave per region runs the scramble chain
struct HashReg {
    uint64_t ptr, len;
    uint32_t blockBase, nBlocks, itemBase, nItems;
    uint32_t zw;    // content pass over a bitpacked column: its width (the pass also
                    // takes the column's largest zigzag delta), else 0
    uint32_t pad;
};
constexpr uint32_t kHG = 16;

struct ColWs {
    ColCtr* ctr;
    ColPlan* plans;
    uint8_t* hdr;                 // kHdrSlot bytes per column
    HashReg* hreg;                // 2K regions
    uint64_t* hout;               // 2K hashes
    uint8_t* dyn;                 // item map, tile sums, hash item map, hash block sums (sized on the device)
    uint64_t end;                 // workspace end address
};

// ws layout: [ctr 256 B][plans K*64][hdr K*96][hreg 2K*32][hout 2K*8][dynamic...]
static ColWs carve(void* w, uint32_t K, size_t wsBytes)
{
    uint8_t* p = (uint8_t*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
    ColWs ws;
    ws.ctr = (ColCtr*)p; p += 256;
    ws.plans = (ColPlan*)p; p += (size_t)K * sizeof(ColPlan);
    ws.hdr = p; p += ((size_t)K * kHdrSlot + 255) & ~(size_t)255;
    ws.hreg = (HashReg*)p; p += (size_t)2 * K * sizeof(HashReg);
    ws.hout = (uint64_t*)p; p += (size_t)2 * K * 8;
    ws.dyn = (uint8_t*)(((uintptr_t)p + 255) & ~(uintptr_t)255);
    ws.end = (uint64_t)(uintptr_t)w + wsBytes;
    return ws;
}
static size_t fixed_bytes(uint32_t K) { return 256 + 256 + (size_t)K * sizeof(ColPlan) + (size_t)K * kHdrSlot + 256 +
                                              (size_t)2 * K * (sizeof(HashReg) + 8) + 256; }

__device__ inline uint32_t* item_col(const ColWs& ws) { return reinterpret_cast<uint32_t*>(ws.dyn); }
__device__ inline uint64_t* tile_sums(const ColWs& ws, uint32_t items)
{
    return reinterpret_cast<uint64_t*>(((uintptr_t)(ws.dyn + 4 * (size_t)items) + 7) & ~(uintptr_t)7);
}
__device__ inline uint32_t* hash_map(const ColWs& ws, uint32_t items)
{
    return reinterpret_cast<uint32_t*>(tile_sums(ws, items) + items);
}
__device__ inline uint64_t* hash_sums(const ColWs& ws, uint32_t items, uint32_t hItems)
{
    return reinterpret_cast<uint64_t*>(((uintptr_t)(hash_map(ws, items) + hItems) + 63) & ~(uintptr_t)63);
}

__device__ inline uint64_t ld_elem(const uint8_t* p, int w) { return ozl_gld_w(p, w); }

__device__ inline uint64_t wmask(int w) { return w == 8 ? ~0ull : (1ull << (8 * w)) - 1; }

// zigzag of the width-w difference b - a (encode_zigzag_kernel.c:8-59)
__device__ inline uint64_t zz_delta(uint64_t a, uint64_t b, int w)
{
    const uint64_t d = (b - a) & wmask(w);
    const uint64_t sign = (d >> (8 * w - 1)) & 1;
    return ((d << 1) ^ (0 - sign)) & wmask(w);
}

__device__ inline uint64_t unzz(uint64_t z, int w) { return ((z >> 1) ^ (0 - (z & 1))) & wmask(w); }

// workgroup-wide item fetch from a device counter; returns the item or ~0u
// Work queue: a workgroup takes g consecutive items per atomic. One counter
// serves thousands of workgroups and its atomics serialise, so each grab
// takes about 1/16 of a workgroup's fair share (grab_size). next_item
// returns the next item of the current range, taking a new range when it is
// used up, or ~0u when the queue is empty; it starts with a barrier, so the
// previous item's LDS is free.
__device__ inline uint32_t grab_size(uint32_t total)
{
    const uint32_t g = total / (gridDim.x * 16u);
    return g ? g : 1u;
}
__device__ inline uint32_t next_item(uint32_t* q, uint32_t total, uint32_t g, uint32_t* sItem, uint32_t& it,
                                     uint32_t& end)
{
    __syncthreads();
    if (it + 1 < end) return ++it;
    if (threadIdx.x == 0) *sItem = atomicAdd(q, g);
    __syncthreads();
    it = *sItem;
    if (it >= total) return ~0u;
    end = it + g < total ? it + g : total;
    return it;
}

template <typename T>
__device__ T block_sum(T v, T* lds)
{
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) lds[wave] = v;
    __syncthreads();
    T t = 0;
    for (int k = 0; k < kCT / 64; k++) t += lds[k];
    return t;
}

// exclusive scan of the workgroup's values (returns the exclusive prefix)
template <typename T>
__device__ T block_excl(T v, T* lds, T& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
    for (int d = 1; d < 64; d <<= 1) {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T off = 0, all = 0;
    for (int k = 0; k < kCT / 64; k++) {
        const T s = lds[k];
        if (k < wave) off += s;
        all += s;
    }
    total = all;
    return off + x - v;
}

// ====================================================================== encode
__global__ void __launch_bounds__(kCT) k_col_setup(const uint8_t* __restrict__ src, const uint64_t* __restrict__ colOff,
                                                  const uint8_t* __restrict__ widths, uint32_t K, uint32_t tile, ColWs ws)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 16) reinterpret_cast<uint32_t*>(ws.ctr)[threadIdx.x] = 0;
    if (c >= K) return;
    ColPlan p = {};
    const uint64_t a = colOff[2 * c], b = a + colOff[2 * c + 1];
    const int w = widths[c];
    p.w = (uint8_t)w;
    p.src = (uint64_t)(uintptr_t)(src + a);
    const bool wok = w == 1 || w == 2 || w == 4 || w == 8;
    if (!wok || b <= a || (b - a) % (uint64_t)(wok ? w : 1) || ((p.src) & (uint64_t)(wok ? w - 1 : 0))) {
        p.status = OZL_ERR_NODE_INVALID_INPUT;
    } else {
        const uint64_t n = (b - a) / (uint64_t)w;
        p.m = n - 1;
        if (n * (uint64_t)w < kMinStream) p.kind = kTinyIn;
        else if ((n - 1) * (uint64_t)w < kMinStream) p.kind = kTinyDz;
        else {
            const uint32_t tmax = w <= 4 ? kTileMax : kTileWide;
            const uint32_t t = tile < tmax ? tile : tmax;
            p.tile = t;
            p.nItems = (uint32_t)((p.m + t - 1) / t);
        }
    }
    ws.plans[c] = p;
}

// item bases: exclusive scan of nItems (one workgroup). When the item map
// and the tile sums would pass the workspace end, no item runs and every
// column reports OZL_ERR_PARAMETER_INVALID.
__global__ void __launch_bounds__(kCT) k_col_scan_items(uint32_t K, ColWs ws)
{
    __shared__ uint32_t lds[kCT / 64];
    const uint32_t per = (K + kCT - 1) / kCT, b0 = threadIdx.x * per, b1 = b0 + per < K ? b0 + per : K;
    uint32_t s = 0, wide = 0;
    for (uint32_t i = b0; i < b1; i++) {
        s += ws.plans[i].nItems;
        wide += ws.plans[i].nItems && ws.plans[i].w > 4;
    }
    uint32_t total, nWide;
    uint32_t off = block_excl<uint32_t>(s, lds, total);
    block_excl<uint32_t>(wide, lds, nWide);
    for (uint32_t i = b0; i < b1; i++) {
        ws.plans[i].itemBase = off;
        off += ws.plans[i].nItems;
    }
    if (threadIdx.x == 0) {
        ws.ctr->nWide = nWide;
        const bool fits = (uint64_t)(uintptr_t)(hash_map(ws, total)) <= ws.end;
        ws.ctr->totalItems = fits ? total : 0u;
        if (!fits) ws.ctr->error |= 1u;
    }
}

// item-to-column map: one workgroup per column writes its items
__global__ void __launch_bounds__(64) k_col_map(uint32_t K, ColWs ws)
{
    const uint32_t c = blockIdx.x;
    if (c >= K || ws.ctr->error) return;
    const ColPlan& p = ws.plans[c];
    for (uint32_t i = threadIdx.x; i < p.nItems; i += 64) item_col(ws)[p.itemBase + i] = c;
}

// ---------------------------------------------------------------- hash passes
// Every region's XXH3: k_hash_scan places blocks and items, k_hash_map maps
// items to regions, k_hash_sums (persistent, one wave per workgroup) sums the
// 16 stripes of each block (ozl_xxh3_block_sums), k_hash_chain runs each
// region's scramble chain over its block sums, the tail stripes and the merge
// (one wave per region) into hout.
__device__ inline void hreg_set(HashReg& r, uint64_t ptr, uint64_t len, uint32_t zw = 0)
{
    r.ptr = ptr;
    r.len = len;
    r.zw = zw;
    r.nBlocks = len > 240 ? (uint32_t)((len - 1) / 1024) : 0u;
    r.nItems = (r.nBlocks + kHG - 1) / kHG;
    r.blockBase = r.itemBase = 0;
}

__global__ void __launch_bounds__(kCT) k_hreg_content(uint32_t K, ColWs ws)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    hreg_set(ws.hreg[c], p.src, p.status ? 0 : (p.m + 1) * (uint64_t)p.w, p.status || p.kind != kBitpack ? 0u : p.w);
}

// compressed checksum input: the chunk and the content checksum
__global__ void __launch_bounds__(kCT) k_hreg_frames(uint32_t K, ColWs ws, const uint8_t* __restrict__ dst,
                                                    const uint64_t* __restrict__ frameOff)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    hreg_set(ws.hreg[c], (uint64_t)(uintptr_t)(dst + frameOff[c] + p.fhLen),
             p.status ? 0 : p.hdrLen - p.fhLen + p.packed + 4);
}

// decode: region c = decoded column, region K + c = the frame's chunk and content checksum
__global__ void __launch_bounds__(kCT) k_hreg_decode(uint32_t K, ColWs ws, const uint8_t* __restrict__ frames,
                                                    const uint64_t* __restrict__ foff, const uint8_t* __restrict__ dst,
                                                    const uint64_t* __restrict__ doff)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    const bool ok = p.status == 0 && !ws.ctr->error;
    hreg_set(ws.hreg[c], (uint64_t)(uintptr_t)(dst + doff[2 * c]), ok ? (p.m + 1) * (uint64_t)p.w : 0);
    hreg_set(ws.hreg[K + c], (uint64_t)(uintptr_t)(frames + foff[c] + p.fhLen),
             ok ? p.hdrLen - p.fhLen + p.packed + 4 : 0);
}

__global__ void __launch_bounds__(kCT) k_hash_scan(uint32_t R, ColWs ws)
{
    __shared__ uint32_t lds[kCT / 64];
    const uint32_t per = (R + kCT - 1) / kCT, b0 = threadIdx.x * per, b1 = b0 + per < R ? b0 + per : R;
    uint32_t sb = 0, si = 0;
    for (uint32_t i = b0; i < b1; i++) { sb += ws.hreg[i].nBlocks; si += ws.hreg[i].nItems; }
    uint32_t tb, ti;
    uint32_t ob = block_excl<uint32_t>(sb, lds, tb);
    __syncthreads();
    uint32_t oi = block_excl<uint32_t>(si, lds, ti);
    for (uint32_t i = b0; i < b1; i++) {
        ws.hreg[i].blockBase = ob;
        ws.hreg[i].itemBase = oi;
        ob += ws.hreg[i].nBlocks;
        oi += ws.hreg[i].nItems;
    }
    if (threadIdx.x == 0) {
        const uint32_t items = ws.ctr->totalItems;
        const bool fits = (uint64_t)(uintptr_t)(hash_sums(ws, items, ti) + 8 * (size_t)tb) <= ws.end;
        ws.ctr->hTotalItems = fits ? ti : 0u;
        ws.ctr->hTotalBlocks = fits ? tb : 0u;
        if (!fits) ws.ctr->error |= 2u;
    }
}

__global__ void __launch_bounds__(64) k_hash_map(uint32_t R, ColWs ws)
{
    const uint32_t r = blockIdx.x;
    if (r >= R || (ws.ctr->error & 2u)) return;
    const HashReg& h = ws.hreg[r];
    uint32_t* map = hash_map(ws, ws.ctr->totalItems);
    for (uint32_t i = threadIdx.x; i < h.nItems; i += 64) map[h.itemBase + i] = r;
}

// low 8w bits of a 32-bit word (w <= 4)
constexpr uint32_t mask32(int W) { return W >= 4 ? 0xFFFFFFFFu : (1u << (8 * (W & 3))) - 1u; }

// largest zigzag delta among the deltas whose first element lies in a lane's
// 16 bytes (v0, v1) of a column of width W; nxt is the element after them
template <int W>
__device__ inline unsigned long long lane_zz_max(uint64_t v0, uint64_t v1, uint64_t nxt, bool hasNext,
                                                 unsigned long long mx)
{
    constexpr int nE = 16 / W;
    if constexpr (W <= 4) {
        constexpr uint32_t M = mask32(W);
        const uint32_t dw[4] = { (uint32_t)v0, (uint32_t)(v0 >> 32), (uint32_t)v1, (uint32_t)(v1 >> 32) };
        uint32_t prev = dw[0] & M, m32 = 0;
#pragma unroll
        for (int e = 1; e <= nE; e++) {
            const uint32_t cur = e < nE ? (dw[(e * W) / 4] >> (8 * ((e * W) % 4))) & M : (uint32_t)nxt & M;
            const uint32_t d = (cur - prev) & M;
            const uint32_t z = ((d << 1) ^ (0u - (d >> (8 * W - 1)))) & M;
            if (e < nE || hasNext) m32 = z > m32 ? z : m32;
            prev = cur;
        }
        return (unsigned long long)m32 > mx ? (unsigned long long)m32 : mx;
    } else {
        uint64_t z = zz_delta(v0, v1, 8);
        mx = z > mx ? z : mx;
        if (hasNext) {
            z = zz_delta(v1, nxt, 8);
            mx = z > mx ? z : mx;
        }
        return mx;
    }
}

__global__ void __launch_bounds__(64) k_hash_sums(ColWs ws, int q)
{
    const uint32_t total = ws.ctr->hTotalItems, items = ws.ctr->totalItems;
    const uint32_t* map = hash_map(ws, items);
    uint64_t* sums = hash_sums(ws, items, total);
    const ozl_xxh3_lane L = ozl_xxh3_lane_init();
    const int lane = threadIdx.x;
    const uint32_t g = grab_size(total);
    uint32_t it = 0, end = 0;
    for (;;) {
        if (it + 1 < end) {
            it++;
        } else {
            uint32_t v = 0;
            if (lane == 0) v = atomicAdd(&ws.ctr->queue[q], g);
            it = (uint32_t)__shfl((int)v, 0, 64);
            if (it >= total) break;
            end = it + g < total ? it + g : total;
        }
        const HashReg h = ws.hreg[map[it]];
        const uint32_t b0 = (it - h.itemBase) * kHG, b1 = b0 + kHG < h.nBlocks ? b0 + kHG : h.nBlocks;
        const uint8_t* in = reinterpret_cast<const uint8_t*>((uintptr_t)h.ptr);
        unsigned long long mx = 0;
        // the item's blocks are read up front (addresses clamped into the
        // item, so every load is issued unconditionally), then summed
        uint64_t bw[kHG][2];
#pragma unroll
        for (uint32_t k = 0; k < kHG; k++) {
            const uint32_t b = b0 + k < b1 ? b0 + k : b1 - 1;
            ozl_xxh3_block_load(in + (size_t)b * 1024, ozl_xxh3_mode(in, h.len, b), bw[k][0], bw[k][1]);
        }
        const uint64_t msk = wmask(h.zw ? (int)h.zw : 1);
        const int w = (int)h.zw;
        // the element after the item's last block, for the last delta
        uint64_t after = 0;
        const uint64_t nOff = (uint64_t)b1 * 1024;
        const bool hasAfter = w && nOff < h.len;
        if (hasAfter && lane == 63) after = ld_elem(in + nOff, w);
#pragma unroll
        for (uint32_t k = 0; k < kHG; k++) {
            if (b0 + k >= b1) continue;
            const uint32_t b = b0 + k;
            uint64_t sa, sb;
            ozl_xxh3_sums_of(bw[k][0], bw[k][1], L, sa, sb);
            if (lane < 4) {
                uint64_t* o = sums + 8 * ((size_t)h.blockBase + b) + 2 * L.j;
                ozl_gst8(o, sa);
                ozl_gst8(o + 1, sb);
            }
            if (w) {
                // deltas whose first element is in this lane's 16 bytes
                const uint64_t v0 = bw[k][0], v1 = bw[k][1];
                uint64_t nxt = (uint64_t)__shfl_down((unsigned long long)(v0 & msk), 1, 64);
                // first element of the next block (lane 0's first word)
                const uint64_t nb0 =
                    k + 1 < kHG ? (uint64_t)__shfl((unsigned long long)(bw[k + 1 < kHG ? k + 1 : k][0] & msk), 0, 64) : 0;
                bool hasNext = true;
                if (lane == 63) {
                    const bool last = b0 + k + 1 >= b1;
                    hasNext = last ? hasAfter : true;
                    nxt = last ? after : nb0;
                }
                switch (w) {
                case 1: mx = lane_zz_max<1>(v0, v1, nxt, hasNext, mx); break;
                case 2: mx = lane_zz_max<2>(v0, v1, nxt, hasNext, mx); break;
                case 4: mx = lane_zz_max<4>(v0, v1, nxt, hasNext, mx); break;
                default: mx = lane_zz_max<8>(v0, v1, nxt, hasNext, mx); break;
                }
            }
        }
        if (h.zw) {
            for (int d = 32; d >= 1; d >>= 1) {
                const unsigned long long o = __shfl_xor(mx, d, 64);
                mx = o > mx ? o : mx;
            }
            if (lane == 0 && mx) atomicMax(&ws.plans[map[it]].maxZ, mx);
        }
    }
}

__global__ void __launch_bounds__(64) k_hash_chain(uint32_t R, ColWs ws)
{
    const uint32_t r = blockIdx.x;
    if (r >= R) return;
    const HashReg h = ws.hreg[r];
    const int lane = threadIdx.x;
    const uint8_t* in = reinterpret_cast<const uint8_t*>((uintptr_t)h.ptr);
    if (h.len == 0 || (ws.ctr->error & 2u)) {
        if (lane == 0) ws.hout[r] = 0;
        return;
    }
    if (h.zw) {
        // deltas whose first element follows the summed blocks
        const int w = (int)h.zw;
        const uint64_t n = h.len / (uint64_t)w, i0 = (uint64_t)h.nBlocks * 1024 / (uint64_t)w;
        unsigned long long mx = 0;
        for (uint64_t i = i0 + lane; i + 1 < n; i += 64) {
            const uint64_t z = zz_delta(ld_elem(in + i * w, w), ld_elem(in + (i + 1) * w, w), w);
            mx = z > mx ? z : mx;
        }
        for (int d = 32; d >= 1; d >>= 1) {
            const unsigned long long o = __shfl_xor(mx, d, 64);
            mx = o > mx ? o : mx;
        }
        if (lane == 0 && mx) atomicMax(&ws.plans[r].maxZ, mx);
    }
    uint64_t hv;
    if (h.len <= 240) {
        hv = ozl_xxh3_short(in, h.len);
    } else {
        const uint64_t* sums = hash_sums(ws, ws.ctr->totalItems, ws.ctr->hTotalItems) + 8 * (size_t)h.blockBase;
        const int i = lane & 7;
        uint64_t acc = ozl_xxh3_init_acc(i);
        const uint64_t c = ozl_sec64(192 - 64 + 8 * i);
        uint32_t b = 0;
        for (; b + 8 <= h.nBlocks; b += 8) {
            uint64_t v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = sums[8 * (size_t)(b + k) + i];
#pragma unroll
            for (int k = 0; k < 8; k++) acc = ozl_xxh3_scr(acc + v[k], c);
        }
        for (; b < h.nBlocks; b++) acc = ozl_xxh3_scr(acc + sums[8 * (size_t)b + i], c);
        hv = ozl_xxh3_finish(in, h.len, h.nBlocks, acc);
    }
    if (lane == 0) ws.hout[r] = hv;
}

template <int W> struct ElemT;
template <> struct ElemT<1> { using T = uint8_t; };
template <> struct ElemT<2> { using T = uint16_t; };
template <> struct ElemT<4> { using T = uint32_t; };
template <> struct ElemT<8> { using T = uint64_t; };

// zigzag deltas d0+tid+256k (k < kPer) of an item: every load of the item is
// issued before the first use, so their latency overlaps
template <int W, int KP>
__device__ inline void item_zz(const uint8_t* x, uint64_t d0, uint64_t d1, uint64_t* z)
{
    using T = typename ElemT<W>::T;
    T a[KP], b[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const uint64_t d = d0 + threadIdx.x + (uint64_t)kCT * k;
        const uint64_t dd = d < d1 ? d : d0;
        a[k] = (T)ozl_gld_w(x + dd * W, W);
        b[k] = (T)ozl_gld_w(x + (dd + 1) * W, W);
    }
#pragma unroll
    for (int k = 0; k < KP; k++) z[k] = zz_delta((uint64_t)a[k], (uint64_t)b[k], W);
}

// the same for narrow columns, in 32 bits
template <int W, int KP>
__device__ inline void item_zz32(const uint8_t* x, uint64_t d0, uint64_t d1, uint32_t* z)
{
    constexpr uint32_t M = mask32(W);
    uint32_t a[KP], b[KP];
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const uint64_t d = d0 + threadIdx.x + (uint64_t)kCT * k;
        const uint64_t dd = d < d1 ? d : d0;
        a[k] = (uint32_t)ozl_gld_w(x + dd * W, W);
        b[k] = (uint32_t)ozl_gld_w(x + (dd + 1) * W, W);
    }
#pragma unroll
    for (int k = 0; k < KP; k++) {
        const uint32_t d = (b[k] - a[k]) & M;
        z[k] = ((d << 1) ^ (0u - (d >> ((8 * W - 1) & 31)))) & M;
    }
}

template <int KP>
__device__ inline void item_zz_w(int w, const uint8_t* x, uint64_t d0, uint64_t d1, uint64_t* z)
{
    switch (w) {
    case 1: item_zz<1, KP>(x, d0, d1, z); break;
    case 2: item_zz<2, KP>(x, d0, d1, z); break;
    case 4: item_zz<4, KP>(x, d0, d1, z); break;
    default: item_zz<8, KP>(x, d0, d1, z); break;
    }
}

// the chunk of each layout (ozl_frame.h); `stored` = bytes of the stored
// stream, `first` = the delta header value. Returns the chunk header bytes.
__device__ inline size_t col_chunk_header(uint8_t* out, int kind, int w, int nb, uint64_t m, uint64_t stored,
                                          uint64_t first)
{
    const int lg = w == 1 ? 0 : w == 2 ? 1 : w == 4 ? 2 : 3;
    ozl_fnode nodes[3];
    ozl_fstream st[4];
    for (int k = 0; k < 4; k++) st[k] = ozl_fstream{ 0, 0, OZL_SRC_NONE, 0 };
    int nn;
    if (kind == kTinyIn) {
        nodes[0] = ozl_fnode{ OZL_TR_CONV_NUM_TO_SERIAL_LE, 1, 0, 0, 1, {} };
        nodes[0].hdr[0] = (uint8_t)lg;
        nn = 1;
    } else {
        nodes[0] = ozl_fnode{ OZL_TR_DELTA_INT, (uint8_t)w, 0, 0, 1, {} };
        for (int k = 0; k < w; k++) nodes[0].hdr[k] = (uint8_t)(first >> (8 * k));
        if (kind == kTinyDz) {
            nodes[1] = ozl_fnode{ OZL_TR_CONV_NUM_TO_SERIAL_LE, 1, 0, 1, 2, {} };
            nodes[1].hdr[0] = (uint8_t)lg;
            nn = 2;
        } else {
            nodes[1] = ozl_fnode{ OZL_TR_ZIGZAG, 0, 0, 1, 2, {} };
            nodes[2] = ozl_fnode{ OZL_TR_BITPACK_INT, 1, 0, 2, 3, {} };
            nodes[2].hdr[0] = (uint8_t)((lg << 6) | (nb - 1));
            const uint64_t maxElts = stored * 8 / (uint64_t)nb;
            if (maxElts > m) {
                nodes[2].hdr[1] = (uint8_t)(maxElts - m);
                nodes[2].hdrLen = 2;
            }
            nn = 3;
        }
    }
    st[nn].size = (uint32_t)stored;
    st[nn].store = 1;
    int order[4], nbStored = 0;
    return ozl_write_chunk_header(out, nodes, nn, st, nn + 1, order, &nbStored);
}

// per column: bits, packed size, frame and chunk headers; frame sizes go to
// frameOff[c] for the scan. Tiny layouts also put their stored bytes (<= 9)
// after the header in the column's header slot.
__global__ void __launch_bounds__(64) k_col_plan(uint32_t K, ColWs ws, uint64_t* __restrict__ frameOff)
{
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    ColPlan p = ws.plans[c];
    if (p.status) {
        if (threadIdx.x == 0) frameOff[c] = 0;
        return;
    }
    const int w = p.w;
    const uint8_t* x = reinterpret_cast<const uint8_t*>((uintptr_t)p.src);
    const uint64_t n = p.m + 1;
    const uint64_t h = ws.hout[c];  // content hash (hash pass over the columns)
    if (threadIdx.x != 0) return;
    uint8_t* hb = ws.hdr + (size_t)c * kHdrSlot;
    const size_t fh = ozl_write_frame_header(hb, OZL_T_NUMERIC, n * (uint64_t)w, 1, 1);
    const uint64_t first = ld_elem(x, w);
    int nb = 0;
    uint64_t stored;
    if (p.kind == kTinyIn) {
        stored = n * (uint64_t)w;
    } else if (p.kind == kTinyDz) {
        stored = p.m * (uint64_t)w;
    } else {
        nb = 1 + (p.maxZ ? ozl_hb64(p.maxZ) : 0);
        stored = (p.m * (uint64_t)nb + 7) / 8;
    }
    const size_t ch = col_chunk_header(hb + fh, p.kind, w, nb, p.m, stored, first);
    if (p.kind == kTinyIn) {
        for (uint64_t k = 0; k < stored; k++) hb[fh + ch + k] = x[k];
    } else if (p.kind == kTinyDz) {
        for (uint64_t d = 0; d < p.m; d++) {
            const uint64_t dv = (ld_elem(x + (d + 1) * w, w) - ld_elem(x + d * w, w)) & wmask(w);
            for (int k = 0; k < w; k++) hb[fh + ch + d * w + k] = (uint8_t)(dv >> (8 * k));
        }
    }
    p.nb = (uint8_t)nb;
    p.packed = stored;
    p.fhLen = (uint8_t)fh;
    p.hdrLen = (uint8_t)(fh + ch);
    p.hashes[0] = (uint32_t)h;
    ws.plans[c] = p;
    frameOff[c] = fh + ch + stored + 9;
}

// frame offsets: exclusive scan of the sizes in frameOff[0..K) -> frameOff[0..K]
__global__ void __launch_bounds__(kCT) k_col_scan_frames(uint32_t K, uint64_t* __restrict__ frameOff, size_t dstCap,
                                                        ColWs ws)
{
    __shared__ uint64_t lds[kCT / 64];
    const uint32_t per = (K + kCT - 1) / kCT, b0 = threadIdx.x * per, b1 = b0 + per < K ? b0 + per : K;
    uint64_t s = 0;
    for (uint32_t i = b0; i < b1; i++) s += frameOff[i];
    uint64_t total;
    uint64_t off = block_excl<uint64_t>(s, lds, total);
    for (uint32_t i = b0; i < b1; i++) {
        const uint64_t v = frameOff[i];
        frameOff[i] = off;
        if (off + v > dstCap && ws.plans[i].status == 0) ws.plans[i].status = OZL_ERR_DSTCAPACITY_TOOSMALL;
        off += v;
    }
    if (threadIdx.x == 0) frameOff[K] = total;
}

// An item's zigzag deltas go to LDS (coalesced loads); then every thread
// builds whole 32-bit words of the item's bits from the deltas that overlap
// them (items start at multiples of 32 deltas, so on word boundaries of the
// stream) and consecutive threads store consecutive words.
// NARROW takes the items of columns of width <= 4, the wide variant the rest
template <bool NARROW>
__global__ void __launch_bounds__(kCT) k_col_pack(uint32_t K, ColWs ws, uint8_t* __restrict__ dst,
                                                 const uint64_t* __restrict__ frameOff)
{
    constexpr int KP = per_of<NARROW>();
    __shared__ uint32_t sItem, sCol;
    // narrow: u32 element e at e + e / 32; wide: u64 element e at e + e / 16
    __shared__ uint64_t Z[kTileWide + kTileWide / 16 + 2];
    if (!NARROW && ws.ctr->nWide == 0) return;
    const uint32_t total = ws.ctr->totalItems;
    const uint32_t g = grab_size(total);
    uint32_t cur = 0, end = 0;
    for (;;) {
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)next_item(&ws.ctr->queue[NARROW ? 1 : 5], total, g, &sItem, cur, end));
        if (it == ~0u) break;
        if (threadIdx.x == 0) sCol = item_col(ws)[it];
        __syncthreads();
        const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)sCol);
        const ColPlan& P = ws.plans[c];
        if (P.status || (P.w <= 4) != NARROW) continue;
        const int nb = P.nb;
        const uint32_t t = it - P.itemBase;
        const uint64_t d0 = (uint64_t)t * P.tile, d1 = d0 + P.tile < P.m ? d0 + P.tile : P.m;
        const uint32_t cnt = (uint32_t)(d1 - d0);
        uint8_t* fr = dst + frameOff[c];
        if (t == 0 && threadIdx.x < P.hdrLen) ozl_gst1(fr + threadIdx.x, ws.hdr[(size_t)c * kHdrSlot + threadIdx.x]);
        uint8_t* o = fr + P.hdrLen + d0 * (uint64_t)nb / 8;
        const uint64_t nbits = (uint64_t)cnt * nb;
        const uint32_t nW = (uint32_t)(nbits / 32), rem = (uint32_t)(nbits % 32);
        if constexpr (NARROW) {
            // 32-bit deltas and words
            uint32_t* Z32 = reinterpret_cast<uint32_t*>(Z);
            {
                uint32_t z[KP];
                const uint8_t* x = reinterpret_cast<const uint8_t*>((uintptr_t)P.src);
                switch (P.w) {
                case 1: item_zz32<1, KP>(x, d0, d1, z); break;
                case 2: item_zz32<2, KP>(x, d0, d1, z); break;
                default: item_zz32<4, KP>(x, d0, d1, z); break;
                }
#pragma unroll
                for (int k = 0; k < KP; k++) {
                    const uint32_t i = threadIdx.x + kCT * k;
                    if (i < cnt) Z32[i + (i >> 5)] = z[k];
                }
            }
            __syncthreads();
            const uint32_t nb32 = (uint32_t)nb, nbits32 = (uint32_t)nbits;
            for (uint32_t wi = threadIdx.x; wi < nW + (rem ? 1u : 0u); wi += kCT) {
                const uint32_t bit0 = 32 * wi;
                uint32_t e = bit0 / nb32;
                const uint32_t bend = bit0 + 32 < nbits32 ? bit0 + 32 : nbits32;
                uint32_t word = 0;
                for (; e * nb32 < bend; e++) {
                    const int off = (int)(e * nb32) - (int)bit0;
                    const uint32_t v = Z32[e + (e >> 5)];
                    word |= off >= 0 ? v << off : v >> (-off);
                }
                if (wi < nW) {
                    ozl_gst4(o + 4 * (size_t)wi, word);
                } else {
                    for (uint32_t k = 0; 8 * k < rem; k++) ozl_gst1(o + 4 * (size_t)wi + k, word >> (8 * k));
                }
            }
            continue;
        }
        {
            uint64_t z[KP];
            item_zz_w<KP>(P.w, reinterpret_cast<const uint8_t*>((uintptr_t)P.src), d0, d1, z);
#pragma unroll
            for (int k = 0; k < KP; k++) {
                const uint32_t i = threadIdx.x + kCT * k;
                if (i < cnt) Z[i + (i >> 4)] = z[k];
            }
        }
        __syncthreads();
        for (uint32_t wi = threadIdx.x; wi < nW + (rem ? 1u : 0u); wi += kCT) {
            const uint64_t bit0 = 32 * (uint64_t)wi;
            uint32_t e = (uint32_t)(bit0 / nb);
            const uint64_t bend = bit0 + 32 < nbits ? bit0 + 32 : nbits;
            uint64_t word = 0;
            for (; (uint64_t)e * nb < bend; e++) {
                const int64_t off = (int64_t)e * nb - (int64_t)bit0;
                const uint64_t v = Z[e + (e >> 4)];
                word |= off >= 0 ? v << off : v >> (-off);
            }
            if (wi < nW) {
                ozl_gst4(o + 4 * (size_t)wi, (uint32_t)word);
            } else {
                for (uint32_t k = 0; 8 * k < rem; k++) ozl_gst1(o + 4 * (size_t)wi + k, (uint32_t)(word >> (8 * k)));
            }
        }
    }
}

// tiny frames' bytes and every frame's content checksum (before the hash
// pass over the chunks)
__global__ void __launch_bounds__(64) k_col_finish1(uint32_t K, ColWs ws, uint8_t* __restrict__ dst,
                                                   const uint64_t* __restrict__ frameOff)
{
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    if (p.status) return;
    uint8_t* fr = dst + frameOff[c];
    const uint64_t e = p.hdrLen + p.packed;
    if (p.kind != kBitpack && threadIdx.x < e) fr[threadIdx.x] = ws.hdr[(size_t)c * kHdrSlot + threadIdx.x];
    if (threadIdx.x < 4) fr[e + threadIdx.x] = (uint8_t)(p.hashes[0] >> (8 * threadIdx.x));
}

// compressed checksum (hash pass over chunk + content checksum), end byte, status
__global__ void __launch_bounds__(kCT) k_col_finish2(uint32_t K, ColWs ws, uint8_t* __restrict__ dst,
                                                    const uint64_t* __restrict__ frameOff, int* __restrict__ status)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    int st = (int)p.status;
    if (!st && ws.ctr->error) st = OZL_ERR_PARAMETER_INVALID;
    if (!st) {
        uint8_t* fr = dst + frameOff[c];
        const uint64_t e = p.hdrLen + p.packed;
        const uint64_t h = ws.hout[c];
        for (int k = 0; k < 4; k++) fr[e + 4 + k] = (uint8_t)(h >> (8 * k));
        fr[e + 8] = 0;
    }
    status[c] = st;
}

// ====================================================================== decode
// Frame check: the width and the layout follow from the frame's input size
// and the candidate width; the bytes before the stored stream must be exactly
// what the encoder writes for them (col_chunk_header), with nbBits tried
// until the rebuilt header matches. Tiny layouts are decoded here.
__global__ void __launch_bounds__(64) k_cold_parse(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                  uint32_t K, uint8_t* __restrict__ dst, const uint64_t* __restrict__ doff,
                                                  uint32_t tile, ColWs ws)
{
    const uint32_t c = blockIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < 16) reinterpret_cast<uint32_t*>(ws.ctr)[threadIdx.x] = 0;
    if (c >= K || threadIdx.x != 0) return;
    ColPlan p = {};
    p.status = OZL_ERR_CORRUPTION;
    const uint8_t* fr = frames + foff[c];
    const uint64_t flen = foff[c + 1] - foff[c];
    const uint64_t outBytes = doff[2 * c + 1];
    uint8_t* out = dst + doff[2 * c];
    ws.plans[c] = p;
    if (flen < 16) return;
    uint64_t sz = 0;
    int q = 6;
    for (int sh = 0; q < 16; q++, sh += 7) {
        sz |= (uint64_t)(fr[q] & 127) << sh;
        if (!(fr[q] & 128)) break;
    }
    if (q >= 16 || sz == 0) return;
    const uint64_t bytes = sz - 1;
    if (bytes != outBytes) { ws.plans[c].status = OZL_ERR_DSTCAPACITY_TOOSMALL; return; }
    uint8_t hb[kHdrSlot];
    for (int lg = 0; lg < 4; lg++) {
        const int w = 1 << lg;
        if (bytes == 0 || bytes % (uint64_t)w) continue;
        if (((uintptr_t)out) & (uintptr_t)(w - 1)) continue;
        const uint64_t n = bytes / (uint64_t)w, m = n - 1;
        const size_t fh = ozl_write_frame_header(hb, OZL_T_NUMERIC, bytes, 1, 1);
        if (fh + 9 > flen) continue;
        bool same = true;
        for (size_t k = 0; k < fh; k++) same &= hb[k] == fr[k];
        if (!same) continue;
        const int kind = bytes < kMinStream ? kTinyIn : m * (uint64_t)w < kMinStream ? kTinyDz : kBitpack;
        const int nbLo = kind == kBitpack ? 1 : 0, nbHi = kind == kBitpack ? 8 * w : 0;
        for (int nb = nbLo; nb <= nbHi; nb++) {
            const uint64_t stored = kind == kTinyIn ? bytes : kind == kTinyDz ? m * (uint64_t)w
                                                                               : (m * (uint64_t)nb + 7) / 8;
            // length with a placeholder first value, then the real one from the frame
            const size_t ch = col_chunk_header(hb + fh, kind, w, nb, m, stored, 0);
            if (fh + ch + stored + 9 != flen || fh + ch > kHdrSlot) continue;
            uint64_t first = 0;
            if (kind != kTinyIn)
                for (int k = 0; k < w; k++) first |= (uint64_t)fr[fh + ch - w + k] << (8 * k);
            col_chunk_header(hb + fh, kind, w, nb, m, stored, first);
            bool eq = fr[flen - 1] == 0;
            for (size_t k = fh; k < fh + ch; k++) eq &= hb[k] == fr[k];
            if (!eq) continue;
            const uint8_t* body = fr + fh + ch;
            p.status = 0;
            p.kind = (uint8_t)kind;
            p.w = (uint8_t)w;
            p.nb = (uint8_t)nb;
            p.m = m;
            p.packed = stored;
            p.maxZ = first;
            p.fhLen = (uint8_t)fh;
            p.hdrLen = (uint8_t)(fh + ch);
            p.src = (uint64_t)(uintptr_t)body;
            const uint8_t* cs = body + stored;
            p.hashes[0] = (uint32_t)cs[0] | ((uint32_t)cs[1] << 8) | ((uint32_t)cs[2] << 16) | ((uint32_t)cs[3] << 24);
            p.hashes[1] = (uint32_t)cs[4] | ((uint32_t)cs[5] << 8) | ((uint32_t)cs[6] << 16) | ((uint32_t)cs[7] << 24);
            if (kind == kTinyIn) {
                for (uint64_t k = 0; k < bytes; k++) out[k] = body[k];
            } else if (kind == kTinyDz) {
                uint64_t v = first;
                for (int k = 0; k < w; k++) out[k] = (uint8_t)(v >> (8 * k));
                for (uint64_t d = 0; d < m; d++) {
                    uint64_t dv = 0;
                    for (int k = 0; k < w; k++) dv |= (uint64_t)body[d * w + k] << (8 * k);
                    v = (v + dv) & wmask(w);
                    for (int k = 0; k < w; k++) out[(d + 1) * w + k] = (uint8_t)(v >> (8 * k));
                }
            } else {
                const uint32_t tmax = w <= 4 ? kTileMax : kTileWide;
                const uint32_t t = tile < tmax ? tile : tmax;
                p.tile = t;
                p.nItems = (uint32_t)((m + t - 1) / t);
            }
            ws.plans[c] = p;
            return;
        }
    }
}

// The packed words of an item (the stream bytes from bit d0*nb, a word
// boundary) go to LDS with coalesced loads; one word more than the item's
// bits is read (a value may straddle into it), which the frame's 9 trailing
// bytes always back.
// LDS slot of word i: one pad word after every 32, so lanes that read runs
// of words at a fixed stride (16 deltas per lane) fall into different banks
__device__ inline uint32_t sk(uint32_t i) { return i + (i >> 5); }
// packed words of an item (<= 64 bits per delta), later its staged output
// (<= 16 KiB), skewed
constexpr uint32_t kWdWords = 2 * kTileWide + 4 + (2 * kTileWide + 4) / 32 + 2;

__device__ inline uint32_t load_item_words(const uint8_t* stream, uint64_t d0, uint32_t cnt, int nb, uint32_t* Wd)
{
    const uint8_t* sp = stream + d0 * (uint64_t)nb / 8;
    const uint32_t nW = (uint32_t)(((uint64_t)cnt * nb + 31) / 32) + 1;
    for (uint32_t k = threadIdx.x; k < nW; k += kCT) Wd[sk(k)] = ozl_gld4(sp + 4 * (size_t)k);
    return nW;
}

// zigzag value of delta e of the item (bits [e*nb, e*nb+nb) of the words)
__device__ inline uint64_t item_value(const uint32_t* Wd, uint32_t e, int nb)
{
    const uint64_t bit = (uint64_t)e * nb;
    const uint32_t q = (uint32_t)(bit >> 5), sh = (uint32_t)(bit & 31);
    uint64_t v = ((uint64_t)Wd[sk(q)] | ((uint64_t)Wd[sk(q + 1)] << 32)) >> sh;
    if (sh && nb + (int)sh > 64) v |= (uint64_t)Wd[sk(q + 2)] << (64 - sh);
    return nb == 64 ? v : v & ((1ull << nb) - 1);
}

// Narrow columns (width <= 4, so nbBits <= 32): value e is one funnel shift
// of two skewed words, and sums and zigzag run in 32 bits (sums are only ever
// used modulo 2^(8w)).
__device__ inline uint32_t item_value32(const uint32_t* Wd, uint32_t e, int nb)
{
    const uint32_t bit = e * (uint32_t)nb;
    const uint32_t q = bit >> 5, sh = bit & 31;
    const uint32_t v = __builtin_amdgcn_alignbit(Wd[sk(q + 1)], Wd[sk(q)], sh);
    return nb == 32 ? v : v & ((1u << nb) - 1u);
}
__device__ inline uint32_t unzz32(uint32_t z) { return (z >> 1) ^ (0u - (z & 1u)); }

// tile sums of the deltas (mod 2^64; reduced mod 2^(8w) when used)
template <bool NARROW>
__global__ void __launch_bounds__(kCT) k_cold_sum(uint32_t K, ColWs ws)
{
    constexpr int KP = per_of<NARROW>();
    __shared__ uint32_t sItem, sCol;
    __shared__ uint64_t lds[kCT / 64];
    __shared__ __attribute__((aligned(16))) uint32_t Wd[kWdWords];
    if (!NARROW && ws.ctr->nWide == 0) return;
    const uint32_t total = ws.ctr->totalItems;
    const uint32_t g = grab_size(total);
    uint32_t cur = 0, end = 0;
    for (;;) {
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)next_item(&ws.ctr->queue[NARROW ? 0 : 4], total, g, &sItem, cur, end));
        if (it == ~0u) break;
        if (threadIdx.x == 0) sCol = item_col(ws)[it];
        __syncthreads();
        const ColPlan& P = ws.plans[(uint32_t)__builtin_amdgcn_readfirstlane((int)sCol)];
        if ((P.w <= 4) != NARROW) continue;
        const int w = P.w, nb = P.nb;
        const uint64_t d0 = (uint64_t)(it - P.itemBase) * P.tile, d1 = d0 + P.tile < P.m ? d0 + P.tile : P.m;
        const uint32_t cnt = (uint32_t)(d1 - d0);
        load_item_words(reinterpret_cast<const uint8_t*>((uintptr_t)P.src), d0, cnt, nb, Wd);
        __syncthreads();
        uint64_t s = 0;
        if constexpr (NARROW) {
            uint32_t s32 = 0;
#pragma unroll
            for (int k = 0; k < KP; k++) {
                const uint32_t e = threadIdx.x + kCT * k;
                if (e < cnt) s32 += unzz32(item_value32(Wd, e, nb));
            }
            s = s32;
        } else {
#pragma unroll
            for (int k = 0; k < KP; k++) {
                const uint32_t e = threadIdx.x + kCT * k;
                if (e < cnt) s += unzz(item_value(Wd, e, nb), w);
            }
        }
        s = block_sum<uint64_t>(s, lds);
        if (threadIdx.x == 0) tile_sums(ws, total)[it] = s;
    }
}

// per column: exclusive scan of its tile sums seeded with the first element
__global__ void __launch_bounds__(64) k_cold_carry(uint32_t K, ColWs ws)
{
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    const ColPlan p = ws.plans[c];
    if (p.status || ws.ctr->error) return;
    const int lane = threadIdx.x;
    uint64_t* sums = tile_sums(ws, ws.ctr->totalItems);
    uint64_t carry = p.maxZ;
    for (uint32_t t0 = 0; t0 < p.nItems; t0 += 64) {
        const uint32_t t = t0 + lane;
        const uint64_t v = t < p.nItems ? sums[p.itemBase + t] : 0;
        uint64_t x = v;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (t < p.nItems) sums[p.itemBase + t] = carry + x - v;
        carry += __shfl(x, 63, 64);
    }
}

// unpack, prefix sum within the item (thread t owns deltas [16t, 16t+16)),
// add the item's carry; the elements are staged in LDS and leave in
// consecutive dwords
template <bool NARROW>
__global__ void __launch_bounds__(kCT) k_cold_apply(uint32_t K, ColWs ws, uint8_t* __restrict__ dst,
                                                   const uint64_t* __restrict__ doff)
{
    constexpr int KP = per_of<NARROW>();
    __shared__ uint32_t sItem, sCol;
    __shared__ uint64_t lds[kCT / 64];
    __shared__ __attribute__((aligned(16))) uint32_t Wd[kWdWords];
    if (!NARROW && ws.ctr->nWide == 0) return;
    const uint32_t total = ws.ctr->totalItems;
    const uint32_t g = grab_size(total);
    uint32_t cur = 0, end = 0;
    for (;;) {
        const uint32_t it = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)next_item(&ws.ctr->queue[NARROW ? 1 : 5], total, g, &sItem, cur, end));
        if (it == ~0u) break;
        if (threadIdx.x == 0) sCol = item_col(ws)[it];
        __syncthreads();
        const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)sCol);
        const ColPlan& P = ws.plans[c];
        if ((P.w <= 4) != NARROW) continue;
        const int w = P.w, nb = P.nb;
        const uint32_t t = it - P.itemBase;
        const uint64_t d0 = (uint64_t)t * P.tile, d1 = d0 + P.tile < P.m ? d0 + P.tile : P.m;
        const uint32_t cnt = (uint32_t)(d1 - d0);
        uint8_t* out = dst + doff[2 * c];
        const uint64_t mask = wmask(w);
        if (t == 0 && threadIdx.x == 0)
            for (int k = 0; k < w; k++) out[k] = (uint8_t)(P.maxZ >> (8 * k));
        load_item_words(reinterpret_cast<const uint8_t*>((uintptr_t)P.src), d0, cnt, nb, Wd);
        __syncthreads();
        uint64_t v[KP];
        uint64_t pre;
        const uint32_t e0 = KP * threadIdx.x;
        const uint64_t carry = tile_sums(ws, total)[it];
        if constexpr (NARROW) {
            uint32_t run = 0;
#pragma unroll
            for (int k = 0; k < KP; k++) {
                const uint32_t e = e0 + k;
                run += e < cnt ? unzz32(item_value32(Wd, e, nb)) : 0u;
                v[k] = run;
            }
            uint32_t tot;
            pre = (uint32_t)(block_excl<uint32_t>(run, reinterpret_cast<uint32_t*>(lds), tot) + (uint32_t)carry);
        } else {
            uint64_t run = 0;
#pragma unroll
            for (int k = 0; k < KP; k++) {
                const uint32_t e = e0 + k;
                run += e < cnt ? unzz(item_value(Wd, e, nb), w) : 0;
                v[k] = run;
            }
            uint64_t tot;
            pre = block_excl<uint64_t>(run, lds, tot) + carry;
        }
        __syncthreads();  // every thread is done with the words
        // staged in skewed words (byte a at word sk(a / 4), byte a % 4)
        uint8_t* xb = reinterpret_cast<uint8_t*>(Wd);
#pragma unroll
        for (int k = 0; k < KP; k++) {
            const uint32_t e = e0 + k;
            if (e < cnt) {
                const uint64_t x = (pre + v[k]) & mask;
                const uint32_t a = e * (uint32_t)w;
                switch (w) {
                case 1: xb[4 * sk(a >> 2) + (a & 3)] = (uint8_t)x; break;
                case 2: reinterpret_cast<uint16_t*>(xb + 4 * sk(a >> 2))[(a >> 1) & 1] = (uint16_t)x; break;
                case 4: Wd[sk(a >> 2)] = (uint32_t)x; break;
                default:
                    Wd[sk(a >> 2)] = (uint32_t)x;
                    Wd[sk((a >> 2) + 1)] = (uint32_t)(x >> 32);
                    break;
                }
            }
        }
        __syncthreads();
        uint8_t* o = out + (d0 + 1) * (uint64_t)w;
        const uint32_t nbytes = cnt * (uint32_t)w, nd = nbytes / 4;
        for (uint32_t k = threadIdx.x; k < nd; k += kCT) ozl_gst4(o + 4 * (size_t)k, Wd[sk(k)]);
        for (uint32_t k = 4 * nd + threadIdx.x; k < nbytes; k += kCT) ozl_gst1(o + k, xb[4 * sk(k >> 2) + (k & 3)]);
    }
}

__global__ void __launch_bounds__(kCT) k_cold_verify(uint32_t K, ColWs ws, int* __restrict__ status)
{
    const uint32_t c = blockIdx.x * kCT + threadIdx.x;
    if (c >= K) return;
    const ColPlan& p = ws.plans[c];
    int st = ws.ctr->error ? (int)OZL_ERR_PARAMETER_INVALID : (int)p.status;
    if (st == 0) {
        if ((uint32_t)ws.hout[K + c] != p.hashes[1]) st = OZL_ERR_COMPRESSED_CHECKSUM;
        else if ((uint32_t)ws.hout[c] != p.hashes[0]) st = OZL_ERR_CONTENT_CHECKSUM;
    }
    status[c] = st;
}

static int persistent_grid(uint32_t wgPerCu)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256 * 4;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    }
    return cus * (int)(wgPerCu ? wgPerCu : 4);
}

// one-wave workgroups of the hash sums: 16 per CU
static int persistent_grid_hash() { return persistent_grid(4) * 4; }

// one hash pass over the R regions the caller's kernel has set
static int hash_pass(uint32_t R, const ColWs& ws, int q, hipStream_t st)
{
    OZL_LAUNCH("k_hash_scan", k_hash_scan, dim3(1), dim3(kCT), 0, st, R, ws);
    OZL_LAUNCH("k_hash_map", k_hash_map, dim3(R), dim3(64), 0, st, R, ws);
    OZL_LAUNCH("k_hash_sums", k_hash_sums, dim3(persistent_grid_hash()), dim3(64), 0, st, ws, q);
    OZL_LAUNCH("k_hash_chain", k_hash_chain, dim3(R), dim3(64), 0, st, R, ws);
    return OZL_OK;
}

// ---------------------------------------------------------------- autotuner
// (K, total bytes) -> best (tile, workgroups per CU), measured on the caller's
// batch with HIP events over the whole encode call
struct TuneKey {
    uint32_t K;
    uint64_t bytes;
    bool operator<(const TuneKey& o) const { return std::tie(K, bytes) < std::tie(o.K, o.bytes); }
};
std::mutex g_tuneMu;
std::map<TuneKey, std::pair<uint32_t, uint32_t>> g_tune;

}  // namespace

extern "C" {

// Workspace bytes for K columns of `totalBytes` input bytes in all, cut into
// items of `tile` deltas: per column the plan, headers and hash regions; per
// item its map entry and the decoder's tile sum; per hashed 1 KiB block its
// 64 bytes of stripe sums.
size_t ozl_columns_workspace(uint32_t K, size_t totalBytes, uint32_t tile)
{
    if (tile < 32) tile = 32;
    const size_t items = totalBytes / tile + K + 1;  // deltas <= bytes; one partial item per column
    // hashed bytes: columns and frames (frames <= columns + 128 B each), per pass
    const size_t hblocks = (2 * totalBytes + (size_t)128 * K) / 1024 + 2 * (size_t)K + 1;
    const size_t hitems = hblocks / kHG + 2 * (size_t)K + 1;
    return fixed_bytes(K) + items * 12 + 64 + hitems * 4 + 64 + hblocks * 64 + 256;
}

// Upper bound of the frame of a column of `bytes` input bytes.
size_t ozl_columns_frame_bound(size_t bytes) { return bytes + 128; }

// Column c of a batch is colSpan[2c] (byte offset into src, a multiple of its
// width) and colSpan[2c+1] (its bytes, a multiple of the width, > 0); on
// decode, dstSpan gives the same pairs for the outputs. Frame c is
// dst[frameOff[c], frameOff[c+1]).

int ozl_columns_compress_batch(const void* src, const uint64_t* colOff, const uint8_t* widths, uint32_t K, void* dst,
                               size_t dstCap, uint64_t* frameOff, int* status, void* workspace, size_t wsBytes,
                               uint32_t tile, uint32_t wgPerCu, void* stream)
{
    if (K == 0) return OZL_OK;
    if (!src || !colOff || !widths || !dst || !frameOff || !status || !workspace) return OZL_ERR_PARAMETER_INVALID;
    if (tile < 32 || (tile & 31) || wsBytes < fixed_bytes(K) + 1024)
        return OZL_ERR_PARAMETER_INVALID;
    hipStream_t st = (hipStream_t)stream;
    ColWs ws = carve(workspace, K, wsBytes);
    const uint8_t* s = (const uint8_t*)src;
    uint8_t* d = (uint8_t*)dst;
    const int G = persistent_grid(wgPerCu);
    const dim3 gk((K + kCT - 1) / kCT);
    OZL_LAUNCH("k_col_setup", k_col_setup, gk, dim3(kCT), 0, st, s, colOff, widths, K, tile, ws);
    OZL_LAUNCH("k_col_scan_items", k_col_scan_items, dim3(1), dim3(kCT), 0, st, K, ws);
    OZL_LAUNCH("k_col_map", k_col_map, dim3(K), dim3(64), 0, st, K, ws);
    OZL_LAUNCH("k_hreg_content", k_hreg_content, gk, dim3(kCT), 0, st, K, ws);
    hash_pass(K, ws, 2, st);
    OZL_LAUNCH("k_col_plan", k_col_plan, dim3(K), dim3(64), 0, st, K, ws, frameOff);
    OZL_LAUNCH("k_col_scan_frames", k_col_scan_frames, dim3(1), dim3(kCT), 0, st, K, frameOff, dstCap, ws);
    OZL_LAUNCH("k_col_pack", k_col_pack<true>, dim3(G), dim3(kCT), 0, st, K, ws, d, (const uint64_t*)frameOff);
    OZL_LAUNCH("k_col_pack_wide", k_col_pack<false>, dim3(G), dim3(kCT), 0, st, K, ws, d, (const uint64_t*)frameOff);
    OZL_LAUNCH("k_col_finish1", k_col_finish1, dim3(K), dim3(64), 0, st, K, ws, d, (const uint64_t*)frameOff);
    OZL_LAUNCH("k_hreg_frames", k_hreg_frames, gk, dim3(kCT), 0, st, K, ws, (const uint8_t*)d,
               (const uint64_t*)frameOff);
    hash_pass(K, ws, 3, st);
    OZL_LAUNCH("k_col_finish2", k_col_finish2, gk, dim3(kCT), 0, st, K, ws, d, (const uint64_t*)frameOff, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

int ozl_columns_decompress_batch(const void* frames, const uint64_t* frameOff, uint32_t K, void* dst,
                                 const uint64_t* dstOff, int* status, void* workspace, size_t wsBytes, uint32_t tile,
                                 uint32_t wgPerCu, void* stream)
{
    if (K == 0) return OZL_OK;
    if (!frames || !frameOff || !dst || !dstOff || !status || !workspace) return OZL_ERR_PARAMETER_INVALID;
    if (tile < 32 || (tile & 31) || wsBytes < fixed_bytes(K) + 1024)
        return OZL_ERR_PARAMETER_INVALID;
    hipStream_t st = (hipStream_t)stream;
    ColWs ws = carve(workspace, K, wsBytes);
    const int G = persistent_grid(wgPerCu);
    const uint8_t* f = (const uint8_t*)frames;
    uint8_t* d = (uint8_t*)dst;
    const dim3 gk((K + kCT - 1) / kCT);
    OZL_LAUNCH("k_cold_parse", k_cold_parse, dim3(K), dim3(64), 0, st, f, frameOff, K, d, dstOff, tile, ws);
    OZL_LAUNCH("k_col_scan_items", k_col_scan_items, dim3(1), dim3(kCT), 0, st, K, ws);
    OZL_LAUNCH("k_col_map", k_col_map, dim3(K), dim3(64), 0, st, K, ws);
    OZL_LAUNCH("k_cold_sum", k_cold_sum<true>, dim3(G), dim3(kCT), 0, st, K, ws);
    OZL_LAUNCH("k_cold_sum_wide", k_cold_sum<false>, dim3(G), dim3(kCT), 0, st, K, ws);
    OZL_LAUNCH("k_cold_carry", k_cold_carry, dim3(K), dim3(64), 0, st, K, ws);
    OZL_LAUNCH("k_cold_apply", k_cold_apply<true>, dim3(G), dim3(kCT), 0, st, K, ws, d, dstOff);
    OZL_LAUNCH("k_cold_apply_wide", k_cold_apply<false>, dim3(G), dim3(kCT), 0, st, K, ws, d, dstOff);
    OZL_LAUNCH("k_hreg_decode", k_hreg_decode, gk, dim3(kCT), 0, st, K, ws, f, frameOff, (const uint8_t*)d, dstOff);
    hash_pass(2 * K, ws, 2, st);
    OZL_LAUNCH("k_cold_verify", k_cold_verify, gk, dim3(kCT), 0, st, K, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Autotuner: times ozl_columns_compress_batch on this batch for every
// (tile, workgroups per CU) candidate and returns the fastest; the choice is
// cached per (K, total bytes) for the process. The workspace must be sized
// for the smallest candidate tile (1024).
int ozl_columns_autotune(const void* src, const uint64_t* colOff, const uint8_t* widths, uint32_t K, uint64_t totalBytes,
                         void* dst, size_t dstCap, uint64_t* frameOff, int* status, void* workspace, size_t wsBytes,
                         uint32_t* tileOut, uint32_t* wgOut, void* stream)
{
    if (!tileOut || !wgOut) return OZL_ERR_PARAMETER_INVALID;
    {
        std::lock_guard<std::mutex> g(g_tuneMu);
        auto itc = g_tune.find(TuneKey{ K, totalBytes });
        if (itc != g_tune.end()) {
            *tileOut = itc->second.first;
            *wgOut = itc->second.second;
            return OZL_OK;
        }
    }
    hipStream_t st = (hipStream_t)stream;
    static const uint32_t tiles[] = { 1024, 2048, 4096 };
    static const uint32_t wgs[] = { 2, 4, 8 };
    hipEvent_t a, b;
    OZL_HIP_CHECK(hipEventCreate(&a));
    OZL_HIP_CHECK(hipEventCreate(&b));
    float best = 1e30f;
    uint32_t bt = 4096, bw = 4;
    int rc = OZL_OK;
    for (uint32_t t : tiles) {
        if (ozl_columns_workspace(K, totalBytes, t) > wsBytes) continue;
        for (uint32_t wg : wgs) {
            // one warm call, then the timed one
            rc = ozl_columns_compress_batch(src, colOff, widths, K, dst, dstCap, frameOff, status, workspace, wsBytes, t,
                                            wg, stream);
            if (rc) break;
            if (hipEventRecord(a, st) != hipSuccess) { rc = OZL_ERR_HIP; break; }
            rc = ozl_columns_compress_batch(src, colOff, widths, K, dst, dstCap, frameOff, status, workspace, wsBytes, t,
                                            wg, stream);
            if (rc) break;
            if (hipEventRecord(b, st) != hipSuccess || hipEventSynchronize(b) != hipSuccess) { rc = OZL_ERR_HIP; break; }
            float ms = 0;
            if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { rc = OZL_ERR_HIP; break; }
            if (ms < best) { best = ms; bt = t; bw = wg; }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> g(g_tuneMu);
        g_tune[TuneKey{ K, totalBytes }] = { bt, bw };
    }
    *tileOut = bt;
    *wgOut = bw;
    return OZL_OK;
}

// hipGraph of one batch call: captures ozl_columns_compress_batch (dir 0) or
// ozl_columns_decompress_batch (dir 1) with these pointers on a private
// stream; ozl_graph_launch replays it on `stream`.
struct ozl_graph {
    hipGraph_t g;
    hipGraphExec_t x;
};

void* ozl_columns_graph_create(int dir, const void* a0, const uint64_t* a1, const uint8_t* widths, uint32_t K, void* a3,
                               size_t dstCap, uint64_t* a4, int* status, void* workspace, size_t wsBytes, uint32_t tile,
                               uint32_t wgPerCu)
{
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return nullptr;
    ozl_graph* G = new ozl_graph{};
    int rc = OZL_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = OZL_ERR_HIP;
    if (!rc) {
        if (dir == 0)
            rc = ozl_columns_compress_batch(a0, a1, widths, K, a3, dstCap, a4, status, workspace, wsBytes, tile, wgPerCu,
                                            cs);
        else
            rc = ozl_columns_decompress_batch(a0, a1, K, a3, a4, status, workspace, wsBytes, tile, wgPerCu, cs);
        if (hipStreamEndCapture(cs, &G->g) != hipSuccess) rc = OZL_ERR_HIP;
    }
    if (!rc && hipGraphInstantiate(&G->x, G->g, nullptr, nullptr, 0) != hipSuccess) rc = OZL_ERR_HIP;
    (void)hipStreamDestroy(cs);
    if (rc) {
        delete G;
        return nullptr;
    }
    return G;
}

int ozl_graph_launch(void* graph, void* stream)
{
    if (!graph) return OZL_ERR_PARAMETER_INVALID;
    OZL_HIP_CHECK(hipGraphLaunch(((ozl_graph*)graph)->x, (hipStream_t)stream));
    return OZL_OK;
}

void ozl_graph_destroy(void* graph)
{
    if (!graph) return;
    ozl_graph* G = (ozl_graph*)graph;
    (void)hipGraphExecDestroy(G->x);
    (void)hipGraphDestroy(G->g);
    delete G;
}

}  // extern "C"
