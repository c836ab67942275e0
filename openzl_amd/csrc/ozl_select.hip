// Batched exhaustive trial over eight candidate graphs, and compression of
// every chunk with the candidate chosen for it, on gfx950. Inputs are
// thousands of numeric chunks (element width 1, 2, 4 or 8; fewer than 100000
// bytes each). The candidates, in trial order:
//   0 ZL_GRAPH_STORE
//   1 ZL_GRAPH_FSE                                   (num->serial, then FSE)
//   2 ZL_GRAPH_BITPACK                               (bitpack_int)
//   3 ZL_NODE_DELTA_INT -> ZL_GRAPH_FSE
//   4 ZL_NODE_DELTA_INT -> ZL_NODE_ZIGZAG -> ZL_GRAPH_BITPACK
//   5 transpose_split -> ZL_GRAPH_FSE on every byte plane
//   6 transpose_split -> ZL_GRAPH_HUFFMAN on every byte plane
//   7 ZL_NODE_DELTA_INT -> transpose_split -> ZL_GRAPH_FSE on every plane
// The trial is the brute-force selector (selector_brute_force.c:9-50): each
// candidate's frame size in a context with checksums off (tryGraph,
// cctx.c:1799-1865); the first strictly smallest size below the input size
// wins, else ZL_GRAPH_STORE. A selector leaves nothing in the frame, so the
// selected frame is the chosen candidate's own frame.
//
// The node and stream order of each candidate follows the engine's graph
// walk (cctx.c:1116-1185 successors with the minStreamSize auto-store,
// cctx.c:534-605 implicit conversions, encode_transpose_binding.c:42-79,
// encode_delta_binding.c:11-45, encode_bitpack_binding.c:92-147, and the
// entropy graphs of encode_entropy_binding.c:601-753 through
// ozl_entropy_layout).
//
// Work per call (one stream, one host read of the block counts):
//   k_sel_setup   per chunk: bytes and entropy blocks of the streams the
//                 requested candidates need
//   k_sel_scan    staging offsets and block numbers (one workgroup)
//   k_sel_prep    per chunk: raw copy, deltas, byte planes and delta byte
//                 planes into the staging area; largest value and zigzag delta
//   the entropy coders over the staged blocks: Huffman over the planes of
//                 candidate 6, FSE over the rest
//   trial         k_sel_trial_sizes (one lane per chunk and candidate) ->
//                 k_sel_choose
//   compress      k_sel_frame_size -> k_sel_scan_frames -> k_sel_pack ->
//                 k_sel_write -> k_sel_hash (content) -> k_sel_hash
//                 (compressed) and the end byte
// Staging groups keep both coders' block lists contiguous: planes only the
// Huffman coder reads, then planes both coders read, then whole streams only
// the FSE coder reads (raw, deltas), then planes only the FSE coder reads.
// The Huffman list is the first two groups, the FSE list the last three; the
// longest FSE blocks (whole chunks) precede the planes, so the FSE coder's
// first workgroups, dispatched first, hold the longest state chains.
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_entropy_api.h"
#include "ozl_profile.h"
#include "ozl_gmem.h"

#include <string.h>

namespace {

constexpr int kNC = 8;
enum { C_STORE, C_FSE, C_BITPACK, C_DELTA_FSE, C_DELTA_ZZ_BP, C_TS_FSE, C_TS_HUF, C_DELTA_TS_FSE };
constexpr uint32_t kMinStream = OZL_MIN_STREAM_SIZE;
constexpr uint32_t kMaxChunk = 100000;
constexpr uint32_t kTinySlot = 256;   // per chunk: tiny deltas at 0, tiny planes at 16, tiny delta planes at 96
constexpr uint32_t kNoBlk = 0xFFFFFFFFu;
constexpr int kT = 256;

struct SelChunk {
    uint64_t src;
    unsigned long long maxV, maxZ;
    uint64_t first;
    uint64_t gBytes[4];      // staged bytes per group, then the exclusive prefix
    uint32_t gBlocks[4];     // entropy blocks per group, then the exclusive prefix
    uint32_t n, w, need, choice, status;
    uint64_t packBytes;      // then the pack area offset
    uint64_t stTs, stRaw, stDelta, stDts;   // staging offsets
    uint32_t bTs, bRaw, bDelta, bDts;       // global block numbers (kNoBlk: none)
    uint32_t sizes[kNC];
    uint32_t pad[2];
};

struct SelCtr {
    uint64_t gBase[4];       // staging offset of each group
    uint32_t bBase[4];       // first block of each group
    uint32_t blocks[4];
    uint64_t stgBytes, packBytes;
    uint32_t error;
    uint32_t pad[5];
};

struct SelFrame {
    uint32_t hlen, nbStored;
    uint8_t hdr[1024];       // frame header + chunk header
    uint64_t ptr[80];        // stored streams in frame order (at most the layout's stream count)
    uint32_t size[80];
};

struct SelWs {
    SelCtr* ctr;
    SelChunk* ch;
    uint64_t* off;           // block offsets in the staging area (blocks + 1)
    uint64_t* fsz;           // frame sizes (compress)
    uint8_t* tiny;
    uint8_t* stg;
    uint8_t* pack;
    uint8_t* ews;            // entropy workspaces
    struct SelFrame* fr;     // header bytes and stored streams of each chunk's frame (compress)
    size_t stgCap, packCap, ewsCap;
};

__device__ inline bool has(uint32_t need, int c) { return (need >> c) & 1u; }

__device__ inline uint64_t ld_el(const uint8_t* p, int w)
{
    switch (w) {
    case 1: return *p;
    case 2: return *reinterpret_cast<const uint16_t*>(p);
    case 4: return *reinterpret_cast<const uint32_t*>(p);
    default: return *reinterpret_cast<const uint64_t*>(p);
    }
}
__device__ inline uint64_t wmask(int w) { return w == 8 ? ~0ull : ((1ull << (8 * w)) - 1); }
__device__ inline uint64_t zz(uint64_t d, int w)
{
    const int sb = 8 * w - 1;
    const uint64_t s = (uint64_t)((int64_t)(d << (63 - sb)) >> 63);
    return ((d << 1) ^ s) & wmask(w);
}

// which staged streams chunk c needs
struct Needs {
    bool raw, delta, tsHuf, tsFse, dts, pack;
    bool small, deltaTiny, tsTiny, dtsTiny;
};
__device__ inline Needs needs_of(uint32_t need, uint32_t n, uint32_t w)
{
    Needs q;
    const uint64_t bytes = (uint64_t)n * w;
    q.small = bytes < kMinStream;
    q.deltaTiny = (uint64_t)(n - 1) * w < kMinStream;
    q.tsTiny = n < kMinStream;
    q.dtsTiny = n - 1 < kMinStream;
    q.raw = has(need, C_FSE);
    q.delta = has(need, C_DELTA_FSE) || ((has(need, C_DELTA_ZZ_BP) || has(need, C_DELTA_TS_FSE)) && q.deltaTiny);
    q.tsHuf = has(need, C_TS_HUF);
    q.tsFse = has(need, C_TS_FSE);
    q.dts = has(need, C_DELTA_TS_FSE) && !q.deltaTiny;
    q.pack = has(need, C_BITPACK) || (has(need, C_DELTA_ZZ_BP) && !q.deltaTiny);
    return q;
}

// ------------------------------------------------------------------ setup and scans
__global__ void __launch_bounds__(kT) k_sel_setup(const uint8_t* __restrict__ src, const uint64_t* __restrict__ spans,
                                                  const uint8_t* __restrict__ widths, const uint8_t* __restrict__ choice,
                                                  uint32_t K, uint64_t maxChunk, SelWs ws)
{
    const uint32_t c = blockIdx.x * kT + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x < sizeof(SelCtr) / 4) reinterpret_cast<uint32_t*>(ws.ctr)[threadIdx.x] = 0;
    if (c >= K) return;
    SelChunk h;
    memset(&h, 0, sizeof(h));
    const uint64_t a = spans[2 * c], bytes = spans[2 * c + 1];
    const uint32_t w = widths[c];
    h.src = (uint64_t)(uintptr_t)(src + a);
    h.w = w;
    h.bTs = h.bRaw = h.bDelta = h.bDts = kNoBlk;
    const bool wok = w == 1 || w == 2 || w == 4 || w == 8;
    if (!wok || bytes == 0 || bytes > maxChunk || bytes % (wok ? w : 1) || (h.src & (uint64_t)(wok ? w - 1 : 0))) {
        h.status = OZL_ERR_NODE_INVALID_INPUT;
        ws.ch[c] = h;
        return;
    }
    h.n = (uint32_t)(bytes / w);
    h.need = choice ? (1u << (choice[c] & 7)) : 0xFFu;
    h.choice = choice ? choice[c] & 7 : 0;
    const Needs q = needs_of(h.need, h.n, w);
    if (!q.small) {
        const uint64_t n = h.n, m = n - 1;
        if (q.raw) { h.gBytes[2] += n * w; h.gBlocks[2] += 1; }
        if (q.delta && !q.deltaTiny) { h.gBytes[2] += m * w; h.gBlocks[2] += 1; }
        if ((q.tsHuf || q.tsFse) && !q.tsTiny) {
            const int g = q.tsHuf && q.tsFse ? 1 : q.tsHuf ? 0 : 3;
            h.gBytes[g] += n * w;
            h.gBlocks[g] += w;
        }
        if (q.dts && !q.dtsTiny) { h.gBytes[3] += m * w; h.gBlocks[3] += w; }
        if (q.pack) h.packBytes = n * w + 16;
    }
    ws.ch[c] = h;
}

template <typename T>
__device__ T wg_excl(T v, T* lds, T& total)
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
    for (int k = 0; k < kT / 64; k++) {
        const T s = lds[k];
        if (k < wave) off += s;
        all += s;
    }
    total = all;
    return off + x - v;
}

// exclusive prefixes of the group bytes, group blocks and pack bytes, then
// the group bases and every chunk's staging offsets and block numbers. Thread
// t sums a contiguous range of chunks, one workgroup scan of the seven
// totals, then each thread writes its range's prefixes.
__global__ void __launch_bounds__(kT) k_sel_scan(uint32_t K, SelWs ws)
{
    __shared__ uint64_t l64[kT / 64];
    __shared__ uint32_t l32[kT / 64];
    const uint32_t per = (K + kT - 1) / kT;
    const uint32_t c0 = threadIdx.x * per, c1 = c0 + per < K ? c0 + per : K;
    uint64_t sb[4] = { 0, 0, 0, 0 }, sp = 0;
    uint32_t sbl[4] = { 0, 0, 0, 0 };
    for (uint32_t c = c0; c < c1; c++) {
        const SelChunk& h = ws.ch[c];
        for (int g = 0; g < 4; g++) { sb[g] += h.gBytes[g]; sbl[g] += h.gBlocks[g]; }
        sp += h.packBytes;
    }
    uint64_t cb[4], cp;
    uint32_t cbl[4];
    uint64_t eb[4], ep;
    uint32_t ebl[4];
    for (int g = 0; g < 4; g++) {
        eb[g] = wg_excl<uint64_t>(sb[g], l64, cb[g]);
        ebl[g] = wg_excl<uint32_t>(sbl[g], l32, cbl[g]);
    }
    ep = wg_excl<uint64_t>(sp, l64, cp);
    for (uint32_t c = c0; c < c1; c++) {
        SelChunk& h = ws.ch[c];
        for (int g = 0; g < 4; g++) {
            const uint64_t b = h.gBytes[g];
            const uint32_t bl = h.gBlocks[g];
            h.gBytes[g] = eb[g];
            h.gBlocks[g] = ebl[g];
            eb[g] += b;
            ebl[g] += bl;
        }
        const uint64_t pb = h.packBytes;
        h.packBytes = ep;
        ep += pb;
    }
    __syncthreads();
    const uint64_t gBase[4] = { 0, cb[0], cb[0] + cb[1], cb[0] + cb[1] + cb[2] };
    const uint32_t bBase[4] = { 0, cbl[0], cbl[0] + cbl[1], cbl[0] + cbl[1] + cbl[2] };
    if (threadIdx.x == 0) {
        SelCtr* t = ws.ctr;
        for (int g = 0; g < 4; g++) { t->gBase[g] = gBase[g]; t->bBase[g] = bBase[g]; t->blocks[g] = cbl[g]; }
        t->stgBytes = cb[0] + cb[1] + cb[2] + cb[3];
        t->packBytes = cp;
        if (t->stgBytes > ws.stgCap || cp > ws.packCap) t->error = 1;
        ws.off[bBase[3] + cbl[3]] = t->stgBytes;
    }
    for (uint32_t c = threadIdx.x; c < K; c += kT) {
        SelChunk& h = ws.ch[c];
        if (h.status) continue;
        const Needs q = needs_of(h.need, h.n, h.w);
        if (q.small) continue;
        const uint64_t n = h.n;
        const uint32_t w = h.w;
        if ((q.tsHuf || q.tsFse) && !q.tsTiny) {
            const int g = q.tsHuf && q.tsFse ? 1 : q.tsHuf ? 0 : 3;
            h.stTs = gBase[g] + h.gBytes[g];
            h.bTs = bBase[g] + h.gBlocks[g];
        }
        // group 2 per chunk: raw, delta; group 3: FSE-only planes, delta planes
        uint64_t s = gBase[2] + h.gBytes[2];
        uint32_t b = bBase[2] + h.gBlocks[2];
        if (q.raw) { h.stRaw = s; h.bRaw = b; s += n * w; b += 1; }
        if (q.delta && !q.deltaTiny) { h.stDelta = s; h.bDelta = b; }
        s = gBase[3] + h.gBytes[3];
        b = bBase[3] + h.gBlocks[3];
        if (q.tsFse && !q.tsHuf && !q.tsTiny) { s += n * w; b += w; }
        if (q.dts && !q.dtsTiny) { h.stDts = s; h.bDts = b; }
    }
}

// ------------------------------------------------------------------ staging
// Thread t takes the quads of elements [4v, 4v + 4) for v = t, t + 256, ...:
// raw copy and deltas as whole elements, each byte plane's 4 bytes as one
// dword (consecutive threads, consecutive dwords); elements past the last
// quad one at a time.
template <int W>
__device__ void prep_w(const uint8_t* x, uint32_t n, uint8_t* raw, uint8_t* dl, uint8_t* ts, uint8_t* dts,
                       unsigned long long& mv, unsigned long long& mz)
{
    const uint32_t m = n - 1, nq = n / 4;
    const uint64_t msk = wmask(W);
    for (uint32_t v = threadIdx.x; v < nq; v += kT) {
        const uint32_t i0 = 4 * v;
        uint64_t e[5];
#pragma unroll
        for (int j = 0; j < 4; j++) e[j] = ozl_gld_w(x + (size_t)(i0 + j) * W, W);
        e[4] = i0 + 4 < n ? ozl_gld_w(x + (size_t)(i0 + 4) * W, W) : e[3];
        uint64_t d[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            mv = e[j] > mv ? e[j] : mv;
            d[j] = (e[j + 1] - e[j]) & msk;
            if (i0 + j < m) {
                const uint64_t z = zz(d[j], W);
                mz = z > mz ? z : mz;
            }
        }
        if (raw)
#pragma unroll
            for (int j = 0; j < 4; j++) ozl_gst_w(raw + (size_t)(i0 + j) * W, e[j], W);
        if (dl)
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (i0 + j < m) ozl_gst_w(dl + (size_t)(i0 + j) * W, d[j], W);
        if (ts)
#pragma unroll
            for (int p = 0; p < W; p++) {
                uint32_t q = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) q |= (uint32_t)((e[j] >> (8 * p)) & 255) << (8 * j);
                ozl_gst4(ts + (size_t)p * n + i0, q);
            }
        if (dts)
#pragma unroll
            for (int p = 0; p < W; p++) {
                uint32_t q = 0;
#pragma unroll
                for (int j = 0; j < 4; j++) q |= (uint32_t)((d[j] >> (8 * p)) & 255) << (8 * j);
                if (i0 + 3 < m) {
                    ozl_gst4(dts + (size_t)p * m + i0, q);
                } else {
                    for (int j = 0; j < 4; j++)
                        if (i0 + j < m) ozl_gst1(dts + (size_t)p * m + i0 + j, q >> (8 * j));
                }
            }
    }
    for (uint32_t i = 4 * nq + threadIdx.x; i < n; i += kT) {
        const uint64_t v = ozl_gld_w(x + (size_t)i * W, W);
        mv = v > mv ? v : mv;
        if (raw) ozl_gst_w(raw + (size_t)i * W, v, W);
        if (ts)
            for (int k = 0; k < W; k++) ozl_gst1(ts + (size_t)k * n + i, (uint32_t)(v >> (8 * k)));
        if (i < m) {
            const uint64_t d = (ozl_gld_w(x + (size_t)(i + 1) * W, W) - v) & msk;
            const uint64_t z = zz(d, W);
            mz = z > mz ? z : mz;
            if (dl) ozl_gst_w(dl + (size_t)i * W, d, W);
            if (dts)
                for (int k = 0; k < W; k++) ozl_gst1(dts + (size_t)k * m + i, (uint32_t)(d >> (8 * k)));
        }
    }
}

// one workgroup per chunk; element i and its delta d[i] = x[i+1] - x[i]
__global__ void __launch_bounds__(kT) k_sel_prep(uint32_t K, SelWs ws)
{
    const uint32_t c = blockIdx.x;
    if (c >= K || ws.ctr->error) return;
    SelChunk& H = ws.ch[c];
    if (H.status) return;
    const uint32_t n = H.n, w = H.w, m = n - 1;
    const Needs q = needs_of(H.need, n, w);
    const uint8_t* x = reinterpret_cast<const uint8_t*>((uintptr_t)H.src);
    uint8_t* tiny = ws.tiny + (size_t)c * kTinySlot;
    if (threadIdx.x == 0) H.first = ld_el(x, (int)w);
    if (q.small) return;
    uint8_t* raw = q.raw ? ws.stg + H.stRaw : nullptr;
    uint8_t* dl = q.delta ? (q.deltaTiny ? tiny : ws.stg + H.stDelta) : nullptr;
    uint8_t* ts = (q.tsHuf || q.tsFse) ? (q.tsTiny ? tiny + 16 : ws.stg + H.stTs) : nullptr;
    uint8_t* dts = q.dts ? (q.dtsTiny ? tiny + 96 : ws.stg + H.stDts) : nullptr;
    const bool wantV = has(H.need, C_BITPACK), wantZ = has(H.need, C_DELTA_ZZ_BP) && !q.deltaTiny;
    unsigned long long mv = 0, mz = 0;
    switch (w) {
    case 1: prep_w<1>(x, n, raw, dl, ts, dts, mv, mz); break;
    case 2: prep_w<2>(x, n, raw, dl, ts, dts, mv, mz); break;
    case 4: prep_w<4>(x, n, raw, dl, ts, dts, mv, mz); break;
    default: prep_w<8>(x, n, raw, dl, ts, dts, mv, mz); break;
    }
    // block offsets
    if (threadIdx.x < w) {
        const uint32_t k = threadIdx.x;
        if (ts && !q.tsTiny) ws.off[H.bTs + k] = H.stTs + (uint64_t)k * n;
        if (dts && !q.dtsTiny) ws.off[H.bDts + k] = H.stDts + (uint64_t)k * m;
        if (k == 0 && raw) ws.off[H.bRaw] = H.stRaw;
        if (k == 0 && dl && !q.deltaTiny) ws.off[H.bDelta] = H.stDelta;
    }
    for (int s = 32; s >= 1; s >>= 1) {
        const unsigned long long o = __shfl_xor(mv, s, 64), oz = __shfl_xor(mz, s, 64);
        mv = o > mv ? o : mv;
        mz = oz > mz ? oz : mz;
    }
    if ((threadIdx.x & 63) == 0) {
        if (wantV && mv) atomicMax(&H.maxV, mv);
        if (wantZ && mz) atomicMax(&H.maxZ, mz);
    }
}

// ------------------------------------------------------------------ layouts
constexpr int kMaxN = 48, kMaxS = 80;

struct EwsPair {
    ozl_ews huf, fse;
    uint32_t hufB, fseFirst;   // Huffman blocks; first global block of the FSE list
};

struct Layout {
    ozl_fnode nd[kMaxN];
    ozl_fstream st[kMaxS];
    const uint8_t* ptr[kMaxS];
    int nn, ns, err;

    __device__ void init() { nn = ns = err = 0; }
    __device__ int stream(uint32_t size, const uint8_t* p)
    {
        if (ns >= kMaxS) { err = OZL_ERR_LOGIC; return kMaxS - 1; }
        st[ns].size = size;
        st[ns].store = 0;
        st[ns].src = 0;
        st[ns].srcOff = 0;
        ptr[ns] = p;
        return ns++;
    }
    __device__ ozl_fnode* node(uint8_t trid, int in)
    {
        if (nn >= kMaxN) { err = OZL_ERR_LOGIC; return &nd[kMaxN - 1]; }
        ozl_fnode* q = &nd[nn++];
        q->trid = trid;
        q->hdrLen = 0;
        q->nbVO = 0;
        q->in = (uint32_t)in;
        q->firstOut = (uint32_t)ns;
        return q;
    }
    // num->serial conversion of a numeric stream of width w (cctx.c:563-605)
    __device__ int to_serial(int sid, int w)
    {
        ozl_fnode* q = node(OZL_TR_CONV_NUM_TO_SERIAL_LE, sid);
        q->hdr[0] = (uint8_t)ozl_next_pow2((uint64_t)w);
        q->hdrLen = 1;
        return stream(st[sid].size, ptr[sid]);
    }
    // STORE1 -> serial_store (selector_store.c:19-36)
    __device__ void store(int sid, int w, bool numeric)
    {
        if (numeric) sid = to_serial(sid, w);
        st[sid].store = 1;
    }
    // the FSE / Huffman graph of one serial stream: its block's plan, side
    // area and payload (ozl_entropy_layout), renumbered after this layout's
    // streams
    __device__ void entropy(int sid, const ozl_ews& e, uint32_t b)
    {
        const ozl_eplan& p = e.plans[b];
        if (p.status) { err = (int)p.status; return; }
        ozl_fnode tn[8];
        ozl_fstream ts[12];
        int tns = 0;
        const int tnn = ozl_entropy_layout(&p, tn, ts, &tns);
        const int base = ns;
        const uint8_t* aux = e.aux + (size_t)b * OZL_AUX_BYTES;
        const uint8_t* pay = e.payload + (size_t)b * e.payloadCap;
        if (ts[0].store) st[sid].store = 1;
        for (int k = 1; k < tns; k++) {
            const uint8_t* from = ts[k].src == OZL_SRC_PAYLOAD ? pay : ts[k].src == OZL_SRC_AUX ? aux + ts[k].srcOff
                                                                                                : ptr[sid];
            const int id = stream(ts[k].size, from);
            st[id].store = ts[k].store;
        }
        for (int k = 0; k < tnn; k++) {
            if (nn >= kMaxN) { err = OZL_ERR_LOGIC; return; }
            ozl_fnode& q = nd[nn++];
            q = tn[k];
            q.in = tn[k].in == 0 ? (uint32_t)sid : (uint32_t)(base + tn[k].in - 1);
            q.firstOut = (uint32_t)(base + tn[k].firstOut - 1);
        }
    }
    // bitpack_int over m values of width w, largest `mx`
    // (encode_bitpack_binding.c:51-147)
    __device__ void bitpack(int sid, uint64_t m, int w, uint64_t mx, const uint8_t* packed)
    {
        static constexpr uint8_t lg[9] = { 0, 0, 1, 0, 2, 0, 0, 0, 3 };
        const int nb = 1 + (mx ? ozl_hb64(mx) : 0);
        const uint64_t size = (m * (uint64_t)nb + 7) / 8;
        ozl_fnode* q = node(OZL_TR_BITPACK_INT, sid);
        q->hdr[0] = (uint8_t)((lg[w] << 6) | (nb - 1));
        q->hdrLen = 1;
        const uint64_t maxElts = size * 8 / (uint64_t)nb;
        if (maxElts > m) { q->hdr[1] = (uint8_t)(maxElts - m); q->hdrLen = 2; }
        const int o = stream((uint32_t)size, packed);
        st[o].store = 1;
    }
};

struct LayoutCtx {
    SelWs ws;
    EwsPair e;
};

__device__ inline const ozl_ews& fse_of(const LayoutCtx& L) { return L.e.fse; }

// Builds candidate `cand`'s final graph for chunk c. Pointers of stored
// streams are filled (the packed streams at the chunk's pack area).
__device__ int build_layout(const LayoutCtx& X, uint32_t c, int cand, Layout& L)
{
    const SelChunk& H = X.ws.ch[c];
    const uint32_t n = H.n, w = H.w, m = n - 1;
    const Needs q = needs_of(1u << cand, n, w);
    const uint8_t* x = reinterpret_cast<const uint8_t*>((uintptr_t)H.src);
    const uint8_t* tiny = X.ws.tiny + (size_t)c * kTinySlot;
    L.init();
    const int in = L.stream(n * w, x);
    if (q.small) {   // below minStreamSize: stored whatever the graph
        L.store(in, (int)w, true);
        return L.err;
    }
    const uint32_t fse0 = X.e.fseFirst;
    switch (cand) {
    case C_STORE: L.store(in, (int)w, true); break;
    case C_FSE: {
        const int s = L.to_serial(in, (int)w);
        L.ptr[s] = X.ws.stg + H.stRaw;
        L.entropy(s, X.e.fse, H.bRaw - fse0);
        break;
    }
    case C_BITPACK: L.bitpack(in, n, (int)w, H.maxV, X.ws.pack + H.packBytes); break;
    case C_DELTA_FSE:
    case C_DELTA_ZZ_BP:
    case C_DELTA_TS_FSE: {
        ozl_fnode* d = L.node(OZL_TR_DELTA_INT, in);
        for (uint32_t k = 0; k < w; k++) d->hdr[k] = (uint8_t)(H.first >> (8 * k));
        d->hdrLen = (uint8_t)w;
        const uint8_t* dp = q.deltaTiny ? tiny : cand == C_DELTA_FSE ? X.ws.stg + H.stDelta : nullptr;
        const int ds = L.stream(m * w, dp);
        if (q.deltaTiny) { L.store(ds, (int)w, true); break; }
        if (cand == C_DELTA_FSE) {
            const int s = L.to_serial(ds, (int)w);
            L.entropy(s, X.e.fse, H.bDelta - fse0);
        } else if (cand == C_DELTA_ZZ_BP) {
            L.node(OZL_TR_ZIGZAG, ds);
            const int zs = L.stream(m * w, nullptr);
            L.bitpack(zs, m, (int)w, H.maxZ, X.ws.pack + H.packBytes);
        } else {
            L.node(OZL_TR_CONV_NUM_TO_STRUCT_LE, ds);
            const int ss = L.stream(m * w, nullptr);
            ozl_fnode* t = L.node(OZL_TR_TRANSPOSE_SPLIT, ss);
            t->nbVO = (uint16_t)w;
            const uint8_t* pl = q.dtsTiny ? tiny + 96 : X.ws.stg + H.stDts;
            const int p0 = L.ns;
            for (uint32_t k = 0; k < w; k++) L.stream(m, pl + (size_t)k * m);
            for (uint32_t k = 0; k < w; k++) {
                if (q.dtsTiny) L.st[p0 + k].store = 1;
                else L.entropy(p0 + (int)k, X.e.fse, H.bDts + k - fse0);
            }
        }
        break;
    }
    case C_TS_FSE:
    case C_TS_HUF: {
        L.node(OZL_TR_CONV_NUM_TO_STRUCT_LE, in);
        const int ss = L.stream(n * w, x);
        ozl_fnode* t = L.node(OZL_TR_TRANSPOSE_SPLIT, ss);
        t->nbVO = (uint16_t)w;
        const uint8_t* pl = q.tsTiny ? tiny + 16 : X.ws.stg + H.stTs;
        const int p0 = L.ns;
        for (uint32_t k = 0; k < w; k++) L.stream(n, pl + (size_t)k * n);
        for (uint32_t k = 0; k < w; k++) {
            if (q.tsTiny) L.st[p0 + k].store = 1;
            else if (cand == C_TS_HUF) L.entropy(p0 + (int)k, X.e.huf, H.bTs + k);
            else L.entropy(p0 + (int)k, X.e.fse, H.bTs + k - fse0);
        }
        break;
    }
    default: return OZL_ERR_LOGIC;
    }
    return L.err;
}

// header bytes of a layout (frame header, chunk header) into `hdr`; returns
// their count, the stored-stream order in `order`
__device__ size_t layout_header(const Layout& L, uint32_t bytes, bool checksums, uint8_t* hdr, int* order,
                                int* nbStored)
{
    const size_t fh = ozl_write_frame_header(hdr, OZL_T_NUMERIC, bytes, checksums, checksums);
    return fh + ozl_write_chunk_header(hdr + fh, L.nd, L.nn, L.st, L.ns, order, nbStored);
}

__device__ inline uint64_t layout_stored(const Layout& L, const int* order, int nbStored)
{
    uint64_t t = 0;
    for (int k = 0; k < nbStored; k++) t += L.st[order[k]].size;
    return t;
}

// ------------------------------------------------------------------ trial
__global__ void __launch_bounds__(64) k_sel_trial_sizes(uint32_t K, LayoutCtx X)
{
    const uint32_t t = blockIdx.x * 64 + threadIdx.x;
    const uint32_t c = t / kNC;
    const int cand = (int)(t % kNC);
    if (c >= K) return;
    SelChunk& H = X.ws.ch[c];
    if (H.status || X.ws.ctr->error) { H.sizes[cand] = 0; return; }
    Layout L;
    uint8_t hdr[1024];
    int order[kMaxS], nbStored = 0;
    if (build_layout(X, c, cand, L)) { H.sizes[cand] = 0; return; }
    const size_t hl = layout_header(L, H.n * H.w, false, hdr, order, &nbStored);
    H.sizes[cand] = (uint32_t)(hl + layout_stored(L, order, nbStored) + 1);
}

__global__ void __launch_bounds__(kT) k_sel_choose(uint32_t K, SelWs ws, uint32_t* __restrict__ sizes,
                                                   uint8_t* __restrict__ choice)
{
    const uint32_t c = blockIdx.x * kT + threadIdx.x;
    if (c >= K) return;
    const SelChunk& H = ws.ch[c];
    uint64_t best = (uint64_t)H.n * H.w;
    int bi = C_STORE;
    for (int k = 0; k < kNC; k++) {
        const uint32_t s = H.status || ws.ctr->error ? 0u : H.sizes[k];
        if (sizes) sizes[(size_t)c * kNC + k] = s;
        if (s && s < best) { best = s; bi = k; }
    }
    if (choice) choice[c] = (uint8_t)bi;
}

// ------------------------------------------------------------------ compress
__global__ void __launch_bounds__(64) k_sel_frame_size(uint32_t K, LayoutCtx X)
{
    const uint32_t c = blockIdx.x * 64 + threadIdx.x;
    if (c >= K) return;
    SelChunk& H = X.ws.ch[c];
    if (H.status || X.ws.ctr->error) { X.ws.fsz[c] = 0; return; }
    Layout L;
    int order[kMaxS], nbStored = 0;
    SelFrame& F = X.ws.fr[c];
    const int rc = build_layout(X, c, (int)H.choice, L);
    if (rc) { H.status = (uint32_t)rc; X.ws.fsz[c] = 0; return; }
    const size_t hl = layout_header(L, H.n * H.w, true, F.hdr, order, &nbStored);
    uint64_t stored = 0;
    for (int k = 0; k < nbStored; k++) {
        const int s = order[k];
        if (!L.ptr[s] && L.st[s].size) { H.status = OZL_ERR_LOGIC; X.ws.fsz[c] = 0; return; }
        F.ptr[k] = (uint64_t)(uintptr_t)L.ptr[s];
        F.size[k] = L.st[s].size;
        stored += L.st[s].size;
    }
    F.hlen = (uint32_t)hl;
    F.nbStored = (uint32_t)nbStored;
    X.ws.fsz[c] = hl + stored + 4 + 4 + 1;
}

// frame offsets: exclusive scan of the frame sizes (one workgroup); frames
// that would pass dstCap fail
__global__ void __launch_bounds__(kT) k_sel_scan_frames(uint32_t K, SelWs ws, uint64_t* __restrict__ frameOff,
                                                        size_t dstCap)
{
    __shared__ uint64_t lds[kT / 64];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < K; base += kT) {
        const uint32_t c = base + threadIdx.x;
        const uint64_t v = c < K ? ws.fsz[c] : 0;
        uint64_t t;
        const uint64_t x = wg_excl<uint64_t>(v, lds, t);
        if (c < K) frameOff[c] = carry + x;
        carry += t;
    }
    if (threadIdx.x == 0) {
        frameOff[K] = carry;
        if (carry > dstCap) ws.ctr->error |= 2;
    }
}

// LSB-first packing (common_bitpack_kernel.c:131-283): output dword j holds
// bits [32j, 32j+32) of the value stream; nbBits == 8*width is a copy
__global__ void __launch_bounds__(kT) k_sel_pack(uint32_t K, SelWs ws)
{
    const uint32_t c = blockIdx.x;
    if (c >= K || ws.ctr->error) return;
    const SelChunk& H = ws.ch[c];
    if (H.status) return;
    const int cand = (int)H.choice;
    const uint32_t n = H.n, w = H.w;
    const Needs q = needs_of(H.need, n, w);
    if (q.small || !q.pack) return;
    const bool dz = cand == C_DELTA_ZZ_BP;
    const uint64_t m = dz ? n - 1 : n;
    const uint64_t mx = dz ? H.maxZ : H.maxV;
    const int nb = 1 + (mx ? ozl_hb64(mx) : 0);
    const uint8_t* x = reinterpret_cast<const uint8_t*>((uintptr_t)H.src);
    uint32_t* out = reinterpret_cast<uint32_t*>(ws.pack + H.packBytes);
    const uint64_t msk = wmask((int)w);
    const uint64_t words = (m * (uint64_t)nb + 31) / 32;
    for (uint64_t j = threadIdx.x; j < words; j += kT) {
        const uint64_t b0 = 32 * j;
        const uint64_t e0 = b0 / (uint64_t)nb, e1l = (b0 + 31) / (uint64_t)nb;
        const uint64_t e1 = e1l < m ? e1l : m - 1;
        uint32_t v = 0;
        for (uint64_t e = e0; e <= e1; e++) {
            uint64_t val = ld_el(x + e * w, (int)w);
            if (dz) val = zz((ld_el(x + (e + 1) * w, (int)w) - val) & msk, (int)w);
            const long long sh = (long long)(e * (uint64_t)nb) - (long long)b0;
            v |= sh >= 0 ? (uint32_t)(val << sh) : (uint32_t)(val >> (-sh));
        }
        // the packed stream may start at any byte of the pack area
        uint8_t* o = reinterpret_cast<uint8_t*>(out) + 4 * j;
        o[0] = (uint8_t)v; o[1] = (uint8_t)(v >> 8); o[2] = (uint8_t)(v >> 16); o[3] = (uint8_t)(v >> 24);
    }
}

// header and stored streams of chunk c's frame (recorded by k_sel_frame_size)
__global__ void __launch_bounds__(kT) k_sel_write(uint32_t K, SelWs ws, uint8_t* __restrict__ dst,
                                                  const uint64_t* __restrict__ frameOff)
{
    const uint32_t c = blockIdx.x;
    if (c >= K || ws.ctr->error) return;
    if (ws.ch[c].status) return;
    const SelFrame& F = ws.fr[c];
    uint8_t* out = dst + frameOff[c];
    const uint32_t hl = F.hlen;
    for (uint32_t i = threadIdx.x; i < hl; i += kT) ozl_gst1(out + i, F.hdr[i]);
    size_t pos = hl;
    const uint32_t ns = F.nbStored;
    for (uint32_t k = 0; k < ns; k++) {
        const uint8_t* from = reinterpret_cast<const uint8_t*>((uintptr_t)F.ptr[k]);
        const uint32_t sz = F.size[k];
        uint8_t* to = out + pos;
        const uint32_t nd = sz / 4;
        for (uint32_t i = threadIdx.x; i < nd; i += kT) ozl_gst4(to + 4 * (size_t)i, ozl_gld4(from + 4 * (size_t)i));
        for (uint32_t i = 4 * nd + threadIdx.x; i < sz; i += kT) ozl_gst1(to + i, ozl_gld1(from + i));
        pos += sz;
    }
}

// content checksum (XXH3 of the chunk) into the frame, then (second call) the
// compressed checksum over [chunk start, end of content checksum) and the end
// byte; one wavefront per frame
__global__ void __launch_bounds__(64) k_sel_hash(uint32_t K, SelWs ws, uint8_t* __restrict__ dst,
                                                 const uint64_t* __restrict__ frameOff, int compressed,
                                                 int* __restrict__ status)
{
    const uint32_t c = blockIdx.x;
    if (c >= K) return;
    const SelChunk& H = ws.ch[c];
    const int st = ws.ctr->error ? (int)OZL_ERR_DSTCAPACITY_TOOSMALL : (int)H.status;
    if (st) {
        if (compressed && threadIdx.x == 0) status[c] = st;
        return;
    }
    uint8_t* fr = dst + frameOff[c];
    const size_t flen = frameOff[c + 1] - frameOff[c];
    if (!compressed) {
        const uint64_t h = ozl_xxh3_wave(reinterpret_cast<const uint8_t*>((uintptr_t)H.src), (size_t)H.n * H.w, nullptr, 0);
        if (threadIdx.x == 0)
            for (int k = 0; k < 4; k++) fr[flen - 9 + k] = (uint8_t)(h >> (8 * k));
        return;
    }
    uint8_t fhb[24];
    const size_t fh = ozl_write_frame_header(fhb, OZL_T_NUMERIC, (uint64_t)H.n * H.w, 1, 1);
    const size_t clen = flen - fh - 4 - 1;
    const uint64_t h = ozl_xxh3_wave(fr + fh, clen, nullptr, 0);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; k++) fr[fh + clen + k] = (uint8_t)(h >> (8 * k));
        fr[flen - 1] = 0;
        status[c] = 0;
    }
}

// ------------------------------------------------------------------ workspace
inline size_t a256(size_t v) { return (v + 255) & ~(size_t)255; }
constexpr uint32_t kMaxBlocksPerChunk = 18;   // raw, delta, 8 planes, 8 delta planes

size_t ews_bytes(uint32_t B, size_t maxBlock)
{
    return (size_t)B * (sizeof(ozl_eplan) + OZL_TAB_BYTES + OZL_AUX_BYTES + ozl_payload_cap(maxBlock)) + 512;
}

struct Carved {
    SelWs ws;
    size_t need;
};

Carved carve(void* base, uint32_t K, size_t maxChunk)
{
    Carved r;
    uint8_t* p0 = reinterpret_cast<uint8_t*>(a256((uintptr_t)base));
    uint8_t* p = p0;
    SelWs& w = r.ws;
    w.ctr = reinterpret_cast<SelCtr*>(p); p += a256(sizeof(SelCtr));
    w.ch = reinterpret_cast<SelChunk*>(p); p += a256((size_t)K * sizeof(SelChunk));
    w.off = reinterpret_cast<uint64_t*>(p); p += a256(((size_t)K * kMaxBlocksPerChunk + 1) * 8);
    w.fsz = reinterpret_cast<uint64_t*>(p); p += a256((size_t)K * 8);
    w.fr = reinterpret_cast<SelFrame*>(p); p += a256((size_t)K * sizeof(SelFrame));
    w.tiny = p; p += a256((size_t)K * kTinySlot);
    w.stgCap = (size_t)K * 4 * maxChunk;
    w.stg = p; p += a256(w.stgCap + 64);
    w.packCap = (size_t)K * (maxChunk + 16);
    w.pack = p; p += a256(w.packCap + 64);
    w.ewsCap = ews_bytes(K * 8, maxChunk) + ews_bytes(K * kMaxBlocksPerChunk, maxChunk);
    w.ews = p; p += a256(w.ewsCap);
    r.need = (size_t)(p - p0) + 256;
    return r;
}

// setup, scans, staging and the entropy coders; returns the layout context
int sel_front(const uint8_t* src, const uint64_t* spans, const uint8_t* widths, const uint8_t* choice, uint32_t K,
              size_t maxChunk, const SelWs& ws, LayoutCtx& X, hipStream_t st)
{
    const int G = (int)((K + kT - 1) / kT);
    OZL_LAUNCH("k_sel_setup", k_sel_setup, dim3(G), dim3(kT), 0, st, src, spans, widths, choice, K, (uint64_t)maxChunk, ws);
    OZL_LAUNCH("k_sel_scan", k_sel_scan, dim3(1), dim3(kT), 0, st, K, ws);
    SelCtr ctr;
    OZL_HIP_CHECK(hipMemcpyAsync(&ctr, ws.ctr, sizeof(ctr), hipMemcpyDeviceToHost, st));
    OZL_HIP_CHECK(hipStreamSynchronize(st));
    if (ctr.error) return OZL_ERR_INTERNAL_BUFFER;
    OZL_LAUNCH("k_sel_prep", k_sel_prep, dim3(K), dim3(kT), 0, st, K, ws);
    const uint32_t hufB = ctr.blocks[0] + ctr.blocks[1];
    const uint32_t fseB = ctr.blocks[1] + ctr.blocks[2] + ctr.blocks[3];
    if ((size_t)ews_bytes(hufB, maxChunk) + ews_bytes(fseB, maxChunk) > ws.ewsCap) return OZL_ERR_INTERNAL_BUFFER;
    X.ws = ws;
    X.e.huf = ozl_ews_carve(ws.ews, hufB ? hufB : 1, maxChunk);
    X.e.fse = ozl_ews_carve(ws.ews + ews_bytes(hufB, maxChunk), fseB ? fseB : 1, maxChunk);
    X.e.hufB = hufB;
    X.e.fseFirst = ctr.bBase[1];
    if (hufB) {
        int rc = ozl_entropy_encode_chunks(1, ws.stg, ws.off, ws.off + 1, hufB, maxChunk, X.e.huf, st);
        if (rc) return rc;
    }
    if (fseB) {
        int rc = ozl_entropy_encode_chunks(0, ws.stg, ws.off + ctr.bBase[1], ws.off + ctr.bBase[1] + 1, fseB, maxChunk, X.e.fse, st);
        if (rc) return rc;
    }
    return OZL_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

// Workspace bytes for K chunks of at most maxChunk bytes.
size_t ozl_select_workspace(uint32_t K, size_t maxChunk) { return carve(nullptr, K, maxChunk).need; }

// Upper bound of one frame for a chunk of n bytes (any candidate).
size_t ozl_select_frame_bound(size_t n) { return n + n / 64 + 2048; }

// Exhaustive trial over the eight candidates for K numeric chunks (chunk c =
// src[spans[2c], spans[2c] + spans[2c+1]) with element width widths[c], the
// start aligned to the width, fewer than 100000 bytes). sizes (K x 8, may be
// null) receives each candidate's frame size with checksums off (0 when the
// candidate failed), choice (K, may be null) the selected candidate: the first
// strictly smallest below the chunk size, else 0 (ZL_GRAPH_STORE). Device
// pointers; one host synchronisation per call.
int ozl_select_trial(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t K, size_t maxChunk,
                     uint32_t* sizes, uint8_t* choice, void* workspace, size_t wsBytes, void* stream)
{
    if (K == 0) return OZL_OK;
    if (maxChunk == 0 || maxChunk >= kMaxChunk) return OZL_ERR_PARAMETER_INVALID;
    Carved cv = carve(workspace, K, maxChunk);
    if (!workspace || wsBytes < cv.need) return OZL_ERR_INTERNAL_BUFFER;
    hipStream_t st = (hipStream_t)stream;
    LayoutCtx X;
    int rc = sel_front((const uint8_t*)src, spans, widths, nullptr, K, maxChunk, cv.ws, X, st);
    if (rc) return rc;
    OZL_LAUNCH("k_sel_trial_sizes", k_sel_trial_sizes, dim3((K * kNC + 63) / 64), dim3(64), 0, st, K, X);
    OZL_LAUNCH("k_sel_choose", k_sel_choose, dim3((K + kT - 1) / kT), dim3(kT), 0, st, K, cv.ws, sizes, choice);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// Compresses chunk c with candidate choice[c] (0..7) into one frame (format
// version 23, both checksums on): the frame the reference writes for that
// candidate graph, and, when choice comes from ozl_select_trial, the frame of
// the brute-force selector over the eight candidates. Frames go back to back
// into dst; frameOff (K+1) receives their offsets, status (K) 0 or an error
// code per chunk. Device pointers; one host synchronisation per call.
int ozl_select_compress(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t K, size_t maxChunk,
                        const uint8_t* choice, void* dst, size_t dstCap, uint64_t* frameOff, int* status,
                        void* workspace, size_t wsBytes, void* stream)
{
    if (K == 0) return OZL_OK;
    if (maxChunk == 0 || maxChunk >= kMaxChunk || !choice) return OZL_ERR_PARAMETER_INVALID;
    Carved cv = carve(workspace, K, maxChunk);
    if (!workspace || wsBytes < cv.need) return OZL_ERR_INTERNAL_BUFFER;
    hipStream_t st = (hipStream_t)stream;
    LayoutCtx X;
    int rc = sel_front((const uint8_t*)src, spans, widths, choice, K, maxChunk, cv.ws, X, st);
    if (rc) return rc;
    uint8_t* d = (uint8_t*)dst;
    OZL_LAUNCH("k_sel_frame_size", k_sel_frame_size, dim3((K + 63) / 64), dim3(64), 0, st, K, X);
    OZL_LAUNCH("k_sel_scan_frames", k_sel_scan_frames, dim3(1), dim3(kT), 0, st, K, cv.ws, frameOff, dstCap);
    OZL_LAUNCH("k_sel_pack", k_sel_pack, dim3(K), dim3(kT), 0, st, K, cv.ws);
    OZL_LAUNCH("k_sel_write", k_sel_write, dim3(K), dim3(kT), 0, st, K, cv.ws, d, frameOff);
    OZL_LAUNCH("k_sel_content_hash", k_sel_hash, dim3(K), dim3(64), 0, st, K, cv.ws, d, frameOff, 0, status);
    OZL_LAUNCH("k_sel_frame_hash", k_sel_hash, dim3(K), dim3(64), 0, st, K, cv.ws, d, frameOff, 1, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"
