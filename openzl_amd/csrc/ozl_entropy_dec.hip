// Batched decoding of FSE / Huffman graph frames (the universal decompressor
// path of src/openzl/decompress/decompress2.c:1442-1580 restricted to the
// single-chunk frames ZL_GRAPH_FSE and ZL_GRAPH_HUFFMAN produce for inputs
// below 100000 bytes).
//
//   k_dparse       one wave per frame: the first KiB of the frame is copied to
//                  LDS, lane 0 parses it (ozl_dparse.h) and leaves the
//                  normalized counts or the Huffman code lengths of the frame
//                  in the workspace
//   k_fse_decode   6 frames per one-wave workgroup: the wave builds each
//                  frame's decoding table in LDS (fse_decompress.c:73-176),
//                  then a quad of lanes per frame decodes it, one lane per
//                  interleaved state (:255-338); the lanes of a frame share
//                  one bit position, their read offsets come from a quad
//                  prefix sum (DPP)
//   k_huf_decode   8 frames per wave: the wave builds each frame's lookup
//                  table from the code lengths, then one lane per Huffman
//                  stream decodes (huf_decompress.c:587-787)
//   k_pack_decode  store / constant / flatpack / bitpack
//   k_dverify      two wavefronts per frame: compressed and content checksums
//
// Bit streams are read backwards through 64-bit windows over a ring in LDS
// per stream, which the stream's lanes refill from HBM with aligned 16-byte
// loads issued one staging step ahead of use, so no global-memory latency
// sits on a decoding lane's dependency chain. Decoded bytes leave in 4-byte
// stores.
#include <mutex>

#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_dparse.h"
#include "ozl_wave.h"
#include "ozl_profile.h"

#define OZL_DSIDE_BYTES 512  // per frame: FSE normalized counts int16[256] or Huffman code lengths u8[256]

struct ozl_dws {
    ozl_dplan* plans;
    uint8_t* side;
};

namespace {

using namespace ozl_wave;

constexpr int kT = 256;
constexpr int kParseT = 64;
constexpr int kPrefix = OZL_DPREFIX;

__global__ void __launch_bounds__(kParseT) k_dparse(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                    uint32_t B, uint8_t* __restrict__ dst, const uint64_t* __restrict__ doff,
                                                    ozl_dws ws)
{
    __shared__ ozl_dlds L;
    __shared__ __attribute__((aligned(16))) uint8_t pre[kPrefix + 16];
    __shared__ ozl_dplan sD;
    __shared__ uint32_t sM;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    const uint8_t* fr = frames + foff[b];
    size_t flen = foff[b + 1] - foff[b];
    size_t pn = flen < (size_t)kPrefix ? flen : (size_t)kPrefix;
    // the frame's first KiB as words (zeros past the frame), zero counts
    {
        uint32_t* pw = reinterpret_cast<uint32_t*>(pre);
        for (uint32_t i = threadIdx.x; i < (kPrefix + 16) / 4; i += kParseT) {
            uint32_t x = 0;
            for (uint32_t k = 0; k < 4; k++) {
                const size_t j = 4 * (size_t)i + k;
                if (j < pn) x |= (uint32_t)fr[j] << (8 * k);
            }
            pw[i] = x;
        }
        for (int s = threadIdx.x; s < 256; s += kParseT) L.norm[s] = 0;
    }
    if (threadIdx.x == 0) sM = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        ozl_dplan d;
        ozl_fview v = { fr, pre, pn, reinterpret_cast<const uint32_t*>(pre) };
        ozl_dparse_frame(v, flen, L, d, true, &sM);
        d.srcPtr = (uint64_t)(uintptr_t)(fr + d.payOff);
        d.outPtr = (uint64_t)(uintptr_t)(dst + doff[b]);
        if (d.status == 0 && doff[b + 1] - doff[b] != d.n) d.status = OZL_ERR_DSTCAPACITY_TOOSMALL;
        sD = d;
    }
    __syncthreads();
    if (sD.status == 0 && sD.bk == OZL_BK_HUF) {
        // code lengths from the m weights (ozl_huf_from_weights): a weight
        // above 12 or a sum of 2^(w-1) that is not a power of two is corrupt
        const uint32_t m = sM;
        uint32_t tot = 0, bad = 0;
        uint8_t wl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t sy = threadIdx.x + 64u * k;
            const uint32_t w = sy < m ? L.w[sy] : 0u;
            wl[k] = (uint8_t)w;
            bad |= w > 12 ? 1u : 0u;
            tot += w && w <= 12 ? 1u << (w - 1) : 0u;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            tot += __shfl_xor(tot, d, 64);
            bad |= __shfl_xor(bad, d, 64);
        }
        const uint32_t tl = tot ? (uint32_t)ozl_hb32(tot) : 0u;
        const bool okW = !bad && tot && (tot & (tot - 1)) == 0 && tl >= 1 && tl <= 11;
        if (okW) {
#pragma unroll
            for (int k = 0; k < 4; k++) L.wlen[threadIdx.x + 64 * k] = wl[k] ? (uint8_t)(tl + 1 - wl[k]) : 0;
        }
        if (threadIdx.x == 0) {
            if (okW) sD.tl = (uint8_t)tl;
            else sD.status = OZL_ERR_CORRUPTION;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ws.plans[b] = sD;
    const int bk = sD.status == 0 ? sD.bk : 255;
    uint8_t* side = ws.side + (size_t)b * OZL_DSIDE_BYTES;
    if (bk == OZL_BK_FSE)
        for (int s = threadIdx.x; s < 256; s += kParseT) reinterpret_cast<int16_t*>(side)[s] = L.norm[s];
    else if (bk == OZL_BK_HUF)
        for (int s = threadIdx.x; s < 256; s += kParseT) side[s] = L.wlen[s];
}

__device__ inline uint64_t floor16(uint64_t a) { return a & ~15ull; }

// ---------------------------------------------------------------- self-staged backward bit readers
// Each decoding lane owns one stream and a ring of R bytes in LDS holding the
// stream by absolute address (slot = address mod R). Every P rounds the lane
// stores the aligned 16-byte blocks it loaded P rounds earlier and issues the
// loads of the next blocks below, keeping LOOK bytes staged below its window;
// bytes below the stream start read as zero. The window is 64 bits,
// MSB-aligned, refilled 32 bits at a time from 4-aligned addresses.
// global-memory accesses through address-space-1 pointers (global_load /
// global_store rather than flat, which the compiler must order against LDS)
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 gc_u32x4;
typedef __attribute__((address_space(1))) const uint32_t gc_uint32;
__device__ inline uint4 gload16(uint64_t a)
{
    u32x4 v = *(gc_u32x4*)(uintptr_t)a;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ inline uint32_t gload4(uint64_t a) { return *(gc_uint32*)(uintptr_t)a; }
typedef __attribute__((address_space(1))) uint32_t g_uint32;
typedef __attribute__((address_space(1))) uint8_t g_uint8;

template <int R>
__device__ inline uint32_t ring_dword(const uint32_t* ring, uint64_t a)
{
    return ring[(uint32_t)(a >> 2) & (R / 4 - 1)];
}

struct Stager {
    uint64_t lo;     // stream start
    uint64_t aLo;    // lowest staged address (16-aligned)
    uint4 pend[3];
    uint64_t pa[3];
    uint32_t np;
};

// initial fill of [max(floor16(top - look), floor16(lo - 8)), top), by this lane alone
// The dwords of [a0, ceil4(top)) are loaded 8 at a time (the last one may
// cover up to 3 bytes past the stream, which the window never reads; the
// frame holds at least its end byte after every stream).
template <int R>
__device__ inline void stage_init(uint32_t* ring, Stager& g, uint64_t lo, uint64_t top, uint64_t look)
{
    constexpr uint32_t W = R / 4 - 1;
    uint64_t a0 = floor16(top - look), f = floor16(lo - 8);
    if (a0 < f) a0 = f;
    uint64_t a1 = (top + 3) & ~3ull;
    for (uint64_t a = a0; a < a1; a += 32) {
        uint32_t d[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint64_t da = a + 4 * j;
            d[j] = (da < a1 && da + 4 > lo) ? gload4(da) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint64_t da = a + 4 * j;
            uint32_t x = d[j];
            if (da < lo) x = da + 4 <= lo ? 0u : x & (~0u << (8 * (uint32_t)(lo - da)));
            if (da < a1) ring[(uint32_t)(da >> 2) & W] = x;
        }
    }
    g.lo = lo;
    g.aLo = a0;
    g.np = 0;
}

template <int R, int MAXB>
__device__ inline void stage_step(uint32_t* ring, Stager& g, uint64_t winLo, uint64_t look)
{
    constexpr uint32_t W = R / 4 - 1;
#pragma unroll
    for (int k = 0; k < MAXB; k++) {
        if ((uint32_t)k >= g.np) break;
        uint32_t d[4] = { g.pend[k].x, g.pend[k].y, g.pend[k].z, g.pend[k].w };
        uint32_t w = (uint32_t)(g.pa[k] >> 2);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            uint64_t da = g.pa[k] + 4 * j;
            uint32_t x = d[j];
            if (da + 4 <= g.lo) x = 0;
            else if (da < g.lo) x &= ~0u << (8 * (uint32_t)(g.lo - da));
            ring[(w + j) & W] = x;
        }
    }
    uint64_t want = floor16(winLo - look), f = floor16(g.lo - 8);
    if (want < f) want = f;
    uint32_t np = 0;
#pragma unroll
    for (int k = 0; k < MAXB; k++) {
        uint64_t a = g.aLo - 16 * (uint64_t)(k + 1);
        if (a >= want && a < g.aLo) {
            g.pa[k] = a;
            g.pend[k] = a + 16 <= g.lo ? make_uint4(0, 0, 0, 0) : gload16(a);
            np = k + 1;
        }
    }
    g.np = np;
    g.aLo -= 16 * (uint64_t)np;
}

struct Win {
    uint64_t acc;
    int valid;
    int lb;  // byte offset of the window bottom in the stream (4-aligned address after init)
};

// Window start: the top 4..7 bytes of the stream (so that the bottom lands on
// a 4-aligned address), minus the bits above and including the end mark.
// Returns false when the last byte is zero.
template <int R>
__device__ inline bool win_init(Win& w, const uint32_t* ring, uint64_t s, uint32_t len)
{
    uint64_t top = s + len;
    uint64_t a0 = (top - 8) & ~3ull;
    uint32_t d0 = ring_dword<R>(ring, a0), d1 = ring_dword<R>(ring, a0 + 4), d2 = ring_dword<R>(ring, a0 + 8);
    uint32_t sh = (uint32_t)((top - 8) & 3) * 8;
    uint64_t lo64 = (uint64_t)d0 | ((uint64_t)d1 << 32);
    uint64_t v = sh ? (lo64 >> sh) | ((uint64_t)d2 << (64 - sh)) : lo64;  // bytes [top-8, top)
    uint32_t take = 4 + (uint32_t)(top & 3);
    w.acc = v & ~((1ull << (64 - 8 * take)) - 1);
    w.valid = (int)(8 * take);
    w.lb = (int)len - (int)take;
    uint32_t topByte = (uint32_t)(w.acc >> 56);
    if (len == 0 || topByte == 0) return false;
    int skip = 8 - ozl_hb32(topByte);
    w.acc <<= skip;
    w.valid -= skip;
    return true;
}

// adds the 32 bits below the window when it holds <= 32 (the dword `d` was read
// at the window bottom before it moved)
__device__ inline void win_refill32(Win& w, uint32_t d)
{
    bool need = w.valid <= 32;
    uint32_t x = w.lb > 0 ? d : 0u;
    uint64_t add = need ? ((uint64_t)x << (32 - w.valid)) : 0ull;
    w.acc |= add;
    w.valid += need ? 32 : 0;
    w.lb -= need ? 4 : 0;
}

__device__ inline uint32_t win_hi(const Win& w) { return (uint32_t)(w.acc >> 32); }

// ---------------------------------------------------------------- FSE decoding
// A quad of lanes decodes one frame of 4 interleaved states
// (fse_decompress.c:255-338), lane q holding state q: per round every lane
// looks its state up, the quad adds up its 4 bit counts (each lane's count
// subtracted from bytes q..3 of a word, the words summed over the quad with
// two DPP exchanges, so that byte k holds 63 - the bits of states 0..k), and
// lane q takes its bits below those of the states before it from a 64-bit
// window that every lane reads from the frame's LDS ring at the quad's bit
// position. After 4 rounds the quad's 16 symbols are transposed across its
// lanes (two DPP byte permutes) into one dword per lane, and after 16 rounds
// the dwords once more, so that a lane stores 16 consecutive bytes. Frames of
// 2 states are decoded by lane 0
// of their quad, state after state. 8 frames per one-wave workgroup: the
// 16-bit table entries carry an identifier of the cell's symbol in the bits
// their state base leaves zero (ozl_wave::wave_fse_dtable_ids), so a frame
// takes its 4 KB table, a 256-byte symbol table and its ring; 4 workgroups
// (one per SIMD) share a CU, and the 16,384 frames of a 1 GiB batch of
// 64 KiB blocks run in two rounds of workgroups.
constexpr int kFQ = 8;
constexpr int kFRing = 256;
constexpr int kFRingW = kFRing / 4;
constexpr int kFLook = 128;     // one-lane frames
constexpr int kFPhase = 8;      // one-lane frames: rounds between staging steps (<= 44 bytes consumed)
constexpr int kQLook = 192;     // quads: bytes kept staged below the window
constexpr int kQPhase = 16;     // quads: rounds between staging steps (<= 88 bytes consumed)

// Frame f's decoding table starts at byte 4096 f of dt, so a lane keeps one
// LDS byte offset per state: 4096 f + 2 state.
struct FseDecLds {
    // the rings first: their addresses fit the LDS instructions' offsets
    uint32_t ring[kFQ][kFRingW + 4];  // dwords kFRingW, kFRingW+1 mirror 0, 1
    uint16_t dt[kFQ][2048];
    uint8_t sys[kFQ][256];
    FseBuildHdr bs;
    ozl_dplan plan[kFQ];
};

// the symbol of table entry e (wave_fse_dtable_ids): m1 = 2^(nbBits+1) - 1
__device__ inline uint32_t fse_entry_symbol(const uint8_t* sys, uint32_t e, uint32_t nb, uint32_t m1)
{
    const uint32_t id = e & m1;
    const uint32_t ys = sys[(m1 - 1 + id) & 255];   // in the table whatever nb (read, then chosen)
    return nb >= 7 ? id : ys;
}

// A frame's quad stages its stream into the ring 128 bytes per step, lane q
// loading blocks q and q+4 of the 16-byte blocks below the staged range; each
// step first stores the blocks the previous step loaded (kQPhase rounds
// earlier, so their latency is hidden). The initial fill covers the top
// kQLook bytes. Blocks may hold bytes below the stream start (down to 8 bytes
// below it, which the frame header precedes) and above its end: the window
// reads them, but a valid stream never takes their bits.
struct QStage {
    uint64_t aLo;     // lowest address loaded or requested (16-aligned)
    uint64_t f;       // floor16(stream start - 8)
    uint4 p0, p1;     // blocks loaded by the last step
    uint64_t a0, a1;  // their addresses (0: none)
};


__device__ inline void qring_put(uint32_t* ring, uint64_t a, uint4 v)
{
    const uint32_t w = (uint32_t)(a >> 2) & (kFRingW - 1);
    *reinterpret_cast<uint4*>(ring + w) = v;
    if (w == 0) *reinterpret_cast<uint2*>(ring + kFRingW) = make_uint2(v.x, v.y);
}

__device__ inline void qstage_init(uint32_t* ring, QStage& g, uint64_t lo, uint64_t top, uint32_t q)
{
    g.f = floor16(lo - 8);
    uint64_t a0 = floor16(top - kQLook);
    if (a0 < g.f) a0 = g.f;
    for (uint64_t a = a0 + 16 * q; a < top; a += 64) qring_put(ring, a, gload16(a));
    g.aLo = a0;
    g.a0 = g.a1 = 0;
}

// a step is qstage_put (waits for the previous step's loads) then
// qstage_load; stores issued between the two do not delay the next put
__device__ inline void qstage_put(uint32_t* ring, QStage& g)
{
    if (g.a0) qring_put(ring, g.a0, g.p0);
    if (g.a1) qring_put(ring, g.a1, g.p1);
}

__device__ inline void qstage_load(QStage& g, uint64_t winLo, uint32_t q)
{
    uint64_t want = floor16(winLo - kQLook);
    if (want < g.f) want = g.f;
    uint64_t nb = g.aLo > want ? (g.aLo - want) >> 4 : 0;
    if (nb > 8) nb = 8;
    const uint64_t x0 = g.aLo - 16 * (uint64_t)(q + 1), x1 = x0 - 64;
    g.p0 = gload16(q < nb ? x0 : g.f);
    g.p1 = gload16(q + 4 < nb ? x1 : g.f);
    g.a0 = q < nb ? x0 : 0;
    g.a1 = q + 4 < nb ? x1 : 0;
    g.aLo -= 16 * nb;
}

__device__ inline void qstage_step(uint32_t* ring, QStage& g, uint64_t winLo, uint32_t q)
{
    qstage_put(ring, g);
    qstage_load(g, winLo, q);
}

// the 64 stream bits above absolute bit address Qm (mod 2^32): the window of
// a reader at Qm + 64, with the dword index taken by one bit-field extract
__device__ inline uint64_t qwindow_m(const uint32_t* ring, uint32_t Qm)
{
    static_assert(kFRingW == 64, "ring index: 6 bits");
    const uint32_t j = __builtin_amdgcn_ubfe(Qm, 5, 6);
    const uint32_t d0 = ring[j], d1 = ring[j + 1], d2 = ring[j + 2];
    return ((uint64_t)__builtin_amdgcn_alignbit(d2, d1, Qm) << 32) | __builtin_amdgcn_alignbit(d1, d0, Qm);
}

// the 64 stream bits below absolute bit address Q (mod 2^32)
__device__ inline uint64_t qwindow(const uint32_t* ring, uint32_t Q)
{
    const uint32_t j = ((Q >> 5) - 2) & (kFRingW - 1);
    const uint32_t d0 = ring[j], d1 = ring[j + 1], d2 = ring[j + 2];
    return ((uint64_t)__builtin_amdgcn_alignbit(d2, d1, Q) << 32) | __builtin_amdgcn_alignbit(d1, d0, Q);
}

// Frames of any state count, one lane: per round one lookup per state, then
// the states take their bits from the window in state order; the window is
// refilled after each pair of states so that it always holds the <= 22 bits a
// pair reads.
__device__ bool fse_decode_lane(const ozl_dplan& P, const uint16_t* DT, const uint8_t* SYS, uint32_t* ring)
{
    const uint32_t S = P.nbStates, n = P.n, tl = P.tl;
    const uint64_t s = P.srcPtr;
    uint8_t* out = reinterpret_cast<uint8_t*>((uintptr_t)P.outPtr);
    Stager g;
    stage_init<kFRing>(ring, g, s, s + P.payLen, kFLook);
    Win w;
    bool ok = win_init<kFRing>(w, ring, s, P.payLen);
    uint32_t st[4] = { 0, 0, 0, 0 };
    if (ok) {
        win_refill32(w, ring_dword<kFRing>(ring, s + (uint64_t)(int64_t)w.lb - 4));
        for (uint32_t j = 0; j < S; j++) {
            st[j] = win_hi(w) >> (32 - tl);
            w.acc <<= tl;
            w.valid -= (int)tl;
            if (j == 1) win_refill32(w, ring_dword<kFRing>(ring, s + (uint64_t)(int64_t)w.lb - 4));
        }
        win_refill32(w, ring_dword<kFRing>(ring, s + (uint64_t)(int64_t)w.lb - 4));
    }
    const uint32_t R = ok ? (n + S - 1) / S : 0u;
    for (uint32_t r = 0; r < R; r++) {
        if ((r & (kFPhase - 1)) == 0) stage_step<kFRing, 3>(ring, g, (uint64_t)((int64_t)s + w.lb), kFLook);
        for (uint32_t j = 0; j < S; j++) {
            const uint32_t p = r * S + j;
            const uint32_t d = DT[st[j]];
            const uint32_t nbd = d >> 12, m1 = (2u << nbd) - 1u;
            if (p < n) out[p] = (uint8_t)fse_entry_symbol(SYS, d, nbd, m1);
            // a state updates unless its symbol is among the last S
            const uint32_t nb = p + S < n ? nbd : 0u;
            const uint32_t bits = __builtin_amdgcn_ubfe(win_hi(w), 32 - nb, nb);
            w.acc <<= nb;
            w.valid -= (int)nb;
            if (p + S < n) st[j] = ((d & 0xFFFu & ~m1) >> 1) + bits;
            if (j & 1) win_refill32(w, ring_dword<kFRing>(ring, s + (uint64_t)(int64_t)w.lb - 4));
        }
        if (S & 1) win_refill32(w, ring_dword<kFRing>(ring, s + (uint64_t)(int64_t)w.lb - 4));
    }
    long pos = 8L * w.lb + w.valid;
    return ok && pos == 0;
}

// a dword of decoded bytes to global memory (byte stores when the
// destination is not 4-aligned); four of them `stride` bytes apart
__device__ inline void put_dword(uint64_t a, uint32_t d, bool wide)
{
    if (wide) {
        *(g_uint32*)(uintptr_t)a = d;
    } else {
        g_uint8* ob = (g_uint8*)(uintptr_t)a;
        ob[0] = (uint8_t)d;
        ob[1] = (uint8_t)(d >> 8);
        ob[2] = (uint8_t)(d >> 16);
        ob[3] = (uint8_t)(d >> 24);
    }
}
__device__ inline void put4_dwords(uint64_t a, uint32_t stride, uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3,
                                   bool wide)
{
    put_dword(a, d0, wide);
    put_dword(a + stride, d1, wide);
    put_dword(a + 2 * stride, d2, wide);
    put_dword(a + 3 * stride, d3, wide);
}

// 4x4 transpose of dwords across a quad: lane q's register k becomes lane
// k's register q (two exchange stages, lanes xor 1 then xor 2)
__device__ inline void quad_transpose4(uint32_t& r0, uint32_t& r1, uint32_t& r2, uint32_t& r3, uint32_t q)
{
    const bool o1 = q & 1, o2 = q & 2;
    uint32_t t = quad_xor1(o1 ? r0 : r1);
    r0 = o1 ? t : r0;
    r1 = o1 ? r1 : t;
    t = quad_xor1(o1 ? r2 : r3);
    r2 = o1 ? t : r2;
    r3 = o1 ? r3 : t;
    t = quad_xor2(o2 ? r0 : r2);
    r0 = o2 ? t : r0;
    r2 = o2 ? r2 : t;
    t = quad_xor2(o2 ? r1 : r3);
    r1 = o2 ? t : r1;
    r3 = o2 ? r3 : t;
}

// 16 bytes to global memory (dword stores when the destination is not
// 16-aligned, byte stores when it is not 4-aligned)
__device__ inline void put_16(uint64_t a, uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3, bool wide)
{
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    if (wide && (a & 15) == 0) {
        u32x4 v = { d0, d1, d2, d3 };
        *(__attribute__((address_space(1))) u32x4*)(uintptr_t)a = v;
    }
    else put4_dwords(a, 4, d0, d1, d2, d3, wide);
}

// one round of a 4-state frame's quad: returns this lane's symbol; TAIL
// rounds keep the states of the last 4 symbols (fse_decompress.c:287-336).
// W holds the 64 bits above bit address Qm (the reader sits at Qm + 64); the
// lane's state is the LDS byte offset o = 4096 f + 2 state. The quad sums, in
// bytes, 63 minus the bit counts of the states (lane 0 adds the 63s, each
// lane subtracts its count from bytes q..3, a byte broadcast by v_perm), so
// byte k holds 63 - (bits of states 0..k): shifted right by that (mod 64),
// the window has this lane's bits at bits 1..nbBits, where the entry's
// 2 * newStateBase has zeros (and its symbol identifier, masked off), and
// 4096 f is above it, so the new offset is an AND-OR.
template <bool TAIL>
__device__ inline uint32_t fse_quad_round(const FseDecLds& L, const uint32_t* ring, uint32_t& o, uint64_t& W, uint32_t& Qm,
                                          uint32_t sq, uint32_t selq, uint32_t cq, uint32_t fb, const uint8_t* sys,
                                          uint32_t p, uint32_t n)
{
    const uint32_t e = *reinterpret_cast<const uint16_t*>(reinterpret_cast<const uint8_t*>(&L.dt[0][0]) + o);
    const bool upd = !TAIL || p + 4 < n;
    const uint32_t nbe = e >> 12, m1 = (2u << nbe) - 1u;
    const uint32_t nb = upd ? nbe : 0u;
    uint32_t v = cq - __builtin_amdgcn_perm(0u, nb, selq);
    v += quad_xor1(v);
    const uint32_t M = v + quad_xor2(v);   // byte k: 63 - bits of states 0..k
    const uint32_t X = (uint32_t)(W >> (__builtin_amdgcn_ubfe(M, sq, 8) & 63));
    const uint32_t b2 = (e & (0xFFFu ^ m1)) | fb;
    const uint32_t no = (X & (m1 ^ 1u)) | b2;
    o = upd ? no : o;
    Qm += (M >> 24) - 63u;
    W = qwindow_m(ring, Qm);
    return fse_entry_symbol(sys, e, nbe, m1);
}

__global__ void __launch_bounds__(64) k_fse_decode(uint32_t B, ozl_dws ws)
{
    __shared__ FseDecLds L;
    const int lane = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kFQ;
    for (int i = lane; i < kFQ * 16; i += 64) {
        int f = i >> 4, k = i & 15;
        reinterpret_cast<uint32_t*>(&L.plan[f])[k] =
            b0 + f < B ? reinterpret_cast<const uint32_t*>(&ws.plans[b0 + f])[k] : 0u;
    }
    wave_sync();
    uint32_t factMask = 0;
    for (int f = 0; f < kFQ; f++) {
        const ozl_dplan& P = L.plan[f];
        if (b0 + f < B && P.status == 0 && P.bk == OZL_BK_FSE) factMask |= 1u << f;
    }
    if (!factMask) return;
    for (int f = 0; f < kFQ; f++) {
        if (!((factMask >> f) & 1)) continue;
        // the spread's cells in the table's upper half, its scratch in the lower
        uint8_t* dtb = reinterpret_cast<uint8_t*>(L.dt[f]);
        wave_fse_dtable_ids(L.dt[f], L.sys[f], dtb + 2048, dtb,
                            reinterpret_cast<const int16_t*>(ws.side + (size_t)(b0 + f) * OZL_DSIDE_BYTES), L.plan[f].aux,
                            L.plan[f].tl, L.bs);
    }
    wave_sync();
    const int f = lane >> 2;
    const uint32_t q = lane & 3;
    if (f >= kFQ || !((factMask >> f) & 1)) return;
    const ozl_dplan P = L.plan[f];
    uint32_t* ring = L.ring[f];
    if (P.nbStates != 4) {
        if (q == 0 && !fse_decode_lane(P, L.dt[f], L.sys[f], ring)) ws.plans[b0 + f].status = OZL_ERR_CORRUPTION;
        return;
    }
    const uint32_t n = P.n, tl = P.tl, sq = 8 * q;
    // v_perm selector of the count in bytes q..3 (0: byte 0, 0x0c: zero)
    const uint32_t selq = 0x0c0c0c0cu & ~(0xFFFFFFFFu << sq);
    const uint32_t cq = q == 0 ? 0x3F3F3F3Fu : 0u;
    const uint8_t* sys = L.sys[f];
    const uint32_t fb = 4096u * (uint32_t)f;
    const uint64_t s = P.srcPtr, top = s + P.payLen;
    uint8_t* out = reinterpret_cast<uint8_t*>((uintptr_t)P.outPtr);
    const bool wide = (P.outPtr & 3) == 0;
    // byte transpose of the quad's 4x4 symbols: pairs of lanes swap bytes,
    // then pairs of pairs swap halves
    const uint32_t selA = (q & 1) ? 0x03070105u : 0x06020400u;
    const uint32_t selB = (q & 2) ? 0x03020706u : 0x05040100u;
    QStage g;
    qstage_init(ring, g, s, top, q);
    wave_sync();
    // end mark: the highest set bit of the last byte
    uint64_t Pb = 8 * top;
    uint64_t W = qwindow(ring, (uint32_t)Pb);
    const uint32_t topByte = (uint32_t)(W >> 56);
    bool ok = P.payLen != 0 && topByte != 0;
    Pb -= 8 - (uint32_t)ozl_hb32(topByte | 1);
    W = qwindow(ring, (uint32_t)Pb);
    // initial states, state 0 in the highest tl bits
    uint32_t o = fb + 2 * __builtin_amdgcn_ubfe((uint32_t)(W >> (64 - (q + 1) * tl)), 0, tl);
    Pb -= 4 * tl;
    uint32_t Q = (uint32_t)Pb - 64;   // window base: the reader sits at Q + 64
    W = qwindow_m(ring, Q);
    // a group is 4 rounds (16 symbols, one dword per lane after the
    // transpose); the groups whose states all update run 4 per staging step,
    // and their dwords are stored at the next step between its ring stores
    // and its loads, so that waiting for a step's loads never waits for
    // output stores (both count in vmcnt, which drains in order)
    auto group = [&]() -> uint32_t {
        uint32_t acc = fse_quad_round<false>(L, ring, o, W, Q, sq, selq, cq, fb, sys, 0, n);
        acc |= fse_quad_round<false>(L, ring, o, W, Q, sq, selq, cq, fb, sys, 0, n) << 8;
        acc |= fse_quad_round<false>(L, ring, o, W, Q, sq, selq, cq, fb, sys, 0, n) << 16;
        acc |= fse_quad_round<false>(L, ring, o, W, Q, sq, selq, cq, fb, sys, 0, n) << 24;
        const uint32_t c = __builtin_amdgcn_perm(quad_xor1(acc), acc, selA);
        return __builtin_amdgcn_perm(quad_xor2(c), c, selB);
    };
    const uint64_t oq = P.outPtr + 4 * q;
    const uint32_t G = ok && n > 19 ? (n - 20) / 16 + 1 : 0u, G4 = G & ~3u;
    uint32_t o0 = 0, o1 = 0, o2 = 0, o3 = 0;
    // after 4 groups, lane q holds dword q of each; transposed, it holds the
    // 16 bytes of group q, stored in one instruction
    const uint64_t o16 = P.outPtr + 16 * q;
    for (uint32_t gi = 0; gi < G4; gi += 4) {
        qstage_put(ring, g);
        if (gi) put_16(o16 + 16 * (uint64_t)(gi - 4), o0, o1, o2, o3, wide);
        qstage_load(g, (Pb >> 3) - 12, q);
        wave_sync();
        const uint32_t Q0 = Q;
        o0 = group();
        o1 = group();
        o2 = group();
        o3 = group();
        quad_transpose4(o0, o1, o2, o3, q);
        Pb -= Q0 - Q;
    }
    if (G4) put_16(o16 + 16 * (uint64_t)(G4 - 4), o0, o1, o2, o3, wide);
    for (uint32_t gi = G4; gi < G; gi++) {
        if ((gi & 3) == 0) {
            qstage_step(ring, g, (Pb >> 3) - 12, q);
            wave_sync();
        }
        const uint32_t Q0 = Q;
        const uint32_t d = group();
        Pb -= Q0 - Q;
        put_dword(oq + 16 * (uint64_t)gi, d, wide);
    }
    const uint32_t R = ok ? (n + 3) / 4 : 0u;
    for (uint32_t r = 4 * G; r < R; r++) {
        if ((r & (kQPhase - 1)) == 0) {
            qstage_step(ring, g, (Pb >> 3) - 12, q);
            wave_sync();
        }
        const uint32_t p = 4 * r + q, Q0 = Q;
        const uint32_t y = fse_quad_round<true>(L, ring, o, W, Q, sq, selq, cq, fb, sys, p, n);
        Pb -= Q0 - Q;
        if (p < n) *(g_uint8*)(uintptr_t)(out + p) = (uint8_t)y;
    }
    if (q == 0 && (!ok || Pb != 8 * s)) ws.plans[b0 + f].status = OZL_ERR_CORRUPTION;
}

// ---------------------------------------------------------------- Huffman decoding
// One lane per Huffman stream, 8 frames (32 streams) per one-wave workgroup,
// so that 4 workgroups (one per SIMD) share a CU's LDS; each lane peeks tl
// bits, looks the code up in its frame's table and refills 32 bits when the
// window holds <= 32 (huf_decompress.c:587-787). The lane stages its stream
// into its ring itself: every kHPhase rounds it stores the two 16-byte blocks
// it loaded kHPhase rounds earlier and loads the next two. Blocks may hold
// bytes below the stream start (down to 8 bytes below it) that the window
// peeks but a valid stream never consumes: every table entry whose top bits
// are a code holds that code whatever the bits below.
constexpr int kHF = 8;
constexpr int kHRing = 128;
constexpr int kHLook = 64;
constexpr int kHPhase = 16;     // rounds between staging steps (<= 22 bytes consumed)

struct HufDecLds {
    uint16_t lut[kHF][2048];
    uint32_t ring[4 * kHF][kHRing / 4];
    HufBuild hb;
    ozl_dplan plan[kHF];
};

struct LStage {
    uint64_t aLo;     // lowest address loaded or requested (16-aligned)
    uint64_t f;       // floor16(stream start - 8)
    uint4 p0, p1;     // blocks loaded by the last step
    uint64_t a0, a1;  // their addresses (0: none)
};

template <int R>
__device__ inline void lring_put(uint32_t* ring, uint64_t a, uint4 v)
{
    *reinterpret_cast<uint4*>(ring + ((uint32_t)(a >> 2) & (R / 4 - 1))) = v;
}

// initial fill of [max(floor16(top - LOOK), floor16(lo - 8)), top)
template <int R, int LOOK>
__device__ inline void lstage_init(uint32_t* ring, LStage& g, uint64_t lo, uint64_t top)
{
    g.f = floor16(lo - 8);
    uint64_t a0 = floor16(top - LOOK);
    if (a0 < g.f) a0 = g.f;
    constexpr int NB = LOOK / 16 + 1;
    uint4 v[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) v[k] = gload16(a0 + 16 * k < top ? a0 + 16 * k : a0);
#pragma unroll
    for (int k = 0; k < NB; k++)
        if (a0 + 16 * k < top) lring_put<R>(ring, a0 + 16 * k, v[k]);
    g.aLo = a0;
    g.a0 = g.a1 = 0;
}

template <int R>
__device__ inline void lstage_put(uint32_t* ring, LStage& g)
{
    if (g.a0) lring_put<R>(ring, g.a0, g.p0);
    if (g.a1) lring_put<R>(ring, g.a1, g.p1);
}

template <int LOOK>
__device__ inline void lstage_load(LStage& g, uint64_t winLo)
{
    uint64_t want = floor16(winLo - LOOK);
    if (want < g.f) want = g.f;
    uint64_t nb = g.aLo > want ? (g.aLo - want) >> 4 : 0;
    if (nb > 2) nb = 2;
    const uint64_t x0 = g.aLo - 16, x1 = g.aLo - 32;
    g.p0 = gload16(nb > 0 ? x0 : g.f);
    g.p1 = gload16(nb > 1 ? x1 : g.f);
    g.a0 = nb > 0 ? x0 : 0;
    g.a1 = nb > 1 ? x1 : 0;
    g.aLo -= 16 * nb;
}

template <int R, int LOOK>
__device__ inline void lstage_step(uint32_t* ring, LStage& g, uint64_t winLo)
{
    lstage_put<R>(ring, g);
    lstage_load<LOOK>(g, winLo);
}

__global__ void __launch_bounds__(64) k_huf_decode(uint32_t B, ozl_dws ws)
{
    __shared__ HufDecLds L;
    const int lane = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kHF;
    for (int i = lane; i < kHF * 16; i += 64) {
        int f = i >> 4, k = i & 15;
        reinterpret_cast<uint32_t*>(&L.plan[f])[k] =
            b0 + f < B ? reinterpret_cast<const uint32_t*>(&ws.plans[b0 + f])[k] : 0u;
    }
    wave_sync();
    uint32_t factMask = 0;
    for (int f = 0; f < kHF; f++) {
        const ozl_dplan& P = L.plan[f];
        if (b0 + f < B && P.status == 0 && P.bk == OZL_BK_HUF && P.tl >= 1 && P.tl <= 11) factMask |= 1u << f;
    }
    if (!factMask) return;
    for (int f = 0; f < kHF; f++)
        if ((factMask >> f) & 1) wave_huf_lut(L.lut[f], ws.side + (size_t)(b0 + f) * OZL_DSIDE_BYTES, L.plan[f].tl, L.hb);
    wave_sync();
    const int f = lane >> 2, k = lane & 3;
    if (f >= kHF || !((factMask >> f) & 1)) return;
    const ozl_dplan P = L.plan[f];
    const uint32_t n = P.n, payLen = P.payLen, tl = P.tl;
    const uint64_t s0 = P.srcPtr;
    uint64_t lo = 0;
    uint32_t len = 0, cnt = 0, outOff = 0;
    if (P.nbStates) {
        uint32_t seg = (n + 3) / 4;
        const uint8_t* jt = reinterpret_cast<const uint8_t*>((uintptr_t)s0);
        uint32_t sz0 = 0, sz1 = 0, sz2 = 0;
        if (payLen >= 6) {
            sz0 = (uint32_t)jt[0] | ((uint32_t)jt[1] << 8);
            sz1 = (uint32_t)jt[2] | ((uint32_t)jt[3] << 8);
            sz2 = (uint32_t)jt[4] | ((uint32_t)jt[5] << 8);
        }
        if (payLen < 6 || 6 + sz0 + sz1 + sz2 > payLen) {
            if (k == 0) ws.plans[b0 + f].status = OZL_ERR_CORRUPTION;
            return;
        }
        uint32_t start = 6 + (k > 0 ? sz0 : 0) + (k > 1 ? sz1 : 0) + (k > 2 ? sz2 : 0);
        len = k == 0 ? sz0 : k == 1 ? sz1 : k == 2 ? sz2 : payLen - 6 - sz0 - sz1 - sz2;
        lo = s0 + start;
        cnt = k < 3 ? seg : n - 3 * seg;
        outOff = k * seg;
    } else {
        if (k != 0) return;
        lo = s0;
        len = payLen;
        cnt = n;
    }
    const uint16_t* LUT = L.lut[f];
    uint32_t* ring = L.ring[lane];
    const uint64_t outA = P.outPtr + outOff;
    const bool wide = (outA & 3) == 0;
    LStage g;
    lstage_init<kHRing, kHLook>(ring, g, lo, lo + len);
    wave_sync();
    Win w;
    bool ok = win_init<kHRing>(w, ring, lo, len);
    if (ok) win_refill32(w, ring_dword<kHRing>(ring, lo + (uint64_t)(int64_t)w.lb - 4));
    const uint32_t R = ok ? cnt : 0u;
    auto sym = [&]() -> uint32_t {
        const uint32_t e = LUT[win_hi(w) >> (32 - tl)];
        const uint32_t d = ring_dword<kHRing>(ring, (uint64_t)((int64_t)lo + w.lb) - 4);
        const uint32_t l = e >> 8;
        w.acc <<= l;
        w.valid -= (int)l;
        win_refill32(w, d);
        return e & 255u;
    };
    // two symbols per refill: the window holds >= 33 bits whenever a pair
    // starts (>= 57 after the set-up; a pair takes <= 22 and a refill adds
    // 32 when <= 32 remain), so both lookups see tl valid bits
    auto sym2 = [&]() -> uint32_t {
        const uint32_t d = ring_dword<kHRing>(ring, (uint64_t)((int64_t)lo + w.lb) - 4);
        const uint32_t e0 = LUT[win_hi(w) >> (32 - tl)];
        const uint32_t l0 = e0 >> 8;
        w.acc <<= l0;
        const uint32_t e1 = LUT[win_hi(w) >> (32 - tl)];
        const uint32_t l1 = e1 >> 8;
        w.acc <<= l1;
        w.valid -= (int)(l0 + l1);
        win_refill32(w, d);
        return (e0 & 255u) | ((e1 & 255u) << 8);
    };
    // 16 rounds per staging step, 4 steps per iteration; the iteration's 64
    // bytes are stored at the next iteration's first step, between its ring
    // stores and its loads, so that waiting for a step's loads never waits
    // for output stores (both count in vmcnt, which drains in order), and
    // whole 64-byte pieces reach L2 (16-byte pieces of 4 streams per frame
    // left partly written lines that were written back several times)
    const uint32_t R64 = R & ~63u;
    uint4 ob[4] = { make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0) };
    auto put64 = [&](uint64_t a) {
        if ((a & 15) == 0) {
            typedef __attribute__((address_space(1))) u32x4 g_u32x4;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                u32x4 v;
                v.x = ob[k].x; v.y = ob[k].y; v.z = ob[k].z; v.w = ob[k].w;
                *(g_u32x4*)(uintptr_t)(a + 16 * k) = v;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 4; k++) put4_dwords(a + 16 * k, 4, ob[k].x, ob[k].y, ob[k].z, ob[k].w, wide);
        }
    };
    for (uint32_t r = 0; r < R64; r += 64) {
#pragma unroll
        for (int q4 = 0; q4 < 4; q4++) {
            lstage_put<kHRing>(ring, g);
            if (q4 == 0 && r) put64(outA + r - 64);
            lstage_load<kHLook>(g, (uint64_t)((int64_t)lo + w.lb) - 4);
            wave_sync();
            uint32_t x[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                uint32_t v = sym2();
                v |= sym2() << 16;
                x[j] = v;
            }
            ob[q4] = make_uint4(x[0], x[1], x[2], x[3]);
        }
    }
    if (R64) put64(outA + R64 - 64);
    for (uint32_t r = R64; r < R; r++) {
        if ((r & (kHPhase - 1)) == 0) {
            lstage_step<kHRing, kHLook>(ring, g, (uint64_t)((int64_t)lo + w.lb) - 4);
            wave_sync();
        }
        *(g_uint8*)(uintptr_t)(outA + r) = (uint8_t)sym();
    }
    long pos = 8L * w.lb + w.valid;
    if (!ok || pos != 0) ws.plans[b0 + f].status = OZL_ERR_CORRUPTION;
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

// two waves per frame: wave 0 hashes the compressed region, wave 1 the
// decoded output (no LDS, so that many frames' waves share a CU and keep
// their loads in flight)
__global__ void __launch_bounds__(128) k_dverify(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                 uint32_t B, const uint8_t* __restrict__ dst,
                                                 const uint64_t* __restrict__ doff, ozl_dws ws, int* status)
{
    __shared__ uint32_t hsh[2];
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_dplan d = ws.plans[b];
    int st = (int)d.status;
    const uint8_t* fr = frames + foff[b];
    const int w = (int)(threadIdx.x >> 6);
    if (st == 0) {
        const uint64_t h = w == 0 ? ozl_xxh3_wave(fr + d.chunkOff, d.csumOff + 4 - d.chunkOff, nullptr, 0)
                                  : ozl_xxh3_wave(dst + doff[b], d.n, nullptr, 0);
        if ((threadIdx.x & 63) == 0) hsh[w] = (uint32_t)h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (st == 0) {
            if (hsh[0] != ozl_le32(fr + d.csumOff + 4)) st = OZL_ERR_COMPRESSED_CHECKSUM;
            else if (hsh[1] != ozl_le32(fr + d.csumOff)) st = OZL_ERR_CONTENT_CHECKSUM;
        }
        status[b] = st;
    }
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

// Validates an item's side stream and leaves its normalized counts / code
// lengths in the workspace, like k_dparse does for batched frames.
__global__ void __launch_bounds__(64) k_prep_items(const ozl_ditem* __restrict__ items, uint32_t B, ozl_dws ws)
{
    __shared__ int16_t norm[256];
    __shared__ uint8_t len[256];
    __shared__ __attribute__((aligned(16))) uint8_t pre[OZL_DPREFIX + 16];
    __shared__ ozl_dplan sD;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_ditem it = items[b];
    // hipcc (ROCm 7.2, gfx950) selected d.bk below on a stale SCC when `huf`
    // came from a byte load it treated as uniform; reading it through
    // readfirstlane makes the comparison a scalar one.
    const uint32_t huf = __builtin_amdgcn_readfirstlane((uint32_t)it.huf);
    const uint8_t* side = (const uint8_t*)(uintptr_t)it.side;
    const uint32_t sideN = it.sideN;
    const bool nc = !huf && (sideN & 0x80000000u);
    // the side stream (NCount header, raw int16 counts or weights) copied to
    // LDS by the wave as words, zeros past it, when it fits the copy
    const uint64_t sideBytes = huf ? (sideN < 256u ? sideN : 256u) : nc ? (sideN & 0x7FFFFFFFu) : 2ull * sideN;
    const bool inLds = sideBytes <= OZL_DPREFIX;
    {
        uint32_t* pw = reinterpret_cast<uint32_t*>(pre);
        for (uint32_t i = threadIdx.x; i < (OZL_DPREFIX + 16) / 4; i += 64) {
            uint32_t x = 0;
            for (uint32_t k = 0; k < 4; k++) {
                const uint64_t j = 4ull * i + k;
                if (inLds && j < sideBytes) x |= (uint32_t)side[j] << (8 * k);
            }
            pw[i] = x;
        }
        for (int s = threadIdx.x; s < 256; s += 64) norm[s] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ozl_dplan d = {};
        d.n = it.n;
        d.payLen = it.bitsLen;
        d.nbStates = it.nbStates;
        d.srcPtr = it.bits;
        d.outPtr = it.out;
        d.bk = huf ? OZL_BK_HUF : OZL_BK_FSE;
        if (!huf) {
            uint32_t sum = 0, cnt = sideN;
            bool bad = false;
            const uint8_t* sp = inLds ? pre : side;
            if (nc) {
                // the side stream is still an NCount header (fse_compress.c:246-357)
                const uint32_t nlen = sideN & 0x7FFFFFFFu;
                unsigned ms = 0, ntl = 0;
                const ozl_fview fv = { side, pre, inLds ? nlen : 0u, reinterpret_cast<const uint32_t*>(pre) };
                const long r = ozl_dread_ncount(norm, &ms, &ntl, fv, 0, nlen, true);
                bad = r < 0 || (uint32_t)r != nlen;
                cnt = bad ? 0u : ms + 1;
                for (uint32_t s = 0; !bad && s < cnt; s++) sum += norm[s] == -1 ? 1 : (uint32_t)norm[s];
            } else {
                bad = cnt == 0 || cnt > 256;
                for (uint32_t s = 0; !bad && s < cnt; s++) {
                    norm[s] = (int16_t)(sp[2 * s] | (sp[2 * s + 1] << 8));
                    if (norm[s] < -1) bad = true;
                    sum += norm[s] == -1 ? 1 : (uint32_t)norm[s];
                }
            }
            unsigned tl = sum ? (unsigned)ozl_hb32(sum) : 0;
            if (bad || sum == 0 || (sum & (sum - 1)) || tl < OZL_FSE_MIN_TABLELOG || tl > 11 ||
                (it.nbStates != 2 && it.nbStates != 4))
                d.status = OZL_ERR_CORRUPTION;
            else {
                d.tl = (uint8_t)tl;
                d.aux = cnt - 1;
            }
        } else if (sideN > 256 || it.nbStates > 1 || (it.nbStates == 1 && it.n < 12)) {
            d.status = OZL_ERR_CORRUPTION;
        }
        sD = d;
    }
    __syncthreads();
    if (huf && sD.status == 0) {
        // code lengths from the weights (ozl_huf_from_weights): a weight
        // above 12 or a sum of 2^(w-1) that is not a power of two is corrupt
        uint32_t tot = 0, bad = 0;
        uint8_t wl[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t sy = threadIdx.x + 64u * k;
            const uint32_t wv = sy < sideN ? pre[sy] : 0u;
            wl[k] = (uint8_t)wv;
            bad |= wv > 12 ? 1u : 0u;
            tot += wv && wv <= 12 ? 1u << (wv - 1) : 0u;
        }
#pragma unroll
        for (int dd = 32; dd >= 1; dd >>= 1) {
            tot += __shfl_xor(tot, dd, 64);
            bad |= __shfl_xor(bad, dd, 64);
        }
        const uint32_t tl = tot ? (uint32_t)ozl_hb32(tot) : 0u;
        const bool okW = !bad && tot && (tot & (tot - 1)) == 0 && tl >= 1 && tl <= 11;
#pragma unroll
        for (int k = 0; k < 4; k++) len[threadIdx.x + 64 * k] = okW && wl[k] ? (uint8_t)(tl + 1 - wl[k]) : 0;
        if (threadIdx.x == 0) {
            if (okW) sD.tl = (uint8_t)tl;
            else sD.status = OZL_ERR_CORRUPTION;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ws.plans[b] = sD;
    const int bk = sD.status == 0 ? sD.bk : 255;
    uint8_t* sd = ws.side + (size_t)b * OZL_DSIDE_BYTES;
    if (bk == OZL_BK_FSE)
        for (int s = threadIdx.x; s < 256; s += 64) reinterpret_cast<int16_t*>(sd)[s] = norm[s];
    else if (bk == OZL_BK_HUF)
        for (int s = threadIdx.x; s < 256; s += 64) sd[s] = len[s];
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
    return (size_t)B * (sizeof(ozl_dplan) + OZL_DSIDE_BYTES) + 4096;
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
    ws.side = w;
    const uint8_t* f = (const uint8_t*)frames;
    uint8_t* o = (uint8_t*)dst;
    OZL_LAUNCH("k_dparse", k_dparse, dim3(B), dim3(kParseT), 0, st, f, foff, B, o, doff, ws);
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3((B + kFQ - 1) / kFQ), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3((B + kHF - 1) / kHF), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_pack_decode", k_pack_decode, dim3(B), dim3(kT), 0, st, f, foff, B, o, doff, ws);
    OZL_LAUNCH("k_dverify", k_dverify, dim3(B), dim3(128), 0, st, f, foff, B, (const uint8_t*)o, doff, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- allocation-free batch decode
// ozl_entropy_decompress_batch's workspace (one plan and one side table per
// frame) for at most OZL_NOALLOC_MAX_FRAMES frames, as a file-scope device
// array: each GPU the library runs on holds its own 9 MiB copy until the
// process ends. Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): the device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
#define OZL_NOALLOC_MAX_FRAMES 16384u
#define OZL_NOALLOC_MAX_DEVICES 64
constexpr size_t kNoAllocBytes = (size_t)OZL_NOALLOC_MAX_FRAMES * (sizeof(ozl_dplan) + OZL_DSIDE_BYTES) + 4096;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
std::mutex g_scratchMu;
void* g_scratchPtr[OZL_NOALLOC_MAX_DEVICES];   // device address of `scratch` per device

int noalloc_scratch(void** out)
{
    int dev = 0;
    OZL_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= OZL_NOALLOC_MAX_DEVICES) return OZL_ERR_PARAMETER_INVALID;
    std::lock_guard<std::mutex> g(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        OZL_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return OZL_ERR_HIP;
        g_scratchPtr[dev] = p;
    }
    *out = g_scratchPtr[dev];
    return OZL_OK;
}
}  // namespace

extern "C" {

// ozl_entropy_decompress_batch with the library's own scratch instead of a
// workspace argument: no allocation of any kind, for callers that cannot
// allocate. nbFrames above OZL_NOALLOC_MAX_FRAMES is internalBuffer_tooSmall.
// Calls that overlap on one GPU share its scratch and race.
int ozl_entropy_decompress_batch_noalloc(const void* frames, const uint64_t* frameOffsets, uint32_t nbFrames,
                                         void* dst, const uint64_t* dstOffsets, int* status, void* stream)
{
    if (nbFrames == 0) return OZL_OK;
    if (nbFrames > OZL_NOALLOC_MAX_FRAMES) return OZL_ERR_INTERNAL_BUFFER;
    void* ws = nullptr;
    int rc = noalloc_scratch(&ws);
    if (rc) return rc;
    // every plan and side table the kernels read is written by k_dparse in
    // the same call: nothing of the previous call's contents is read
    return ozl_entropy_decompress_batch(frames, frameOffsets, nbFrames, dst, dstOffsets, status, ws, kNoAllocBytes,
                                        stream);
}

size_t ozl_entropy_decompress_noalloc_max_frames(void) { return OZL_NOALLOC_MAX_FRAMES; }

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
    ws.side = w;
    OZL_LAUNCH("k_prep_items", k_prep_items, dim3(B), dim3(64), 0, st, items, B, ws);
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3((B + kFQ - 1) / kFQ), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3((B + kHF - 1) / kHF), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_item_status", k_item_status, dim3((B + 63) / 64), dim3(64), 0, st, B, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
