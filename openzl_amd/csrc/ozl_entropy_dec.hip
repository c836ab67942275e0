// Batched decoding of FSE / Huffman graph frames (the universal decompressor
// path of src/openzl/decompress/decompress2.c:1442-1580 restricted to the
// single-chunk frames ZL_GRAPH_FSE and ZL_GRAPH_HUFFMAN produce for inputs
// below 100000 bytes).
//
//   k_dparse       one wave per frame: the first KiB of the frame is copied to
//                  LDS, lane 0 parses it (ozl_dparse.h) and leaves the
//                  normalized counts or the Huffman code lengths of the frame
//                  in the workspace
//   k_fse_decode   4 frames per 64-lane wave: the wave builds each frame's
//                  decoding table in LDS (fse_decompress.c:73-176), then one
//                  lane per interleaved state decodes (:255-338); the lanes of
//                  a frame share one bit window, their read offsets come from
//                  a quad prefix sum (DPP)
//   k_huf_decode   4 frames per wave: the wave builds each frame's lookup
//                  table from the code lengths, then one lane per Huffman
//                  stream decodes (huf_decompress.c:587-787)
//   k_pack_decode  store / constant / flatpack / bitpack
//   k_dverify      one wavefront per frame: content and compressed checksums
//
// Bit streams are read backwards through a 64-bit window per lane. Its bytes
// come from a ring in LDS that every lane of the wave refills once per phase
// of 32 rounds with aligned 16-byte loads issued one phase ahead of use, so
// no global-memory latency sits on a decoding lane's dependency chain.
// Decoded bytes leave in 4-byte stores.
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
constexpr int kPrefix = 1024;

__global__ void __launch_bounds__(kParseT) k_dparse(const uint8_t* __restrict__ frames, const uint64_t* __restrict__ foff,
                                                    uint32_t B, uint8_t* __restrict__ dst, const uint64_t* __restrict__ doff,
                                                    ozl_dws ws)
{
    __shared__ ozl_dlds L;
    __shared__ uint8_t pre[kPrefix];
    __shared__ int sBk;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    const uint8_t* fr = frames + foff[b];
    size_t flen = foff[b + 1] - foff[b];
    size_t pn = flen < (size_t)kPrefix ? flen : (size_t)kPrefix;
    for (size_t i = threadIdx.x; i < pn; i += kParseT) pre[i] = fr[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        ozl_dplan d;
        ozl_fview v = { fr, pre, pn };
        ozl_dparse_frame(v, flen, L, d);
        d.srcPtr = (uint64_t)(uintptr_t)(fr + d.payOff);
        d.outPtr = (uint64_t)(uintptr_t)(dst + doff[b]);
        if (d.status == 0 && doff[b + 1] - doff[b] != d.n) d.status = OZL_ERR_DSTCAPACITY_TOOSMALL;
        ws.plans[b] = d;
        sBk = d.status == 0 ? d.bk : 255;
    }
    __syncthreads();
    uint8_t* side = ws.side + (size_t)b * OZL_DSIDE_BYTES;
    if (sBk == OZL_BK_FSE)
        for (int s = threadIdx.x; s < 256; s += kParseT) reinterpret_cast<int16_t*>(side)[s] = L.norm[s];
    else if (sBk == OZL_BK_HUF)
        for (int s = threadIdx.x; s < 256; s += kParseT) side[s] = L.wlen[s];
}

// ---------------------------------------------------------------- LDS rings of stream bytes
// A ring of R bytes holds stream bytes by absolute address (slot = address
// mod R). read8 returns the 8 bytes [A-8, A) as a little-endian u64.
template <int R>
__device__ inline uint64_t ring_read8(const uint32_t* ring, uint64_t A)
{
    constexpr uint32_t W = R / 4 - 1;
    uint64_t a0 = A - 8;
    uint32_t w = (uint32_t)(a0 >> 2);
    uint32_t d0 = ring[w & W], d1 = ring[(w + 1) & W], d2 = ring[(w + 2) & W];
    uint32_t sh = (uint32_t)(a0 & 3) * 8;
    uint64_t lo = (uint64_t)d0 | ((uint64_t)d1 << 32);
    return sh ? (lo >> sh) | ((uint64_t)d2 << (64 - sh)) : lo;
}

// Stages the 16-byte aligned block at absolute address `a` (loaded earlier into
// `v`) into a ring; bytes below `lo` (the stream start) read as zero.
template <int R>
__device__ inline void ring_store16(uint32_t* ring, uint64_t a, uint4 v, uint64_t lo)
{
    constexpr uint32_t W = R / 4 - 1;
    uint32_t d[4] = { v.x, v.y, v.z, v.w };
    uint32_t w = (uint32_t)(a >> 2);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint64_t da = a + 4 * k;
        uint32_t x = d[k];
        if (da + 4 <= lo) x = 0;
        else if (da < lo) x &= ~0u << (8 * (uint32_t)(lo - da));
        ring[(w + k) & W] = x;
    }
}

// Loads the aligned block at `a` unless it lies wholly below the stream start.
__device__ inline uint4 load_block(uint64_t a, uint64_t lo)
{
    if (a + 16 <= lo) return make_uint4(0, 0, 0, 0);
    return *reinterpret_cast<const uint4*>((uintptr_t)a);
}

// Backward bit window: `acc` holds the `valid` stream bits just below bit
// position 8*lb + valid, most recent first (MSB-aligned); lb is the byte
// offset of the window's low end inside the stream (may go below 0, where the
// stream reads as zeros).
struct BitWin {
    uint64_t acc;
    int valid;
    int lb;
};

// Refill from the 8 bytes below the window (`v`, read before the window last
// moved); applied only when `need` (select-based, no branch).
__device__ inline void win_refill_with(BitWin& w, uint64_t v, bool need)
{
    int k = (64 - w.valid) >> 3;
    if (w.lb <= 0) v = 0;
    uint64_t keep = k >= 8 ? ~0ull : (k <= 0 ? 0ull : ~((1ull << (64 - 8 * k)) - 1));
    uint64_t top = v & keep;
    uint64_t acc = w.acc | (w.valid >= 64 ? 0ull : (w.valid ? (top >> w.valid) : top));
    w.acc = need ? acc : w.acc;
    w.valid = need ? w.valid + 8 * k : w.valid;
    w.lb = need ? w.lb - k : w.lb;
}

template <int R>
__device__ inline void win_refill(BitWin& w, const uint32_t* ring, uint64_t s)
{
    int k = (64 - w.valid) >> 3;
    uint64_t v = w.lb > 0 ? ring_read8<R>(ring, s + (uint64_t)w.lb) : 0;
    uint64_t keep = k >= 8 ? ~0ull : (k <= 0 ? 0ull : ~((1ull << (64 - 8 * k)) - 1));
    uint64_t top = v & keep;
    w.acc |= w.valid >= 64 ? 0ull : (w.valid ? (top >> w.valid) : top);
    w.valid += 8 * k;
    w.lb -= k;
}

// ---------------------------------------------------------------- FSE decoding
constexpr int kFF = 4;          // frames per workgroup, 4 lanes each
constexpr int kFRing = 512;     // staged stream bytes per frame
constexpr int kFLook = 384;     // bytes kept staged below a window
constexpr int kPhase = 32;      // rounds between staging steps

struct FseDecLds {
    uint32_t dt[kFF][2048];
    uint32_t ring[kFF][kFRing / 4];
    FseBuild bs;
    ozl_dplan plan[kFF];
    uint64_t aLo[kFF];    // lowest staged address (16-aligned)
    uint64_t winLo[kFF];  // window bottom at the start of the phase
};

__device__ inline uint64_t floor16(uint64_t a) { return a & ~15ull; }

__global__ void __launch_bounds__(64) k_fse_decode(uint32_t B, ozl_dws ws)
{
    __shared__ FseDecLds L;
    const int lane = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kFF;
    {
        int f = lane >> 4, k = lane & 15;
        reinterpret_cast<uint32_t*>(&L.plan[f])[k] =
            b0 + f < B ? reinterpret_cast<const uint32_t*>(&ws.plans[b0 + f])[k] : 0u;
    }
    __syncthreads();
    uint32_t factMask = 0;
    uint32_t R = 0;
#pragma unroll
    for (int f = 0; f < kFF; f++) {
        const ozl_dplan& P = L.plan[f];
        if (b0 + f < B && P.status == 0 && P.bk == OZL_BK_FSE) {
            factMask |= 1u << f;
            uint32_t rf = (P.n + P.nbStates - 1) / P.nbStates;
            R = rf > R ? rf : R;
        }
    }
    if (R == 0) return;
#pragma unroll
    for (int f = 0; f < kFF; f++) {
        if (!((factMask >> f) & 1)) continue;
        const ozl_dplan& P = L.plan[f];
        wave_fse_dtable(L.dt[f], reinterpret_cast<const int16_t*>(ws.side + (size_t)(b0 + f) * OZL_DSIDE_BYTES), P.aux,
                        P.tl, L.bs);
        uint64_t s = P.srcPtr, top = s + P.payLen;
        uint64_t a0 = floor16(top - kFLook), lo = floor16(s - 8);
        if (a0 < lo) a0 = lo;
        uint8_t* rb = reinterpret_cast<uint8_t*>(L.ring[f]);
        for (uint64_t a = a0 + lane; a < top; a += 64) rb[a & (kFRing - 1)] = a >= s ? *reinterpret_cast<const uint8_t*>((uintptr_t)a) : 0;
        if (lane == 0) L.aLo[f] = a0;
    }
    __syncthreads();

    // decoding lanes: frame fd, state q
    const int fd = lane < 4 * kFF ? lane >> 2 : 0, q = lane & 3;
    const bool dl = lane < 4 * kFF && ((factMask >> fd) & 1);
    const ozl_dplan& P = L.plan[fd];
    const uint32_t S = dl ? P.nbStates : 4u, n = P.n, tl = P.tl;
    const uint32_t* DT = L.dt[fd];
    const uint32_t* ring = L.ring[fd];
    const uint64_t s = P.srcPtr;
    uint8_t* out = reinterpret_cast<uint8_t*>((uintptr_t)P.outPtr);
    const bool wide = dl && S == 4 && (P.outPtr & 3) == 0;
    BitWin w = { 0, 0, (int)P.payLen };
    uint32_t state = 0;
    bool ok = dl;
    if (dl) {
        win_refill<kFRing>(w, ring, s);
        uint32_t topByte = (uint32_t)(w.acc >> 56);
        if (topByte == 0) ok = false;
        else {
            int skip = 8 - ozl_hb32(topByte);
            w.acc <<= skip;
            w.valid -= skip;
            if ((uint32_t)q < S) state = (uint32_t)((w.acc << (q * tl)) >> (64 - tl));
            w.acc <<= S * tl;
            w.valid -= (int)(S * tl);
            win_refill<kFRing>(w, ring, s);
        }
    }
    bool pendOn = false;
    uint64_t pa = 0;
    int pf = 0;
    uint4 pend = make_uint4(0, 0, 0, 0);
    uint32_t wacc = 0;
    const uint32_t R4 = (R + 3) & ~3u;
    for (uint32_t r = 0; r < R4; r++) {
        if (r % kPhase == 0) {
            if (dl && q == 0) L.winLo[fd] = (uint64_t)((int64_t)s + w.lb);
            __syncthreads();
            if (pendOn) ring_store16<kFRing>(L.ring[pf], pa, pend, L.plan[pf].srcPtr);
            const int sf = lane >> 4, si = lane & 15;
            pendOn = false;
            uint64_t newLo = 0;
            const bool sa = (factMask >> sf) & 1;
            if (sa) {
                uint64_t sb = L.plan[sf].srcPtr, cur = L.aLo[sf];
                uint64_t floorLo = floor16(sb - 8), want = floor16(L.winLo[sf] - kFLook);
                if (want < floorLo) want = floorLo;
                uint64_t a = cur - 16 * (uint64_t)(si + 1);
                if (a >= want && a < cur) {
                    pendOn = true;
                    pa = a;
                    pf = sf;
                    pend = load_block(a, sb);
                }
                uint64_t lim = cur - 16 * 16;
                newLo = want < cur ? (want > lim ? want : lim) : cur;
            }
            __syncthreads();
            if (sa && si == 0) L.aLo[sf] = newLo;
        }
        // the table entry and the bytes below the window are read together;
        // the window holds >= 44 bits at the top of every round
        uint32_t p = r * S + (uint32_t)q;
        bool mine = ok && (uint32_t)q < S && p < n;
        uint32_t e = DT[state];
        uint64_t below = ring_read8<kFRing>(ring, s + (uint64_t)(int64_t)w.lb);
        e = mine ? e : 0u;
        bool upd = mine && p + S < n;
        uint32_t nb = upd ? ((e >> 8) & 255u) : 0u;
        uint32_t n0 = quad_lane<0>(nb), n1 = quad_lane<1>(nb), n2 = quad_lane<2>(nb), n3 = quad_lane<3>(nb);
        uint32_t before = (q > 0 ? n0 : 0u) + (q > 1 ? n1 : 0u) + (q > 2 ? n2 : 0u);
        uint32_t all = n0 + n1 + n2 + n3;
        uint64_t hi = w.acc << before;
        uint32_t bits = nb ? (uint32_t)(hi >> (64 - nb)) : 0u;
        state = upd ? (e >> 16) + bits : state;
        w.acc <<= all;
        w.valid -= (int)all;
        win_refill_with(w, below, ok && w.valid < 44);
        if (wide) wacc |= (e & 255u) << (8 * (r & 3));
        else if (mine) out[p] = (uint8_t)e;
        if ((r & 3) == 3) {
            uint32_t t0 = quad_lane<0>(wacc), t1 = quad_lane<1>(wacc), t2 = quad_lane<2>(wacc), t3 = quad_lane<3>(wacc);
            if (wide && ok) {
                uint32_t r0 = r - 3, sh = 8 * (uint32_t)q;
                if ((r0 + 4) * 4 <= n) {
                    uint32_t tw = ((t0 >> sh) & 255u) | (((t1 >> sh) & 255u) << 8) | (((t2 >> sh) & 255u) << 16) |
                                  (((t3 >> sh) & 255u) << 24);
                    *reinterpret_cast<uint32_t*>(out + (r0 + (uint32_t)q) * 4) = tw;
                } else {
                    for (uint32_t j = 0; j < 4; j++) {
                        uint32_t p2 = (r0 + j) * 4 + (uint32_t)q;
                        if (p2 < n) out[p2] = (uint8_t)(wacc >> (8 * j));
                    }
                }
            }
            wacc = 0;
        }
    }
    if (dl && q == 0) {
        long pos = 8L * w.lb + w.valid;
        if (!ok || pos != 0) ws.plans[b0 + fd].status = OZL_ERR_CORRUPTION;
    }
}

// ---------------------------------------------------------------- Huffman decoding
constexpr int kHF = 4;          // frames per workgroup, one lane per stream
constexpr int kHS = 4 * kHF;    // streams per workgroup
constexpr int kHRing = 256;     // staged stream bytes per stream
constexpr int kHLook = 160;

struct HufDecLds {
    uint16_t lut[kHF][2048];
    uint32_t ring[kHS][kHRing / 4];
    HufBuild hb;
    ozl_dplan plan[kHF];
    uint64_t sLo[kHS];    // stream start
    uint64_t sTop[kHS];   // stream end
    uint64_t aLo[kHS];
    uint64_t winLo[kHS];
    uint32_t sAct[kHS];
};

__global__ void __launch_bounds__(64) k_huf_decode(uint32_t B, ozl_dws ws)
{
    __shared__ HufDecLds L;
    const int lane = threadIdx.x;
    const uint32_t b0 = blockIdx.x * kHF;
    {
        int f = lane >> 4, k = lane & 15;
        reinterpret_cast<uint32_t*>(&L.plan[f])[k] =
            b0 + f < B ? reinterpret_cast<const uint32_t*>(&ws.plans[b0 + f])[k] : 0u;
    }
    __syncthreads();
    uint32_t factMask = 0;
#pragma unroll
    for (int f = 0; f < kHF; f++) {
        const ozl_dplan& P = L.plan[f];
        if (b0 + f < B && P.status == 0 && P.bk == OZL_BK_HUF && P.tl >= 1 && P.tl <= 11) factMask |= 1u << f;
    }
    if (!factMask) return;
    // streams: lane k < kHS is stream k & 3 of frame k >> 2
    uint32_t cnt = 0, outOff = 0;
    bool sok = false;
    if (lane < kHS) {
        int f = lane >> 2, k = lane & 3;
        const ozl_dplan& P = L.plan[f];
        uint64_t s = P.srcPtr;
        uint32_t n = P.n, payLen = P.payLen;
        uint64_t lo = 0, top = 0;
        if ((factMask >> f) & 1) {
            if (P.nbStates) {
                uint32_t seg = (n + 3) / 4;
                const uint8_t* jt = reinterpret_cast<const uint8_t*>((uintptr_t)s);
                uint32_t sz[3] = { 0, 0, 0 };
                if (payLen >= 6)
                    for (int j = 0; j < 3; j++) sz[j] = (uint32_t)jt[2 * j] | ((uint32_t)jt[2 * j + 1] << 8);
                uint32_t sum = sz[0] + sz[1] + sz[2];
                if (payLen >= 6 && 6 + sum <= payLen) {
                    uint32_t start = 6;
                    for (int j = 0; j < k; j++) start += sz[j];
                    uint32_t len = k < 3 ? sz[k] : payLen - 6 - sum;
                    lo = s + start;
                    top = lo + len;
                    cnt = k < 3 ? seg : n - 3 * seg;
                    outOff = k * seg;
                    sok = true;
                } else if (k == 0) {
                    ws.plans[b0 + f].status = OZL_ERR_CORRUPTION;
                }
            } else if (k == 0) {
                lo = s;
                top = s + payLen;
                cnt = n;
                sok = true;
            }
        }
        L.sLo[lane] = lo;
        L.sTop[lane] = top;
        L.sAct[lane] = sok;
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < kHF; f++)
        if ((factMask >> f) & 1)
            wave_huf_lut(L.lut[f], ws.side + (size_t)(b0 + f) * OZL_DSIDE_BYTES, L.plan[f].tl, L.hb);
    for (int k = 0; k < kHS; k++) {
        if (!L.sAct[k]) continue;
        uint64_t s = L.sLo[k], top = L.sTop[k];
        uint64_t a0 = floor16(top - kHLook), lo = floor16(s - 8);
        if (a0 < lo) a0 = lo;
        uint8_t* rb = reinterpret_cast<uint8_t*>(L.ring[k]);
        for (uint64_t a = a0 + lane; a < top; a += 64) rb[a & (kHRing - 1)] = a >= s ? *reinterpret_cast<const uint8_t*>((uintptr_t)a) : 0;
        if (lane == 0) L.aLo[k] = a0;
    }
    __syncthreads();

    const int sid = lane < kHS ? lane : 0;
    const int fd = sid >> 2;
    const bool dl = lane < kHS && sok;
    const uint32_t tl = L.plan[fd].tl;
    const uint16_t* LUT = L.lut[fd];
    const uint32_t* ring = L.ring[sid];
    const uint64_t s = L.sLo[sid];
    uint8_t* out = reinterpret_cast<uint8_t*>((uintptr_t)L.plan[fd].outPtr) + outOff;
    const bool wide = dl && (((uintptr_t)out) & 3) == 0;
    BitWin w = { 0, 0, (int)(L.sTop[sid] - s) };
    bool ok = dl;
    if (dl) {
        win_refill<kHRing>(w, ring, s);
        uint32_t topByte = (uint32_t)(w.acc >> 56);
        if (topByte == 0) ok = false;
        else {
            int skip = 8 - ozl_hb32(topByte);
            w.acc <<= skip;
            w.valid -= skip;
        }
    }
    uint32_t R = 0;
    for (int k = 0; k < kHS; k++) {
        // the longest stream of the workgroup sets the loop length
        uint32_t c = __shfl(cnt, k, 64);
        R = c > R ? c : R;
    }
    bool pendOn = false;
    uint64_t pa = 0;
    int pk = 0;
    uint4 pend = make_uint4(0, 0, 0, 0);
    uint32_t wacc = 0;
    const uint32_t R4 = (R + 3) & ~3u;
    for (uint32_t r = 0; r < R4; r++) {
        if (r % kPhase == 0) {
            if (dl) L.winLo[sid] = (uint64_t)((int64_t)s + w.lb);
            __syncthreads();
            if (pendOn) ring_store16<kHRing>(L.ring[pk], pa, pend, L.sLo[pk]);
            const int st = lane >> 2, si = lane & 3;
            pendOn = false;
            uint64_t newLo = 0;
            bool sa = L.sAct[st] != 0;
            if (sa) {
                uint64_t sb = L.sLo[st], cur = L.aLo[st];
                uint64_t floorLo = floor16(sb - 8), want = floor16(L.winLo[st] - kHLook);
                if (want < floorLo) want = floorLo;
                uint64_t a = cur - 16 * (uint64_t)(si + 1);
                if (a >= want && a < cur) {
                    pendOn = true;
                    pa = a;
                    pk = st;
                    pend = load_block(a, sb);
                }
                uint64_t lim = cur - 16 * 4;
                newLo = want < cur ? (want > lim ? want : lim) : cur;
            }
            __syncthreads();
            if (sa && si == 0) L.aLo[st] = newLo;
        }
        // lookup and the bytes below the window are read together; the
        // window holds >= tl bits at the top of every round
        bool mine = ok && r < cnt;
        uint32_t e = LUT[(uint32_t)(w.acc >> (64 - tl))];
        uint64_t below = ring_read8<kHRing>(ring, s + (uint64_t)(int64_t)w.lb);
        uint32_t l = mine ? e >> 8 : 0u;
        w.acc <<= l;
        w.valid -= (int)l;
        win_refill_with(w, below, ok && w.valid < (int)tl);
        if (wide) wacc |= (e & 255u) << (8 * (r & 3));
        else if (mine) out[r] = (uint8_t)e;
        if (wide && (r & 3) == 3) {
            uint32_t r0 = r - 3;
            if (r0 + 4 <= cnt) *reinterpret_cast<uint32_t*>(out + r0) = wacc;
            else
                for (uint32_t j = 0; j < 4; j++)
                    if (r0 + j < cnt) out[r0 + j] = (uint8_t)(wacc >> (8 * j));
            wacc = 0;
        }
    }
    if (lane < kHS && L.sAct[lane]) {
        long pos = 8L * w.lb + w.valid;
        if (!ok || pos != 0) ws.plans[b0 + fd].status = OZL_ERR_CORRUPTION;
    }
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

// Validates an item's side stream and leaves its normalized counts / code
// lengths in the workspace, like k_dparse does for batched frames.
__global__ void __launch_bounds__(64) k_prep_items(const ozl_ditem* __restrict__ items, uint32_t B, ozl_dws ws)
{
    __shared__ int16_t norm[256];
    __shared__ uint8_t w[256];
    __shared__ uint8_t len[256];
    __shared__ uint16_t val[256];
    __shared__ int sBk;
    uint32_t b = blockIdx.x;
    if (b >= B) return;
    ozl_ditem it = items[b];
    // hipcc (ROCm 7.2, gfx950) selected d.bk below on a stale SCC when `huf`
    // came from a byte load it treated as uniform; reading it through
    // readfirstlane makes the comparison a scalar one.
    const uint32_t huf = __builtin_amdgcn_readfirstlane((uint32_t)it.huf);
    if (threadIdx.x == 0) {
        ozl_dplan d = {};
        d.n = it.n;
        d.payLen = it.bitsLen;
        d.nbStates = it.nbStates;
        d.srcPtr = it.bits;
        d.outPtr = it.out;
        d.bk = huf ? OZL_BK_HUF : OZL_BK_FSE;
        const uint8_t* side = (const uint8_t*)(uintptr_t)it.side;
        if (!huf) {
            uint32_t sum = 0;
            bool bad = it.sideN == 0 || it.sideN > 256;
            for (int s = 0; s < 256; s++) norm[s] = 0;
            for (uint32_t s = 0; !bad && s < it.sideN; s++) {
                norm[s] = (int16_t)(side[2 * s] | (side[2 * s + 1] << 8));
                if (norm[s] < -1) bad = true;
                sum += norm[s] == -1 ? 1 : (uint32_t)norm[s];
            }
            unsigned tl = sum ? (unsigned)ozl_hb32(sum) : 0;
            if (bad || sum == 0 || (sum & (sum - 1)) || tl < OZL_FSE_MIN_TABLELOG || tl > 11 ||
                (it.nbStates != 2 && it.nbStates != 4))
                d.status = OZL_ERR_CORRUPTION;
            else {
                d.tl = (uint8_t)tl;
                d.aux = it.sideN - 1;
            }
        } else {
            for (uint32_t s = 0; s < it.sideN && s < 256; s++) w[s] = side[s];
            int tl = it.sideN <= 256 ? ozl_huf_from_weights(len, val, w, it.sideN) : -1;
            if (tl < 1 || tl > 11 || it.nbStates > 1 || (it.nbStates == 1 && it.n < 12)) d.status = OZL_ERR_CORRUPTION;
            else d.tl = (uint8_t)tl;
        }
        ws.plans[b] = d;
        sBk = d.status == 0 ? d.bk : 255;
    }
    __syncthreads();
    uint8_t* sd = ws.side + (size_t)b * OZL_DSIDE_BYTES;
    if (sBk == OZL_BK_FSE)
        for (int s = threadIdx.x; s < 256; s += 64) reinterpret_cast<int16_t*>(sd)[s] = norm[s];
    else if (sBk == OZL_BK_HUF)
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
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3((B + kFF - 1) / kFF), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3((B + kHF - 1) / kHF), dim3(64), 0, st, B, ws);
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
    ws.side = w;
    OZL_LAUNCH("k_prep_items", k_prep_items, dim3(B), dim3(64), 0, st, items, B, ws);
    OZL_LAUNCH("k_fse_decode", k_fse_decode, dim3((B + kFF - 1) / kFF), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_huf_decode", k_huf_decode, dim3((B + kHF - 1) / kHF), dim3(64), 0, st, B, ws);
    OZL_LAUNCH("k_item_status", k_item_status, dim3((B + 63) / 64), dim3(64), 0, st, B, ws, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
