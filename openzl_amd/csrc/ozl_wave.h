// ozl_wave.h -- wave-level (64-lane) building blocks of the entropy kernels:
// lane masks and key matching with ballots, quad broadcasts (DPP), and the
// parallel construction of FSE encoding / decoding tables and Huffman lookup
// tables in LDS. Device only; every function is called by all 64 lanes of one
// wave in uniform control flow (the other waves of its workgroup may idle).
#pragma once
#include "ozl_core.h"

namespace ozl_wave {

__device__ inline uint64_t lanes_below() { return (1ull << (threadIdx.x & 63)) - 1; }

// Orders LDS accesses among the lanes of one wave (a wave's LDS operations
// execute in order; this keeps the compiler from moving them across).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// lanes among `act` whose 8-bit key equals this lane's
__device__ inline uint64_t match8(uint32_t key, bool act)
{
    uint64_t m = __ballot(act);
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
        uint64_t on = __ballot(act && ((key >> bit) & 1));
        m &= ((key >> bit) & 1) ? on : ~on;
    }
    return m;
}

// value of lane K of this lane's quad (DPP quad_perm; every lane must run it)
template <int K>
__device__ inline uint32_t quad_lane(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, K * 0x55, 0xF, 0xF, false);
}

// lanes of this lane's quad exchanged pairwise (DPP quad_perm [1,0,3,2]) and
// in pairs of pairs ([2,3,0,1])
__device__ inline uint32_t quad_xor1(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false); }
__device__ inline uint32_t quad_xor2(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false); }

// ---------------------------------------------------------------- FSE decoding table
struct FseBuildHdr {
    int16_t norm[256];
    uint16_t cum[256];   // exclusive prefix of the positive counts
    uint16_t next[256];  // per symbol: running rank base (decoder) or state slot (encoder)
};
struct FseBuild : FseBuildHdr {
    uint8_t cell[2048];  // spread, when the caller keeps none of its own
};

// Counts of a spread (fse_compress.c:93-155, fse_decompress.c:93-160):
// leaves the counts in S.norm (zero above maxSym), per symbol the exclusive
// prefix of the positive counts in S.cum and, in S.next, the exclusive
// prefix of all counts (-1 counting as 1) when `prefixAll`, else the count
// itself (-1 as 1); writes the low-probability symbols (count -1) to the top
// cells of `cell` in symbol order and returns how many there are.
static __device__ unsigned wave_fse_counts(uint8_t* cell, const int16_t* normSrc, unsigned maxSym, unsigned tl,
                                           FseBuildHdr& S, bool prefixAll)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    const unsigned T = 1u << tl;
    for (int s = lane; s < 256; s += 64) S.norm[s] = s <= (int)maxSym ? normSrc[s] : (int16_t)0;
    wave_sync();
    unsigned low = 0, carry = 0, carryAll = 0;
    for (int c = 0; c < 4; c++) {
        int s = c * 64 + lane;
        int v = S.norm[s];
        bool lp = v == -1;
        uint64_t m = __ballot(lp);
        if (lp) cell[T - 1 - (low + (unsigned)__popcll(m & lt))] = (uint8_t)s;
        low += (unsigned)__popcll(m);
        uint32_t cnt = v > 0 ? (uint32_t)v : 0u, cntAll = lp ? 1u : cnt;
        uint32_t x = cnt, y = cntAll;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t xs = __shfl_up(x, d, 64), ys = __shfl_up(y, d, 64);
            if (lane >= d) { x += xs; y += ys; }
        }
        S.cum[s] = (uint16_t)(carry + x - cnt);
        S.next[s] = (uint16_t)(prefixAll ? carryAll + y - cntAll : cntAll);
        carry += __shfl(x, 63, 64);
        carryAll += __shfl(y, 63, 64);
    }
    wave_sync();
    return low;
}

// Symbol spread with one wave: low-probability symbols take the top cells in
// symbol order (wave_fse_counts); the k-th occurrence of the other symbols (in
// symbol order) goes to the k-th position of the sequence j*step mod T that is
// not above the low-probability cells. Writes the symbol of every cell to
// `cell` and leaves S as wave_fse_counts does.
static __device__ void wave_fse_spread(uint8_t* cell, const int16_t* normSrc, unsigned maxSym, unsigned tl, FseBuildHdr& S,
                                       bool prefixAll)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    const unsigned T = 1u << tl, mask = T - 1, step = (T >> 1) + (T >> 3) + 3;
    const unsigned low = wave_fse_counts(cell, normSrc, maxSym, tl, S, prefixAll);
    const unsigned high = T - 1 - low;
    unsigned run = 0;
    for (unsigned j0 = 0; j0 < T; j0 += 64) {
        unsigned j = j0 + lane;
        unsigned pos = (j * step) & mask;
        bool v = j < T && pos <= high;
        uint64_t m = __ballot(v);
        if (v) {
            unsigned k = run + (unsigned)__popcll(m & lt);
            unsigned s = 0;
            for (unsigned half = 128; half; half >>= 1)
                if (S.cum[s + half] <= k) s += half;
            cell[pos] = (uint8_t)s;
        }
        run += (unsigned)__popcll(m);
    }
    wave_sync();
}

// The same spread without a search per cell: sk[k] (T bytes of scratch) gets
// the symbol of occurrence k, as the running maximum over markers sk[cum[s]] =
// s of the symbols with a positive count (cum increases with the symbol), then
// each cell reads its occurrence's symbol.
static __device__ void wave_fse_spread_sk(uint8_t* cell, uint8_t* sk, const int16_t* normSrc, unsigned maxSym,
                                          unsigned tl, FseBuildHdr& S)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    const unsigned T = 1u << tl, mask = T - 1, step = (T >> 1) + (T >> 3) + 3;
    const unsigned low = wave_fse_counts(cell, normSrc, maxSym, tl, S, false);
    uint32_t* skw = reinterpret_cast<uint32_t*>(sk);
    for (unsigned i = lane; i < T / 4; i += 64) skw[i] = 0;
    wave_sync();
    for (int c = 0; c < 4; c++) {
        const int s = c * 64 + lane;
        if (S.norm[s] > 0) sk[S.cum[s]] = (uint8_t)s;
    }
    wave_sync();
    // lane l takes dwords [D l, D l + D) (D = T / 256, at least 1)
    const unsigned D = T >= 256 ? T / 256 : 1u;
    const bool mine = D * 4 * (unsigned)lane < T;
    uint32_t w[8];
    uint32_t segMax = 0;
#pragma unroll
    for (unsigned i = 0; i < 8; i++) {
        w[i] = (i < D && mine) ? skw[D * lane + i] : 0u;
        const uint32_t a = w[i] & 255, b = (w[i] >> 8) & 255, c = (w[i] >> 16) & 255, d = w[i] >> 24;
        const uint32_t m0 = a > b ? a : b, m1 = c > d ? c : d;
        const uint32_t mm = m0 > m1 ? m0 : m1;
        segMax = segMax > mm ? segMax : mm;
    }
    // maximum over the lanes below (exclusive)
    uint32_t incl = segMax;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t x = __shfl_up(incl, d, 64);
        if (lane >= d) incl = incl > x ? incl : x;
    }
    uint32_t run = __shfl_up(incl, 1, 64);
    if (lane == 0) run = 0;
#pragma unroll
    for (unsigned i = 0; i < 8; i++) {
        uint32_t o = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint32_t x = (w[i] >> (8 * k)) & 255;
            run = run > x ? run : x;
            o |= run << (8 * k);
        }
        if (i < D && mine) skw[D * lane + i] = o;
    }
    wave_sync();
    const unsigned high = T - 1 - low;
    unsigned base = 0;
    for (unsigned j0 = 0; j0 < T; j0 += 64) {
        const unsigned j = j0 + lane;
        const unsigned pos = (j * step) & mask;
        const bool v = j < T && pos <= high;
        const uint64_t m = __ballot(v);
        if (v) cell[pos] = sk[base + (unsigned)__popcll(m & lt)];
        base += (unsigned)__popcll(m);
    }
    wave_sync();
}

// FSE_buildDTable (fse_decompress.c:73-176) with symbol identifiers: every
// cell's next state is its rank among the cells of its symbol (in cell order)
// plus the symbol's count; dt[u] = nbBits << 12 | 2 * newStateBase | id, with
// id in the low nbBits + 1 bits, which 2 * newStateBase leaves zero. A cell
// of nbBits = k has its next state in [2^(tl-k), 2^(tl-k+1)), so its symbol's
// count exceeds 2^(tl-k-1) and fewer than 2^(k+1) symbols have such cells: id
// is the symbol's rank among them (in symbol order) for k <= 6, whose symbol
// sys[2^(k+1) - 2 + id] holds (254 bytes), and the symbol itself for k >= 7.
// `cell` (T bytes) may be the upper half of a 2048-entry dt: a cell's symbol
// is read before the entries written over it. sk is T bytes of scratch
// (the lower half of such a dt).
static __device__ void wave_fse_dtable_ids(uint16_t* dt, uint8_t* sys, uint8_t* cell, uint8_t* sk, const int16_t* normSrc,
                                           unsigned maxSym, unsigned tl, FseBuildHdr& S)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    const unsigned T = 1u << tl;
    wave_fse_spread_sk(cell, sk, normSrc, maxSym, tl, S);
    // identifiers per symbol: S.cum[s] = id at the larger nbBits | id at the
    // smaller one << 8 (S.cum is free after the spread)
    uint32_t base[7];
#pragma unroll
    for (int k = 0; k < 7; k++) base[k] = 0;
    for (int c = 0; c < 4; c++) {
        const int s = c * 64 + lane;
        const int v = S.norm[s];
        const uint32_t n = v < 0 ? 1u : (uint32_t)v;
        const uint32_t kHi = n ? tl - (uint32_t)ozl_hb32(n | 1) : 99u;
        const uint32_t kLo = (n & (n - 1)) ? kHi - 1 : 99u;
        uint32_t ids = 0;
#pragma unroll
        for (uint32_t k = 0; k < 7; k++) {
            const bool mem = kHi == k || kLo == k;
            const uint64_t m = __ballot(mem);
            const uint32_t id = base[k] + (uint32_t)__popcll(m & lt);
            if (mem) {
                ids |= k == kHi ? id : id << 8;
                sys[((2u << k) - 2 + id) & 255] = (uint8_t)s;
            }
            base[k] += (uint32_t)__popcll(m);
        }
        S.cum[s] = (uint16_t)ids;
    }
    wave_sync();
    for (unsigned u0 = 0; u0 < T; u0 += 64) {
        unsigned u = u0 + lane;
        bool act = u < T;
        uint32_t s = act ? cell[u] : 0u;
        uint64_t m = match8(s, act);
        uint32_t nxt = act ? S.next[s] : 0u;
        uint32_t ids = act ? S.cum[s] : 0u;
        int v = act ? S.norm[s] : 0;
        if (act) {
            const uint32_t ns = nxt + (uint32_t)__popcll(m & lt);
            const uint32_t nb = tl - (uint32_t)ozl_hb32(ns);
            const uint32_t n = v < 0 ? 1u : (uint32_t)v;
            const uint32_t kHi = tl - (uint32_t)ozl_hb32(n);
            const uint32_t id = nb >= 7 ? s : nb == kHi ? (ids & 255) : (ids >> 8);
            dt[u] = (uint16_t)((nb << 12) | (((ns << nb) - T) << 1) | id);
        }
        wave_sync();
        if (act && (m & lt) == 0) S.next[s] = (uint16_t)(nxt + (uint32_t)__popcll(m));
        wave_sync();
    }
}

// FSE_buildCTable_wksp (fse_compress.c:69-215): the state table lists, per
// symbol in symbol order, T + cell for its cells in cell order; the symbol
// transforms are packed by ozl_tt_pack. st has T entries, tt maxSym+1.
static __device__ void wave_fse_ctable(uint16_t* st, uint32_t* tt, const int16_t* normSrc, unsigned maxSym, unsigned tl,
                                       FseBuild& S)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    const unsigned T = 1u << tl;
    wave_fse_spread(S.cell, normSrc, maxSym, tl, S, true);
    for (unsigned s = lane; s <= maxSym; s += 64) {
        int c = S.norm[s];
        int cumul = S.next[s];
        if (c == 0) tt[s] = ozl_tt_pack(tl + 1, T, -1);
        else if (c == -1 || c == 1) tt[s] = ozl_tt_pack(tl, T, cumul - 1);
        else {
            unsigned mb = tl - (unsigned)ozl_hb32((uint32_t)c - 1);
            tt[s] = ozl_tt_pack(mb, (unsigned)c << mb, cumul - c);
        }
    }
    wave_sync();
    for (unsigned u0 = 0; u0 < T; u0 += 64) {
        unsigned u = u0 + lane;
        bool act = u < T;
        uint32_t s = act ? S.cell[u] : 0u;
        uint64_t m = match8(s, act);
        uint32_t base = act ? S.next[s] : 0u;
        if (act) st[base + (uint32_t)__popcll(m & lt)] = (uint16_t)(T + u);
        wave_sync();
        if (act && (m & lt) == 0) S.next[s] = (uint16_t)(base + (uint32_t)__popcll(m));
        wave_sync();
    }
}

// ---------------------------------------------------------------- Huffman lookup table
struct HufBuild {
    uint8_t len[256];
    uint8_t list[256];    // symbols by code length, symbol order within a length
    uint16_t start[16];   // first code value of each length
    uint16_t off[16];     // first list index of each length
    uint16_t cnt[16];
};

// Canonical codes of ozl_huf_from_weights (HUF_readDTableX1, huf_decompress.c:
// 342-430): lengths tl..1 take consecutive value ranges from 0 upward, symbol
// order within a length. Entry e of the lookup table (the next tl stream bits)
// holds symbol | length << 8 of the code it starts with.
static __device__ void wave_huf_lut(uint16_t* lut, const uint8_t* lenG, unsigned tl, HufBuild& S)
{
    const int lane = threadIdx.x & 63;
    const uint64_t lt = lanes_below();
    uint32_t run[13];
#pragma unroll
    for (int r = 0; r < 13; r++) run[r] = 0;
    uint8_t myLen[4];
    uint32_t myRank[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int s = c * 64 + lane;
        uint32_t l = lenG[s];
        if (l > tl) l = 0;
        myLen[c] = (uint8_t)l;
        S.len[s] = (uint8_t)l;
        uint32_t rk = 0;
#pragma unroll
        for (int r = 1; r <= 12; r++) {
            uint64_t m = __ballot(l == (uint32_t)r);
            if (l == (uint32_t)r) rk = run[r] + (uint32_t)__popcll(m & lt);
            run[r] += (uint32_t)__popcll(m);
        }
        myRank[c] = rk;
    }
    if (lane == 0) {
        uint32_t mn = 0, o = 0;
        for (int r = (int)tl; r > 0; r--) {
            S.start[r] = (uint16_t)mn;
            mn = (mn + run[r]) >> 1;
        }
        for (int r = 1; r <= 12; r++) {
            S.off[r] = (uint16_t)o;
            S.cnt[r] = (uint16_t)run[r];
            o += run[r];
        }
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < 4; c++)
        if (myLen[c]) S.list[S.off[myLen[c]] + myRank[c]] = (uint8_t)(c * 64 + lane);
    wave_sync();
    const unsigned T = 1u << tl;
    for (unsigned e = lane; e < T; e += 64) {
        uint32_t ent = 0;
        for (unsigned r = 1; r <= tl; r++) {
            uint32_t b0 = (uint32_t)S.start[r] << (tl - r), b1 = ((uint32_t)S.start[r] + S.cnt[r]) << (tl - r);
            if (e >= b0 && e < b1) ent = S.list[S.off[r] + ((e - b0) >> (tl - r))] | (r << 8);
        }
        lut[e] = (uint16_t)ent;
    }
    wave_sync();
}

}  // namespace ozl_wave
