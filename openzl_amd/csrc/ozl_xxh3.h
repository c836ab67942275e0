// ozl_xxh3.h -- XXH3-64 (seed 0, default secret) for host and device.
//
// The frame format stores XXH3-64 checksums (src/openzl/compress/cctx.c:
// 1537-1580, encode_frameheader.c:624-630). Written from the published XXH3
// algorithm; the 192-byte default secret is the specification's table.
//   ozl_xxh3_seq   one lane, any length (headers, small streams)
//   ozl_xxh3_wave  one 64-lane wavefront: each 1 KiB block is one coalesced
//                  wave load; its 16 stripe sums are reduced across lanes and
//                  only the per-block scramble chain is sequential.
#pragma once
#include "ozl_core.h"
#if defined(__HIPCC__)
#include "ozl_gmem.h"
#endif

#define OZL_XXH3_SECRET_BYTES \
    0xb8, 0xfe, 0x6c, 0x39, 0x23, 0xa4, 0x4b, 0xbe, 0x7c, 0x01, 0x81, 0x2c, 0xf7, 0x21, 0xad, 0x1c, \
    0xde, 0xd4, 0x6d, 0xe9, 0x83, 0x90, 0x97, 0xdb, 0x72, 0x40, 0xa4, 0xa4, 0xb7, 0xb3, 0x67, 0x1f, \
    0xcb, 0x79, 0xe6, 0x4e, 0xcc, 0xc0, 0xe5, 0x78, 0x82, 0x5a, 0xd0, 0x7d, 0xcc, 0xff, 0x72, 0x21, \
    0xb8, 0x08, 0x46, 0x74, 0xf7, 0x43, 0x24, 0x8e, 0xe0, 0x35, 0x90, 0xe6, 0x81, 0x3a, 0x26, 0x4c, \
    0x3c, 0x28, 0x52, 0xbb, 0x91, 0xc3, 0x00, 0xcb, 0x88, 0xd0, 0x65, 0x8b, 0x1b, 0x53, 0x2e, 0xa3, \
    0x71, 0x64, 0x48, 0x97, 0xa2, 0x0d, 0xf9, 0x4e, 0x38, 0x19, 0xef, 0x46, 0xa9, 0xde, 0xac, 0xd8, \
    0xa8, 0xfa, 0x76, 0x3f, 0xe3, 0x9c, 0x34, 0x3f, 0xf9, 0xdc, 0xbb, 0xc7, 0xc7, 0x0b, 0x4f, 0x1d, \
    0x8a, 0x51, 0xe0, 0x4b, 0xcd, 0xb4, 0x59, 0x31, 0xc8, 0x9f, 0x7e, 0xc9, 0xd9, 0x78, 0x73, 0x64, \
    0xea, 0xc5, 0xac, 0x83, 0x34, 0xd3, 0xeb, 0xc3, 0xc5, 0x81, 0xa0, 0xff, 0xfa, 0x13, 0x63, 0xeb, \
    0x17, 0x0d, 0xdd, 0x51, 0xb7, 0xf0, 0xda, 0x49, 0xd3, 0x16, 0x55, 0x26, 0x29, 0xd4, 0x68, 0x9e, \
    0x2b, 0x16, 0xbe, 0x58, 0x7d, 0x47, 0xa1, 0xfc, 0x8f, 0xf8, 0xb8, 0xd1, 0x7a, 0xd0, 0x31, 0xce, \
    0x45, 0xcb, 0x3a, 0x8f, 0x95, 0x16, 0x04, 0x28, 0xaf, 0xd7, 0xfb, 0xca, 0xbb, 0x4b, 0x40, 0x7e,
static const uint8_t ozl_xxh3_secret_host[192] = { OZL_XXH3_SECRET_BYTES };
#if defined(__HIPCC__)
static __device__ __constant__ uint8_t ozl_xxh3_secret_dev[192] = { OZL_XXH3_SECRET_BYTES };
#endif
#define ozl_xxh3_secret OZL_TAB(ozl_xxh3_secret)

#define OZL_P32_1 0x9E3779B1u
#define OZL_P32_2 0x85EBCA77u
#define OZL_P32_3 0xC2B2AE3Du
#define OZL_P64_1 0x9E3779B185EBCA87ull
#define OZL_P64_2 0xC2B2AE3D27D4EB4Full
#define OZL_P64_3 0x165667B19E3779F9ull
#define OZL_P64_4 0x85EBCA77C2B2AE63ull
#define OZL_P64_5 0x27D4EB2F165667C5ull
#define OZL_PMX1 0x165667919E3779F9ull
#define OZL_PMX2 0x9FB21C651E98DF25ull

// byte-wise little-endian reads: inputs may sit at any byte offset
OZL_HD uint64_t ozl_le64(const uint8_t* p)
{
    uint64_t v = 0;
    for (int k = 0; k < 8; k++) v |= (uint64_t)p[k] << (8 * k);
    return v;
}
OZL_HD uint32_t ozl_le32(const uint8_t* p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
OZL_HD uint64_t ozl_sec64(int off) { return ozl_le64(ozl_xxh3_secret + off); }
OZL_HD uint64_t ozl_rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
OZL_HD uint64_t ozl_fold128(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t lo = a * b, hi = __umul64hi(a, b);
#else
    unsigned __int128 m = (unsigned __int128)a * b;
    uint64_t lo = (uint64_t)m, hi = (uint64_t)(m >> 64);
#endif
    return lo ^ hi;
}
OZL_HD uint64_t ozl_avalanche64(uint64_t h)
{
    h ^= h >> 33; h *= OZL_P64_2; h ^= h >> 29; h *= OZL_P64_3; h ^= h >> 32; return h;
}
OZL_HD uint64_t ozl_avalanche3(uint64_t h) { h ^= h >> 37; h *= OZL_PMX1; h ^= h >> 32; return h; }
OZL_HD uint64_t ozl_mix16(const uint8_t* in, int so)
{
    return ozl_fold128(ozl_le64(in) ^ ozl_sec64(so), ozl_le64(in + 8) ^ ozl_sec64(so + 8));
}
OZL_HD uint64_t ozl_xxh3_short(const uint8_t* in, size_t len)
{
    if (len == 0) return ozl_avalanche64(ozl_sec64(56) ^ ozl_sec64(64));
    if (len <= 3) {
        uint32_t c1 = in[0], c2 = in[len >> 1], c3 = in[len - 1];
        uint32_t comb = (c1 << 16) | (c2 << 24) | c3 | ((uint32_t)len << 8);
        uint64_t flip = (uint64_t)(ozl_le32(ozl_xxh3_secret) ^ ozl_le32(ozl_xxh3_secret + 4));
        return ozl_avalanche64((uint64_t)comb ^ flip);
    }
    if (len <= 8) {
        uint32_t i1 = ozl_le32(in), i2 = ozl_le32(in + len - 4);
        uint64_t flip = ozl_sec64(8) ^ ozl_sec64(16);
        uint64_t h = (i2 + ((uint64_t)i1 << 32)) ^ flip;
        h ^= ozl_rotl64(h, 49) ^ ozl_rotl64(h, 24);
        h *= OZL_PMX2;
        h ^= (h >> 35) + len;
        h *= OZL_PMX2;
        h ^= h >> 28;
        return h;
    }
    if (len <= 16) {
        uint64_t lo = ozl_le64(in) ^ (ozl_sec64(24) ^ ozl_sec64(32));
        uint64_t hi = ozl_le64(in + len - 8) ^ (ozl_sec64(40) ^ ozl_sec64(48));
        uint64_t acc = len + __builtin_bswap64(lo) + hi + ozl_fold128(lo, hi);
        return ozl_avalanche3(acc);
    }
    if (len <= 128) {
        uint64_t acc = len * OZL_P64_1;
        if (len > 32) {
            if (len > 64) {
                if (len > 96) { acc += ozl_mix16(in + 48, 96); acc += ozl_mix16(in + len - 64, 112); }
                acc += ozl_mix16(in + 32, 64); acc += ozl_mix16(in + len - 48, 80);
            }
            acc += ozl_mix16(in + 16, 32); acc += ozl_mix16(in + len - 32, 48);
        }
        acc += ozl_mix16(in, 0); acc += ozl_mix16(in + len - 16, 16);
        return ozl_avalanche3(acc);
    }
    // 129..240
    uint64_t acc = len * OZL_P64_1;
    size_t rounds = len / 16;
    for (size_t i = 0; i < 8; i++) acc += ozl_mix16(in + 16 * i, 16 * (int)i);
    acc = ozl_avalanche3(acc);
    for (size_t i = 8; i < rounds; i++) acc += ozl_mix16(in + 16 * i, 16 * (int)(i - 8) + 3);
    acc += ozl_mix16(in + len - 16, 136 - 17);
    return ozl_avalanche3(acc);
}

// contribution of one 64-byte stripe to accumulator lane i
OZL_HD uint64_t ozl_stripe_term(const uint8_t* stripe, int secOff, int i)
{
    uint64_t dOther = ozl_le64(stripe + 8 * (i ^ 1));
    uint64_t d = ozl_le64(stripe + 8 * i);
    uint64_t k = d ^ ozl_sec64(secOff + 8 * i);
    return dOther + (uint64_t)(uint32_t)k * (k >> 32);
}
OZL_HD uint64_t ozl_scramble(uint64_t a, int i)
{
    a ^= a >> 47;
    a ^= ozl_sec64(192 - 64 + 8 * i);
    a *= OZL_P32_1;
    return a;
}
OZL_HD uint64_t ozl_xxh3_merge(const uint64_t* acc, size_t len)
{
    uint64_t res = len * OZL_P64_1;
    for (int i = 0; i < 4; i++)
        res += ozl_fold128(acc[2 * i] ^ ozl_sec64(11 + 16 * i), acc[2 * i + 1] ^ ozl_sec64(11 + 16 * i + 8));
    return ozl_avalanche3(res);
}

OZL_HD uint64_t ozl_xxh3_seq(const uint8_t* in, size_t len)
{
    if (len <= 240) return ozl_xxh3_short(in, len);
    uint64_t acc[8] = { OZL_P32_3, OZL_P64_1, OZL_P64_2, OZL_P64_3, OZL_P64_4, OZL_P32_2, OZL_P64_5, OZL_P32_1 };
    size_t nbBlocks = (len - 1) / 1024;
    for (size_t b = 0; b < nbBlocks; b++) {
        for (int s = 0; s < 16; s++)
            for (int i = 0; i < 8; i++) acc[i] += ozl_stripe_term(in + b * 1024 + s * 64, s * 8, i);
        for (int i = 0; i < 8; i++) acc[i] = ozl_scramble(acc[i], i);
    }
    size_t nbStripes = ((len - 1) - 1024 * nbBlocks) / 64;
    for (size_t s = 0; s < nbStripes; s++)
        for (int i = 0; i < 8; i++) acc[i] += ozl_stripe_term(in + nbBlocks * 1024 + s * 64, (int)s * 8, i);
    for (int i = 0; i < 8; i++) acc[i] += ozl_stripe_term(in + len - 64, 192 - 64 - 7, i);
    return ozl_xxh3_merge(acc, len);
}

#if defined(__HIPCC__)
// One wavefront hashes `len` bytes at `in`; every lane returns the hash.
// Each 1 KiB block is one coalesced wave load: lane l takes bytes
// [16l, 16l+16), i.e. words 2j, 2j+1 (j = l mod 4) of stripe l/4, and forms
// their terms for accumulators 2j and 2j+1; the 16 stripes are summed across
// the lanes with equal j (DPP row rotations, then two xor shuffles), and
// every lane keeps the accumulator pair of its j through the per-block
// scramble chain. The tail stripes run on lanes 0..7 as in ozl_xxh3_seq.
// `lds` and `chainCap` are unused (kept for the callers' signature).
template <int CTRL>
__device__ inline uint64_t ozl_dpp64(uint64_t v)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, CTRL, 0xF, 0xF, false);
    uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((uint64_t)hi << 32) | lo;
}
// Lane constants of the block step: lane l holds words 2j, 2j+1 (j = l mod 4)
// of stripe l/4 and the accumulator pair 2j, 2j+1.
struct ozl_xxh3_lane {
    uint64_t k0, k1;  // stripe secrets of the lane's two words
    uint64_t c0, c1;  // scramble secrets of accumulators 2j, 2j+1
    int j;
};
__device__ inline ozl_xxh3_lane ozl_xxh3_lane_init()
{
    const int lane = threadIdx.x & 63;
    ozl_xxh3_lane L;
    L.j = lane & 3;
    const int st = lane >> 2, i0 = 2 * L.j;
    L.k0 = ozl_sec64(8 * st + 8 * i0);
    L.k1 = ozl_sec64(8 * st + 8 * i0 + 8);
    L.c0 = ozl_sec64(192 - 64 + 8 * i0);
    L.c1 = ozl_sec64(192 - 64 + 8 * i0 + 8);
    return L;
}
// this lane's 16 bytes of the 1 KiB block at `blk`, as two words. mode 1:
// the block is 16-byte aligned (one 16-byte load); mode 2: unaligned, with at
// least 16 readable bytes after the block (two aligned 16-byte loads, the
// second one the next lane's first, and a funnel shift: whole cache lines per
// instruction); mode 0: four dword loads.
template <int Q>
__device__ inline void ozl_xxh3_pick(const uint32_t* d, uint32_t bs, uint32_t& d0, uint32_t& d1, uint32_t& d2,
                                     uint32_t& d3)
{
    d0 = __builtin_amdgcn_alignbit(d[Q + 1], d[Q], bs);
    d1 = __builtin_amdgcn_alignbit(d[Q + 2], d[Q + 1], bs);
    d2 = __builtin_amdgcn_alignbit(d[Q + 3], d[Q + 2], bs);
    d3 = __builtin_amdgcn_alignbit(d[Q + 4], d[Q + 3], bs);
}
__device__ inline void ozl_xxh3_block_load(const uint8_t* blk, int mode, uint64_t& v0, uint64_t& v1)
{
    const uint8_t* p = blk + 16 * (threadIdx.x & 63);
    uint32_t d0, d1, d2, d3;
    if (mode == 1) {
        const ozl_u32x4 q = ozl_gld16(p);
        d0 = q.x; d1 = q.y; d2 = q.z; d3 = q.w;
    } else if (mode == 2) {
        const uintptr_t a = (uintptr_t)p;
        const uint8_t* base = reinterpret_cast<const uint8_t*>(a & ~(uintptr_t)15);
        const uint32_t sh = (uint32_t)(a & 15), bs = 8 * (sh & 3);
        const ozl_u32x4 lo = ozl_gld16(base), hi = ozl_gld16(base + 16);
        const uint32_t d[8] = { lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w };
        switch (sh >> 2) {
        case 0: ozl_xxh3_pick<0>(d, bs, d0, d1, d2, d3); break;
        case 1: ozl_xxh3_pick<1>(d, bs, d0, d1, d2, d3); break;
        case 2: ozl_xxh3_pick<2>(d, bs, d0, d1, d2, d3); break;
        default: ozl_xxh3_pick<3>(d, bs, d0, d1, d2, d3); break;
        }
    } else {
        d0 = ozl_gld4(p); d1 = ozl_gld4(p + 4); d2 = ozl_gld4(p + 8); d3 = ozl_gld4(p + 12);
    }
    v0 = d0 | ((uint64_t)d1 << 32);
    v1 = d2 | ((uint64_t)d3 << 32);
}
// load mode of block b of a region of len bytes at `in` (see ozl_xxh3_block_load)
__device__ inline int ozl_xxh3_mode(const uint8_t* in, size_t len, size_t b)
{
    if (((uintptr_t)in & 15) == 0) return 1;
    return (b + 1) * 1024 + 16 <= len ? 2 : 0;
}
// v + v of lane l ^ 16, and of lane l ^ 32: v_permlane16_swap /
// v_permlane32_swap with both operands v hand back the two rows (halves) in
// swapped order, so their sum is the pair's sum in every lane (no LDS)
__device__ inline uint32_t ozl_add_xor16(uint32_t v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return r[0] + r[1];
}
__device__ inline uint64_t ozl_add_xor16_64(uint64_t v)
{
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    return (((uint64_t)hi[0] << 32) | lo[0]) + (((uint64_t)hi[1] << 32) | lo[1]);
}
__device__ inline uint64_t ozl_add_xor32_64(uint64_t v)
{
    const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)v, (uint32_t)v, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(v >> 32), (uint32_t)(v >> 32), false, false);
    return (((uint64_t)hi[0] << 32) | lo[0]) + (((uint64_t)hi[1] << 32) | lo[1]);
}
// sums over the 16 stripes of a 1 KiB block for accumulators 2j and 2j+1,
// in every lane of the pair's group, from the words ozl_xxh3_block_load read
__device__ inline void ozl_xxh3_sums_of(uint64_t v0, uint64_t v1, const ozl_xxh3_lane& L, uint64_t& sa, uint64_t& sb)
{
    const uint64_t x0 = v0 ^ L.k0, x1 = v1 ^ L.k1;
    sa = v1 + (uint64_t)(uint32_t)x0 * (x0 >> 32);
    sb = v0 + (uint64_t)(uint32_t)x1 * (x1 >> 32);
    sa += ozl_dpp64<0x124>(sa);  // row_ror:4
    sb += ozl_dpp64<0x124>(sb);
    sa += ozl_dpp64<0x128>(sa);  // row_ror:8
    sb += ozl_dpp64<0x128>(sb);
    sa = ozl_add_xor16_64(sa);
    sb = ozl_add_xor16_64(sb);
    sa = ozl_add_xor32_64(sa);
    sb = ozl_add_xor32_64(sb);
}
__device__ inline void ozl_xxh3_block_sums(const uint8_t* blk, int mode, const ozl_xxh3_lane& L, uint64_t& sa,
                                           uint64_t& sb)
{
    uint64_t v0, v1;
    ozl_xxh3_block_load(blk, mode, v0, v1);
    ozl_xxh3_sums_of(v0, v1, L, sa, sb);
}
__device__ inline uint64_t ozl_xxh3_scr(uint64_t a, uint64_t c)
{
    a ^= a >> 47;
    a ^= c;
    return a * OZL_P32_1;
}
// initial accumulator i
__device__ inline uint64_t ozl_xxh3_init_acc(int i)
{
    return i == 0 ? OZL_P32_3 : i == 1 ? OZL_P64_1 : i == 2 ? OZL_P64_2 : i == 3 ? OZL_P64_3
         : i == 4 ? OZL_P64_4 : i == 5 ? OZL_P32_2 : i == 6 ? OZL_P64_5 : OZL_P32_1;
}
// lanes 0..7 hold accumulator `lane` after the full blocks; adds the tail
// stripes and the last stripe and returns the hash in every lane. Every lane
// of the wave runs it: lane l reads the 16-byte piece l of the tail (stripe
// l / 4, words 2 (l % 4) and 2 (l % 4) + 1) and lanes 0..3 the pieces of the
// last stripe, with dword loads (any alignment) issued together; the terms
// are summed across lanes like a block's.
__device__ inline uint64_t ozl_xxh3_finish(const uint8_t* in, size_t len, size_t nbBlocks, uint64_t acc)
{
    const int lane = threadIdx.x & 63;
    const uint32_t st = (uint32_t)lane >> 2, j = (uint32_t)lane & 3;
    const uint32_t nbStripes = (uint32_t)(((len - 1) - 1024 * nbBlocks) / 64);   // < 16
    uint32_t d[4] = { 0, 0, 0, 0 }, e[4] = { 0, 0, 0, 0 };
    const uint8_t* p = in + nbBlocks * 1024 + 16 * (size_t)lane;
    const uint8_t* q = in + len - 64 + 16 * (size_t)j;
    if (st < nbStripes)
        for (int k = 0; k < 4; k++) d[k] = ozl_gld4(p + 4 * k);
    if (lane < 4)
        for (int k = 0; k < 4; k++) e[k] = ozl_gld4(q + 4 * k);
    uint64_t ta = 0, tb = 0;   // terms of accumulators 2j and 2j + 1
    if (st < nbStripes) {
        const uint64_t w0 = d[0] | ((uint64_t)d[1] << 32), w1 = d[2] | ((uint64_t)d[3] << 32);
        const uint64_t x0 = w0 ^ ozl_sec64(8 * (int)st + 16 * (int)j), x1 = w1 ^ ozl_sec64(8 * (int)st + 16 * (int)j + 8);
        ta = w1 + (uint64_t)(uint32_t)x0 * (x0 >> 32);
        tb = w0 + (uint64_t)(uint32_t)x1 * (x1 >> 32);
    }
    if (lane < 4) {
        const uint64_t w0 = e[0] | ((uint64_t)e[1] << 32), w1 = e[2] | ((uint64_t)e[3] << 32);
        const uint64_t x0 = w0 ^ ozl_sec64(192 - 64 - 7 + 16 * (int)j), x1 = w1 ^ ozl_sec64(192 - 64 - 7 + 16 * (int)j + 8);
        ta += w1 + (uint64_t)(uint32_t)x0 * (x0 >> 32);
        tb += w0 + (uint64_t)(uint32_t)x1 * (x1 >> 32);
    }
    // sum over the lanes of the same j (rows of 16, then the four rows)
    ta += ozl_dpp64<0x124>(ta);  // row_ror:4
    tb += ozl_dpp64<0x124>(tb);
    ta += ozl_dpp64<0x128>(ta);  // row_ror:8
    tb += ozl_dpp64<0x128>(tb);
    ta = ozl_add_xor32_64(ozl_add_xor16_64(ta));
    tb = ozl_add_xor32_64(ozl_add_xor16_64(tb));
    uint64_t a[8];
    for (int i = 0; i < 8; i++) a[i] = __shfl(acc, i, 64) + __shfl((i & 1) ? tb : ta, i >> 1, 64);
    return ozl_xxh3_merge(a, len);
}

// XXH3 of a buffer by one wave (many waves per CU keep the loads in flight)
__device__ inline uint64_t ozl_xxh3_wave(const uint8_t* in, size_t len, uint64_t* lds, int chainCap)
{
    (void)lds;
    (void)chainCap;
    const int lane = threadIdx.x & 63;
    if (len <= 240) return ozl_xxh3_short(in, len);
    const ozl_xxh3_lane L = ozl_xxh3_lane_init();
    uint64_t A = ozl_xxh3_init_acc(2 * L.j), B = ozl_xxh3_init_acc(2 * L.j + 1);
    const size_t nbBlocks = (len - 1) / 1024;
    for (size_t b = 0; b < nbBlocks; b++) {
        uint64_t sa, sb;
        ozl_xxh3_block_sums(in + b * 1024, ozl_xxh3_mode(in, len, b), L, sa, sb);
        A = ozl_xxh3_scr(A + sa, L.c0);
        B = ozl_xxh3_scr(B + sb, L.c1);
    }
    // accumulator i to lane i, then the tail stripes and the last stripe
    uint64_t acc = 0;
    for (int q = 0; q < 4; q++) {
        const uint64_t a = __shfl(A, q, 64), bq = __shfl(B, q, 64);
        if (lane == 2 * q) acc = a;
        if (lane == 2 * q + 1) acc = bq;
    }
    return ozl_xxh3_finish(in, len, nbBlocks, acc);
}
#endif
