// 2-byte inputs of ZL_GRAPH_HUFFMAN on gfx950
// (src/openzl/codecs/entropy/encode_entropy_binding.c:612-682): 16-bit
// histograms and their statistics, tokenize of small alphabets, the 16-bit
// Huffman coder of huffman_struct_v2 (encode_huffman_kernel.c,
// decode_huffman_kernel.c), and the decoders of those nodes.
//
// Work per call of the graph engine (one launch per kernel for all chunks):
//   k_e16_hist      16 elements per thread, atomics into the chunk's 65,536
//                   counters in global memory (L2)
//   k_e16_stats     one workgroup per chunk: largest value, distinct values,
//                   entropy terms (exact int64 sum)
//   k_e16_tokenize  one workgroup per chunk: distinct values (< 256) in LDS,
//                   first occurrences by atomicMin, ranks, indices by a binary
//                   search over the sorted values
//   k_e16_huf_build one workgroup per chunk: nonzero counts compacted in
//                   symbol order, bitonic sort by (count desc, symbol), the
//                   reference's two-queue tree and height limit on one lane,
//                   canonical codes per length in symbol order (16-bit
//                   counters, as the reference's), weights
//   k_e16_huf_encode one wave per stream: reverse scan of code lengths over
//                   the lanes' segments, codes packed in registers, shared
//                   boundary words OR-ed
//   k_e16_huf_decode one workgroup per chunk: weights checked, symbols sorted
//                   by weight, one lane per stream reads its bits from the end
//                   mark down through the canonical ranges
// Node arrays of up to 2,048 leaves live in LDS; larger alphabets use the
// chunk's workspace in global memory with the same code.
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_gmem.h"
#include "ozl_profile.h"
#include "ozl_entropy16_api.h"

namespace {

constexpr int kT = 256;
constexpr uint32_t kSyms = 65536;
constexpr int kHistPerThread = 16;
constexpr uint32_t kLdsLeaves = 2048;
constexpr int kMaxBits = 20;   // ZS_kLargeHuffmanMaxTableLog / MaxNbBits

// Exclusive scan of one value per thread of the block; *total gets the sum.
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total)
{
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kT; d <<= 1) {
        const uint32_t x = t >= d ? sh[t - d] : 0;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    const uint32_t incl = sh[t];
    *total = sh[kT - 1];
    __syncthreads();
    return incl - v;
}

// ---------------------------------------------------------------- statistics
__global__ void __launch_bounds__(kT) k_e16_hist(const ozl_c16* __restrict__ chunks, uint32_t* __restrict__ hist)
{
    const ozl_c16 c = chunks[blockIdx.y];
    const uint32_t a = (blockIdx.x * kT + threadIdx.x) * kHistPerThread;
    if (a >= c.n) return;
    const uint32_t b = a + kHistPerThread < c.n ? a + kHistPerThread : c.n;
    uint32_t* h = hist + (size_t)blockIdx.y * kSyms;
    const uint8_t* s = reinterpret_cast<const uint8_t*>((uintptr_t)c.src);
    for (uint32_t i = a; i < b; i++) atomicAdd(&h[ozl_gld2(s + 2 * (size_t)i)], 1u);
}

__global__ void __launch_bounds__(kT) k_e16_stats(const ozl_c16* __restrict__ chunks, const uint32_t* __restrict__ hist,
                                                  ozl_s16* __restrict__ stats)
{
    __shared__ int64_t sT[kT];
    __shared__ uint32_t sM[kT], sC[kT];
    const ozl_c16 c = chunks[blockIdx.x];
    const uint32_t* h = hist + (size_t)blockIdx.x * kSyms;
    const int t = threadIdx.x;
    const int64_t norm = c.n ? ((int64_t)1 << 62) / (int64_t)c.n : 0;
    int64_t ts = 0;
    uint32_t mx = 0, cd = 0;
    for (uint32_t s = (uint32_t)t * (kSyms / kT); s < (uint32_t)(t + 1) * (kSyms / kT); s++) {
        const uint32_t x = h[s];
        if (x) {
            ts += ozl_entropy_term(x, norm);
            mx = s;
            cd++;
        }
    }
    sT[t] = ts;
    sM[t] = mx;
    sC[t] = cd;
    __syncthreads();
    for (int d = kT / 2; d > 0; d >>= 1) {
        if (t < d) {
            sT[t] += sT[t + d];
            sC[t] += sC[t + d];
            sM[t] = sM[t + d] > sM[t] ? sM[t + d] : sM[t];
        }
        __syncthreads();
    }
    if (t == 0) {
        ozl_s16 r;
        r.maxSym = sM[0];
        r.card = sC[0];
        r.maxCount = h[sM[0]];
        r.pad = 0;
        r.termSum = sT[0];
        stats[blockIdx.x] = r;
    }
}

// ---------------------------------------------------------------- tokenize
// encode_tokenize_binding.c:124-216 for a struct input of < 256 distinct
// values: alphabet in order of first occurrence, 1-byte indices.
__device__ inline uint32_t slot_of(const uint16_t* syms, uint32_t card, uint32_t v)
{
    uint32_t lo = 0, hi = card;   // syms ascending; v is present
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (syms[m] <= v) lo = m;
        else hi = m;
    }
    return lo;
}

__global__ void __launch_bounds__(kT) k_e16_tokenize(const ozl_tok16* __restrict__ jobs)
{
    __shared__ uint16_t syms[256];
    __shared__ uint32_t first[256];
    __shared__ uint8_t rank[256];
    __shared__ uint32_t sh[kT];
    const ozl_tok16 j = jobs[blockIdx.x];
    const uint32_t* h = reinterpret_cast<const uint32_t*>((uintptr_t)j.hist);
    const uint8_t* src = reinterpret_cast<const uint8_t*>((uintptr_t)j.src);
    const int t = threadIdx.x;
    // distinct values in ascending order
    const uint32_t a = (uint32_t)t * (kSyms / kT), b = a + kSyms / kT;
    uint32_t cnt = 0;
    for (uint32_t s = a; s < b; s++) cnt += h[s] != 0;
    uint32_t total;
    uint32_t pos = block_excl_scan(cnt, sh, &total);
    for (uint32_t s = a; s < b && pos < 256; s++)
        if (h[s]) syms[pos++] = (uint16_t)s;
    first[t] = 0xFFFFFFFFu;
    __syncthreads();
    const uint32_t card = j.card;
    for (uint32_t i = t; i < j.n; i += kT) atomicMin(&first[slot_of(syms, card, ozl_gld2(src + 2 * (size_t)i))], i);
    __syncthreads();
    if ((uint32_t)t < card) {
        uint32_t r = 0;
        for (uint32_t u = 0; u < card; u++) r += first[u] < first[t];
        rank[t] = (uint8_t)r;
        ozl_gst2(reinterpret_cast<uint8_t*>((uintptr_t)j.alpha) + 2 * r, syms[t]);
    }
    __syncthreads();
    uint8_t* idx = reinterpret_cast<uint8_t*>((uintptr_t)j.idx);
    for (uint32_t i = t; i < j.n; i += kT) ozl_gst1(idx + i, rank[slot_of(syms, card, ozl_gld2(src + 2 * (size_t)i))]);
}

// ---------------------------------------------------------------- 16-bit Huffman table
// Bitonic sort of P (a power of two) 64-bit keys, ascending, by the block.
__device__ void bitonic_sort(uint64_t* k, uint32_t P)
{
    for (uint32_t size = 2; size <= P; size <<= 1)
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            for (uint32_t i = threadIdx.x; i < P / 2; i += kT) {
                const uint32_t lo = 2 * i - (i & (stride - 1));
                const uint32_t hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t x = k[lo], y = k[hi];
                if ((x > y) == up) {
                    k[lo] = y;
                    k[hi] = x;
                }
            }
            __syncthreads();
        }
}

// encode_huffman_kernel.c:36-140 ZS_largeHuffmanSetMaxHeight over the leaves
// 0..last (decreasing counts), one lane. Returns the height, or 99 where the
// reference would fail an assertion (no leaf to deepen).
__device__ int h16_set_max_height(uint8_t* nb, const uint32_t* cnt, int last, int maxNbBits)
{
    const int largest = nb[last];
    if (largest <= maxNbBits) return largest;
    int cost = 0;
    const int base = 1 << (largest - maxNbBits);
    int n = last;
    for (; nb[n] > maxNbBits; n--) {
        cost += base - (1 << (largest - nb[n]));
        nb[n] = (uint8_t)maxNbBits;
    }
    while (n >= 0 && nb[n] == maxNbBits) n--;
    cost >>= largest - maxNbBits;
    const uint32_t none = 0xF0F0F0F0u;
    uint32_t rl[22];
    for (int r = 0; r < 22; r++) rl[r] = none;
    {
        int cur = maxNbBits;
        for (int p = n; p >= 0; p--) {
            if (nb[p] >= cur) continue;
            cur = nb[p];
            rl[maxNbBits - cur] = (uint32_t)p;
        }
    }
    while (cost > 0) {
        int k = ozl_hb32((uint32_t)cost) + 1;
        while (k > 1 && rl[k] == none) k--;
        for (; k > 1; k--) {
            const uint32_t hi = rl[k], lo = rl[k - 1];
            if (lo == none) break;
            if (cnt[hi] <= 2 * cnt[lo]) break;
        }
        while (k <= kMaxBits && rl[k] == none) k++;
        if (k > kMaxBits) return 99;
        cost -= 1 << (k - 1);
        if (rl[k - 1] == none) rl[k - 1] = rl[k];
        nb[rl[k]]++;
        if (rl[k] == 0) {
            rl[k] = none;
        } else {
            rl[k]--;
            if (nb[rl[k]] != maxNbBits - k) rl[k] = none;
        }
    }
    while (cost < 0) {
        if (rl[1] == none) {
            while (n >= 0 && nb[n] == maxNbBits) n--;
            if (n + 1 > last) return 99;
            nb[n + 1]--;
            rl[1] = (uint32_t)(n + 1);
            cost++;
            continue;
        }
        if ((int)rl[1] + 1 > last) return 99;
        nb[rl[1] + 1]--;
        rl[1]++;
        cost++;
    }
    return maxNbBits;
}

struct H16Lds {
    uint32_t cnt[2 * kLdsLeaves + 1];    // node counts, [0] is the sentinel of index -1
    union {
        uint64_t keys[kLdsLeaves];       // sort keys, then
        uint32_t par[2 * kLdsLeaves];    // node parents
    };
    uint16_t sym[kLdsLeaves];
    uint8_t nb[2 * kLdsLeaves];
    uint32_t perRank[kMaxBits + 2];
    uint16_t start[kMaxBits + 2];
    uint32_t sh[kT];
    int tl;
};

__host__ __device__ inline uint32_t pow2_at_least(uint32_t k)
{
    uint32_t p = 1;
    while (p < k) p <<= 1;
    return p;
}

// Workspace of one job: the table u32[S] (code | length << 16), then, above
// the LDS limit, keys u64[P] | counts u32[2k+1] | parents u32[2k] |
// symbols u16[k] | lengths u8[2k].
__host__ __device__ inline size_t h16_table_bytes(uint32_t maxSym) { return ((size_t)(maxSym + 1) * 4 + 255) & ~(size_t)255; }

__global__ void __launch_bounds__(kT) k_e16_huf_build(const ozl_h16job* __restrict__ jobs, ozl_h16res* __restrict__ res)
{
    __shared__ H16Lds L;
    const ozl_h16job j = jobs[blockIdx.x];
    const int t = threadIdx.x;
    const uint32_t S = j.maxSym + 1, k = j.card;
    const uint32_t* h = reinterpret_cast<const uint32_t*>((uintptr_t)j.hist);
    uint8_t* ws = reinterpret_cast<uint8_t*>((uintptr_t)j.ws);
    uint32_t* ct = reinterpret_cast<uint32_t*>(ws);
    const uint32_t P = pow2_at_least(k);
    uint64_t* keys;
    uint32_t *cntB, *par;
    uint16_t* sym;
    uint8_t* nb;
    if (k <= kLdsLeaves) {
        keys = L.keys;
        cntB = L.cnt;
        par = L.par;
        sym = L.sym;
        nb = L.nb;
    } else {
        uint8_t* w = ws + h16_table_bytes(j.maxSym);
        keys = reinterpret_cast<uint64_t*>(w); w += (size_t)P * 8;
        cntB = reinterpret_cast<uint32_t*>(w); w += (size_t)(2 * k + 1) * 4;
        par = reinterpret_cast<uint32_t*>(w); w += (size_t)(2 * k) * 4;
        sym = reinterpret_cast<uint16_t*>(w); w += (size_t)k * 2;
        nb = w;
    }
    uint32_t* cnt = cntB + 1;
    // 1. nonzero counts in symbol order as keys (count descending, then symbol)
    {
        const uint32_t per = (S + kT - 1) / kT;
        const uint32_t a = (uint32_t)t * per, b = a + per < S ? a + per : S;
        uint32_t c = 0;
        for (uint32_t s = a; s < b; s++) c += h[s] != 0;
        uint32_t total;
        uint32_t pos = block_excl_scan(c, L.sh, &total);
        for (uint32_t s = a; s < b; s++) {
            const uint32_t x = h[s];
            if (x && pos < k) keys[pos++] = ((uint64_t)(~x) << 16) | s;
        }
        for (uint32_t i = k + t; i < P; i += kT) keys[i] = ~0ull;
        if (total != k) {   // the statistics and the histogram disagree (uniform)
            if (t == 0) res[blockIdx.x].status = OZL_ERR_LOGIC;
            return;
        }
    }
    __syncthreads();
    // 2. sort (encode_huffman_kernel.c:145-229 orders by decreasing count,
    //    equal counts in symbol order: the keys are unique, so any sort agrees)
    bitonic_sort(keys, P);
    // 3. leaves in that order (the parents, which share the keys' LDS, are
    //    written only after the next barrier)
    for (uint32_t i = t; i < k; i += kT) {
        const uint64_t key = keys[i];
        cnt[i] = ~(uint32_t)(key >> 16);
        sym[i] = (uint16_t)(key & 0xFFFF);
    }
    __syncthreads();
    // 4. tree (encode_huffman_kernel.c:231-273) and depths, one lane
    if (t == 0) {
        int maxNbBits = kMaxBits;
        {
            const int bound = ozl_next_pow2((uint64_t)S);
            if (bound + 3 < maxNbBits) maxNbBits = bound + 3;
        }
        const int last = (int)k - 1;
        int nbN = (int)k, lowS = last, lowN = nbN;
        const int root = nbN + lowS - 1;
        cnt[nbN] = cnt[lowS] + cnt[lowS - 1];
        par[lowS] = par[lowS - 1] = (uint32_t)nbN;
        nbN++;
        lowS -= 2;
        for (int q = nbN; q <= root; q++) cnt[q] = 1u << 30;
        cntB[0] = 1u << 31;
        for (; nbN <= root; nbN++) {
            const int a1 = cnt[lowS] < cnt[lowN] ? lowS-- : lowN++;
            const int a2 = cnt[lowS] < cnt[lowN] ? lowS-- : lowN++;
            cnt[nbN] = cnt[a1] + cnt[a2];
            par[a1] = par[a2] = (uint32_t)nbN;
        }
        nb[root] = 0;
        for (int q = root - 1; q >= (int)k; q--) nb[q] = (uint8_t)(nb[par[q]] + 1);
        for (int q = 0; q <= last; q++) nb[q] = (uint8_t)(nb[par[q]] + 1);
        const int tl = h16_set_max_height(nb, cnt, last, maxNbBits);
        L.tl = tl;
        for (int r = 0; r < kMaxBits + 2; r++) L.perRank[r] = 0;
        for (int q = 0; q <= last; q++) L.perRank[nb[q]]++;
        // encode_huffman_kernel.c:275-301: starting value per length, uint16_t
        uint16_t mn = 0;
        for (int r = 0; r < kMaxBits + 2; r++) L.start[r] = 0;
        if (tl <= kMaxBits)
            for (int r = tl; r > 0; r--) {
                L.start[r] = mn;
                mn = (uint16_t)(mn + (uint16_t)L.perRank[r]);
                mn >>= 1;
            }
        ozl_h16res& rr = res[blockIdx.x];
        rr.tableLog = (uint32_t)tl;
        if (tl > kMaxBits) rr.status = OZL_ERR_GENERIC;
    }
    __syncthreads();
    const int tl = L.tl;
    if (tl > kMaxBits) return;
    // 5. weights and the code table: absent symbols first, then the leaves
    uint8_t* wts = reinterpret_cast<uint8_t*>((uintptr_t)j.weights);
    for (uint32_t s = t; s < S; s += kT) {
        ozl_gst1(wts + s, 0);
        ct[s] = 0;
    }
    // leaves re-sorted by (length, symbol): the code of the m-th symbol of a
    // length, in symbol order, is that length's start + m (modulo 2^16)
    for (uint32_t i = t; i < P; i += kT) keys[i] = i < k ? ((uint64_t)nb[i] << 16) | sym[i] : ~0ull;
    __syncthreads();
    bitonic_sort(keys, P);
    for (uint32_t i = t; i < k; i += kT) {
        const uint32_t len = (uint32_t)(keys[i] >> 16), s = (uint32_t)(keys[i] & 0xFFFF);
        // first leaf of this length: leaves of shorter lengths come before
        uint32_t before = 0;
        for (uint32_t r = 1; r < len; r++) before += L.perRank[r];
        const uint16_t code = (uint16_t)(L.start[len] + (uint16_t)(i - before));
        ct[s] = (uint32_t)code | (len << 16);
        ozl_gst1(wts + s, (uint32_t)(tl + 1 - (int)len));
    }
}

// ---------------------------------------------------------------- 16-bit Huffman streams
// Bits of elements b-1 down to a at absolute bit position `bit` of out32:
// the lane's first and last words are shared with its neighbours (OR-ed),
// the words between are its own.
__device__ void put_codes(uint32_t* out32, uint64_t bit, const uint8_t* src, uint32_t a, uint32_t b,
                          const uint32_t* ct)
{
    if (a >= b) return;
    uint64_t word = bit >> 5;
    uint32_t fill = (uint32_t)(bit & 31);
    uint64_t acc = 0;
    bool first = true;
    for (uint32_t i = b; i-- > a;) {
        const uint32_t e = ct[ozl_gld2(src + 2 * (size_t)i)];
        const uint32_t len = e >> 16;
        const uint32_t c = (e & 0xFFFF) & ((1u << len) - 1);
        acc |= (uint64_t)c << fill;
        fill += len;
        while (fill >= 32) {
            const uint32_t w = (uint32_t)acc;
            if (first) atomicOr(&out32[word], w);
            else out32[word] = w;
            first = false;
            word++;
            acc >>= 32;
            fill -= 32;
        }
    }
    if (fill) atomicOr(&out32[word], (uint32_t)acc);
}

// `v` OR-ed into byte position p of out32 (headers and end marks)
__device__ inline void or_byte(uint32_t* out32, size_t p, uint32_t v)
{
    atomicOr(&out32[p >> 2], v << (8 * (p & 3)));
}

__global__ void __launch_bounds__(kT) k_e16_huf_encode(const ozl_h16job* __restrict__ jobs, ozl_h16res* __restrict__ res)
{
    __shared__ uint64_t sTot[4];
    __shared__ uint64_t sStart[4];
    __shared__ int sOk;
    const ozl_h16job j = jobs[blockIdx.x];
    if (res[blockIdx.x].status) return;
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const int ns = j.x4 ? 4 : 1;
    const uint32_t seg = j.x4 ? j.n / 4 + 1 : j.n;
    const uint8_t* src = reinterpret_cast<const uint8_t*>((uintptr_t)j.src);
    const uint32_t* ct = reinterpret_cast<const uint32_t*>((uintptr_t)j.ws);
    // stream w: elements [s0, s0 + cnt)
    uint32_t s0 = 0, cnt = 0;
    if (w < ns) {
        s0 = (uint32_t)w * seg < j.n ? (uint32_t)w * seg : j.n;
        cnt = j.n - s0 < seg ? j.n - s0 : seg;
    }
    const uint32_t per = (cnt + 63) / 64;
    const uint32_t a = s0 + (uint32_t)lane * per < s0 + cnt ? s0 + (uint32_t)lane * per : s0 + cnt;
    const uint32_t b = a + per < s0 + cnt ? a + per : s0 + cnt;
    uint64_t mine = 0;
    for (uint32_t i = a; i < b; i++) mine += ct[ozl_gld2(src + 2 * (size_t)i)] >> 16;
    // bits of the lanes above (later elements sit lower in the stream)
    uint64_t above = mine;
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t x = __shfl_down(above, d, 64);
        if (lane + d < 64) above += x;
    }
    const uint64_t off = above - mine;
    if (lane == 0 && w < ns) sTot[w] = above;
    __syncthreads();
    if (t == 0) {
        // encode_huffman_kernel.c:331-452 with bitstream.h:154-245: each stream
        // needs 8 header bytes, a bit capacity above 8 bytes, and its full
        // bytes 8 bytes below its end
        const uint64_t cap = 2ull * j.n + 32;
        uint64_t pos = 0;
        int ok = 1;
        for (int q = 0; q < ns && ok; q++) {
            if (cap - pos < 8) { ok = 0; break; }
            const uint64_t avail = cap - pos - 8;
            if (avail <= 8 || ((sTot[q] + 1) >> 3) >= avail - 8) { ok = 0; break; }
            sStart[q] = pos;
            pos += 8 + ((sTot[q] + 8) >> 3);
        }
        sOk = ok;
        ozl_h16res& r = res[blockIdx.x];
        if (!ok) r.status = OZL_ERR_GENERIC;
        else r.size = (uint32_t)pos;
    }
    __syncthreads();
    if (!sOk || w >= ns) return;
    uint32_t* out32 = reinterpret_cast<uint32_t*>((uintptr_t)j.out);
    const uint64_t base = sStart[w];
    put_codes(out32, 8 * (base + 8) + off, src, a, b, ct);
    if (lane == 0) {
        const uint64_t T = sTot[w];
        const uint32_t bytes = (uint32_t)((T + 8) >> 3);
        for (int q = 0; q < 4; q++) {
            or_byte(out32, base + q, (cnt >> (8 * q)) & 255);
            or_byte(out32, base + 4 + q, (bytes >> (8 * q)) & 255);
        }
        or_byte(out32, base + 8 + (T >> 3), 1u << (T & 7));   // end mark
    }
}

// ---------------------------------------------------------------- decoders
// decode_entropy_binding.c:348-472 and decode_huffman_kernel.c:13-254.
__global__ void __launch_bounds__(kT) k_e16_huf_decode(const ozl_h16dec* __restrict__ items, int* __restrict__ status)
{
    __shared__ uint32_t cntW[kMaxBits + 1];
    __shared__ uint32_t rangeLo[kMaxBits + 2], symBase[kMaxBits + 2];
    __shared__ uint32_t colT[kMaxBits + 1][kT];
    __shared__ unsigned long long sSum;
    __shared__ uint32_t sNz, sBad;
    __shared__ int sTl;
    __shared__ uint32_t stOff[4], stLen[4], stDst[4], stCnt[4];
    const ozl_h16dec it = items[blockIdx.x];
    const int t = threadIdx.x;
    const uint8_t* wts = reinterpret_cast<const uint8_t*>((uintptr_t)it.weights);
    const uint32_t W = it.nbWeights;
    if (t <= kMaxBits) cntW[t] = 0;
    if (t == 0) { sSum = 0; sNz = 0; sBad = 0; }
    __syncthreads();
    // weights: count, range, Kraft sum
    const uint32_t per = (W + kT - 1) / kT;
    const uint32_t a = (uint32_t)t * per < W ? (uint32_t)t * per : W;
    const uint32_t b = a + per < W ? a + per : W;
    uint32_t my[kMaxBits + 1];
    for (int r = 0; r <= kMaxBits; r++) my[r] = 0;
    unsigned long long sum = 0;
    uint32_t nz = 0, bad = 0;
    for (uint32_t s = a; s < b; s++) {
        const uint32_t x = ozl_gld1(wts + s);
        if (x > kMaxBits) { bad = 1; continue; }
        if (x) { nz++; sum += 1ull << (x - 1); my[x]++; }
    }
    atomicAdd(&sSum, sum);
    atomicAdd(&sNz, nz);
    if (bad) atomicOr(&sBad, 1u);
    for (int r = 1; r <= kMaxBits; r++) {
        colT[r][t] = my[r];
        if (my[r]) atomicAdd(&cntW[r], my[r]);
    }
    __syncthreads();
    if (t == 0) {
        int tl = -1;
        const unsigned long long s = sSum;
        if (W >= 2 && W <= 65536 && !sBad && sNz >= 2 && s && !(s & (s - 1))) tl = 63 - __builtin_clzll(s);
        if (tl > kMaxBits) tl = -1;
        sTl = tl;
        uint32_t lo = 0, sb = 0;
        for (int r = 1; r <= kMaxBits + 1; r++) {
            rangeLo[r] = lo;
            symBase[r] = sb;
            if (r <= kMaxBits) {
                lo += cntW[r] << (r - 1);
                sb += cntW[r];
            }
        }
        // stream headers
        const int ns = it.x4 ? 4 : 1;
        const uint8_t* bits = reinterpret_cast<const uint8_t*>((uintptr_t)it.bits);
        uint64_t ip = 0, op = 0;
        int ok = tl > 0 && it.n >= 2;
        for (int q = 0; q < ns && ok; q++) {
            if (it.bitsLen - ip < 8) { ok = 0; break; }
            uint32_t dn = 0, sn = 0;
            for (int k = 0; k < 4; k++) {
                dn |= ozl_gld1(bits + ip + k) << (8 * k);
                sn |= ozl_gld1(bits + ip + 4 + k) << (8 * k);
            }
            ip += 8;
            if (sn > it.bitsLen - ip || dn > it.n - op || sn == 0 || ozl_gld1(bits + ip + sn - 1) == 0) { ok = 0; break; }
            stOff[q] = (uint32_t)ip;
            stLen[q] = sn;
            stDst[q] = (uint32_t)op;
            stCnt[q] = dn;
            ip += sn;
            op += dn;
        }
        if (ok && (op != it.n || ip != it.bitsLen)) ok = 0;
        if (!ok) sTl = -1;
        status[blockIdx.x] = ok ? 0 : OZL_ERR_CORRUPTION;
    }
    __syncthreads();
    const int tl = sTl;
    if (tl <= 0) return;
    // symbols sorted by (weight, symbol): column scans of the per-thread counts
    if (t >= 1 && t <= kMaxBits) {
        uint32_t acc = 0;
        for (int q = 0; q < kT; q++) {
            const uint32_t x = colT[t][q];
            colT[t][q] = acc;
            acc += x;
        }
    }
    __syncthreads();
    uint16_t* sorted = reinterpret_cast<uint16_t*>((uintptr_t)it.ws);
    for (int r = 1; r <= kMaxBits; r++) my[r] = 0;
    for (uint32_t s = a; s < b; s++) {
        const uint32_t x = ozl_gld1(wts + s);
        if (x) ozl_gst2(sorted + symBase[x] + colT[x][t] + my[x]++, s);
    }
    __syncthreads();
    // one lane per stream, from the end mark down; bits below the stream
    // start read as zero
    const int ns = it.x4 ? 4 : 1;
    if ((t & 63) != 0 || (t >> 6) >= ns) return;
    const int q = t >> 6;
    const uint8_t* s = reinterpret_cast<const uint8_t*>((uintptr_t)it.bits) + stOff[q];
    const uint32_t len = stLen[q];
    int64_t pos = (int64_t)(len - 1) * 8 + ozl_hb32(ozl_gld1(s + len - 1));
    uint8_t* out = reinterpret_cast<uint8_t*>((uintptr_t)it.out) + 2 * (size_t)stDst[q];
    const uint32_t mask = (1u << tl) - 1;
    bool fail = false;
    for (uint32_t i = 0; i < stCnt[q]; i++) {
        const int64_t lo = pos - tl;
        uint32_t v;
        if (lo >= 0) {
            const uint32_t byte = (uint32_t)(lo >> 3);
            uint32_t x = 0;
            for (int k = 0; k < 4; k++)
                if (byte + k < len) x |= ozl_gld1(s + byte + k) << (8 * k);
            v = (x >> (lo & 7)) & mask;
        } else {
            uint32_t x = 0;
            for (int k = 0; k < 4; k++)
                if ((uint32_t)k < len) x |= ozl_gld1(s + k) << (8 * k);
            const int have = (int)pos;   // bits [0, pos)
            v = have > 0 ? ((x & ((1u << have) - 1)) << (tl - have)) & mask : 0;
        }
        int r = tl;
        while (r > 1 && rangeLo[r] > v) r--;
        const uint32_t idx = symBase[r] + ((v - rangeLo[r]) >> (r - 1));
        ozl_gst2(out + 2 * (size_t)i, ozl_gld2(sorted + idx));
        pos -= tl + 1 - r;
        if (pos < 0) { fail = true; break; }
    }
    if (fail && stCnt[q]) {
        // over-read below the stream start: not an encoder's stream
        status[blockIdx.x] = OZL_ERR_CORRUPTION;
    }
}

__global__ void __launch_bounds__(kT) k_tokenize_decode(const uint8_t* __restrict__ alpha, uint32_t card, int w,
                                                        const uint8_t* __restrict__ idx, int iw, size_t n,
                                                        uint8_t* __restrict__ out, int* __restrict__ status)
{
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
        const uint32_t x = (uint32_t)ozl_gld_w(idx + i * iw, iw);
        if (x >= card) {
            atomicOr(status, OZL_ERR_CORRUPTION);
            continue;
        }
        for (int k = 0; k < w; k++) out[i * w + k] = alpha[(size_t)x * w + k];
    }
}

__global__ void __launch_bounds__(kT) k_differs(const uint8_t* __restrict__ src, size_t n, int w, int* __restrict__ flag)
{
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT)
        for (int k = 0; k < w; k++)
            if (src[i * w + k] != src[k]) {
                atomicOr(flag, 1);
                break;
            }
}

__global__ void __launch_bounds__(kT) k_fill_fixed(uint8_t* __restrict__ dst, const uint8_t* __restrict__ val, int w,
                                                   size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n * w; i += (size_t)gridDim.x * kT)
        dst[i] = val[i % w];
}

}  // namespace

int ozl_e16_stats(const ozl_c16* dChunks, uint32_t C, uint32_t maxN, uint32_t* hist, ozl_s16* dStats, hipStream_t st)
{
    if (C == 0) return OZL_OK;
    OZL_HIP_CHECK(hipMemsetAsync(hist, 0, (size_t)C * kSyms * 4, st));
    const unsigned bx = ozl_blocks(maxN, kT * kHistPerThread);
    for (uint32_t y0 = 0; y0 < C; y0 += 65535) {
        const uint32_t cy = C - y0 < 65535 ? C - y0 : 65535;
        OZL_LAUNCH("k_e16_hist", k_e16_hist, dim3(bx, cy), dim3(kT), 0, st, dChunks + y0, hist + (size_t)y0 * kSyms);
    }
    OZL_LAUNCH("k_e16_stats", k_e16_stats, dim3(C), dim3(kT), 0, st, dChunks, hist, dStats);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

int ozl_e16_tokenize(const ozl_tok16* dJobs, uint32_t T, hipStream_t st)
{
    if (T == 0) return OZL_OK;
    OZL_LAUNCH("k_e16_tokenize", k_e16_tokenize, dim3(T), dim3(kT), 0, st, dJobs);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

size_t ozl_e16_huf_ws(uint32_t card, uint32_t maxSym)
{
    size_t b = h16_table_bytes(maxSym);
    if (card > kLdsLeaves) {
        const size_t P = pow2_at_least(card);
        b += P * 8 + (size_t)(2 * card + 1) * 4 + (size_t)(2 * card) * 4 + (size_t)card * 2 + (size_t)(2 * card);
    }
    return (b + 255) & ~(size_t)255;
}

int ozl_e16_huf_encode(const ozl_h16job* dJobs, uint32_t H, ozl_h16res* dRes, hipStream_t st)
{
    if (H == 0) return OZL_OK;
    OZL_HIP_CHECK(hipMemsetAsync(dRes, 0, (size_t)H * sizeof(ozl_h16res), st));
    OZL_LAUNCH("k_e16_huf_build", k_e16_huf_build, dim3(H), dim3(kT), 0, st, dJobs, dRes);
    OZL_LAUNCH("k_e16_huf_encode", k_e16_huf_encode, dim3(H), dim3(kT), 0, st, dJobs, dRes);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

int ozl_e16_huf_decode(const ozl_h16dec* dItems, uint32_t D, int* dStatus, hipStream_t st)
{
    if (D == 0) return OZL_OK;
    OZL_LAUNCH("k_e16_huf_decode", k_e16_huf_decode, dim3(D), dim3(kT), 0, st, dItems, dStatus);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

int ozl_tokenize_decode(const void* alpha, uint32_t card, int w, const void* idx, int iw, size_t n, void* out,
                        int* status, hipStream_t st)
{
    if (n == 0) return OZL_OK;
    OZL_LAUNCH("k_tokenize_decode", k_tokenize_decode, dim3(ozl_blocks(n, kT, 4096)), dim3(kT), 0, st,
               (const uint8_t*)alpha, card, w, (const uint8_t*)idx, iw, n, (uint8_t*)out, status);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

int ozl_fill_fixed(void* dst, const void* val, int w, size_t n, hipStream_t st)
{
    if (n == 0) return OZL_OK;
    OZL_LAUNCH("k_fill_fixed", k_fill_fixed, dim3(ozl_blocks(n * w, kT, 4096)), dim3(kT), 0, st, (uint8_t*)dst,
               (const uint8_t*)val, w, n);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}

// *flag |= 1 when some element of src differs from its first (flag zeroed by the caller)
int ozl_differs_async(const void* src, size_t n, int w, int* flag, hipStream_t st)
{
    if (n <= 1) return OZL_OK;
    OZL_LAUNCH("k_differs", k_differs, dim3(ozl_blocks(n, kT, 4096)), dim3(kT), 0, st, (const uint8_t*)src, n, w, flag);
    OZL_HIP_CHECK(hipGetLastError());
    return OZL_OK;
}
