// Universal decompressor of libozl_hip.so for frames made of the hot-path
// transforms (src/openzl/decompress/decompress2.c:1442-1874 restricted to
// single-input, single-chunk frames of format version 23).
//
// The host reads the frame and chunk headers, rebuilds the stream numbering
// of the encoder (nodes in creation order, singleton then variable outputs,
// decode_frameheader.c), and runs the decoders in frame order on the GPU.
// Independent fse_v2 / huffman_v2 decoders and NCount readers are batched
// into one launch each until a later decoder needs their output.
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "openzl_amd.h"
#include "ozl_internal.h"
#include "ozl_core.h"
#include "ozl_xxh3.h"
#include "ozl_frame.h"
#include "ozl_profile.h"
#include "ozl_engine_util.h"
#include "ozl_hparse.h"
#include "ozl_entropy16_api.h"

extern "C" {
int ozl_delta_decode(const void* first, const void* deltas, size_t n, int width, void* dst, void* scratch, void* stream);
size_t ozl_delta_decode_scratch_size(size_t n, int width);
int ozl_zigzag_decode(const void* src, size_t n, int width, void* dst, void* stream);
int ozl_transpose_join(const void* planes, size_t n, int width, void* dst, void* stream);
int ozl_transpose_join_planes(const void* const* planes, size_t n, int width, void* dst, void* stream);
int ozl_bitpack_decode(const void* src, size_t n, int width, int nbBits, void* dst, void* stream);
}

struct ozl_ditem {
    uint64_t side;
    uint64_t bits;
    uint64_t out;
    uint32_t sideN;
    uint32_t bitsLen;
    uint32_t n;
    uint8_t huf;
    uint8_t nbStates;
    uint8_t pad[2];
};
size_t ozl_entropy_items_workspace(uint32_t B);
int ozl_entropy_decode_items(const ozl_ditem* items, uint32_t B, int* status, void* workspace, size_t wsSize,
                             hipStream_t st);

using namespace ozl_hparse;

namespace {

constexpr int kT = 256;

// constant chunks of one batch: item y fills [pieces of 16 KiB] with its byte
struct FillItem {
    uint64_t dst, n, val;
};
__global__ void __launch_bounds__(kT) k_fill_items(const FillItem* __restrict__ items)
{
    const FillItem it = items[blockIdx.y];
    const uint64_t a = (uint64_t)blockIdx.x * 16384;
    if (a >= it.n) return;
    const uint64_t b = a + 16384 < it.n ? a + 16384 : it.n;
    uint8_t* d = reinterpret_cast<uint8_t*>((uintptr_t)it.dst);
    const uint8_t x = *reinterpret_cast<const uint8_t*>((uintptr_t)it.val);
    for (uint64_t i = a + threadIdx.x; i < b; i += kT) d[i] = x;
}


// flatpack decode (decode_flatpack_kernel.c:359)
__global__ void __launch_bounds__(kT) k_flatpack_decode(const uint8_t* __restrict__ alpha, uint32_t card,
                                                        const uint8_t* __restrict__ pk, size_t n, uint32_t nb,
                                                        uint8_t* __restrict__ dst)
{
    __shared__ uint8_t a[256];
    for (uint32_t i = threadIdx.x; i < 256; i += kT) a[i] = i < card ? alpha[i] : 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
        uint64_t bit = (uint64_t)i * nb;
        uint32_t v = pk[bit >> 3];
        if ((bit & 7) + nb > 8) v |= (uint32_t)pk[(bit >> 3) + 1] << 8;
        v = (v >> (bit & 7)) & ((1u << nb) - 1);
        dst[i] = a[v & 255];
    }
}

__global__ void __launch_bounds__(kT) k_copy(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) dst[i] = src[i];
}

ZL_Report rep_ok(size_t v)
{
    ZL_Report r;
    memset(&r, 0, sizeof(r));
    r._value._value = v;
    return r;
}
// A failed HIP runtime call has no reference code: it surfaces as GENERIC.
ZL_Report rep_err(int code)
{
    ZL_Report r;
    memset(&r, 0, sizeof(r));
    r._code = (ZL_ErrorCode)(code == OZL_ERR_HIP ? ZL_ErrorCode_GENERIC : code);
    return r;
}

struct DStream {
    uint8_t* dev = nullptr;
    size_t bytes = 0;
    int type = 0;   // ZL_Type bit
    int width = 1;
    bool ready = false;
    bool pending = false;
    bool ncount = false;   // still the NCount header (dev, bytes): the FSE item reads it
    bool waiting = false;  // its decoder waits for a queued launch (this round)
};

}  // namespace

// A decompression parameter's applied value: 0 stands for the default
// (enable); only enable turns a check on (decompress2.c:1495, 1538).
inline bool ozl_dparam_on(int v) { return (v ? v : ZL_TernaryParam_enable) == ZL_TernaryParam_enable; }

struct ZL_DCtx_s {
    int sticky = 0, checkCompressed = 0, checkContent = 0;   // requested parameters
    hipStream_t st = nullptr;
    hipStream_t st2 = nullptr;   // compressed checksums, alongside the decoding
    hipEvent_t evIn = nullptr, evHash = nullptr;
    OzlHostStage stage;          // pinned staging of the uploads
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    std::vector<uint8_t*> old;
    ~ZL_DCtx_s()
    {
        if (st2) (void)hipStreamSynchronize(st2);
        if (st) (void)hipStreamSynchronize(st);
        if (base) (void)hipFree(base);
        for (uint8_t* p : old) (void)hipFree(p);
        if (st) (void)hipStreamDestroy(st);
        if (st2) (void)hipStreamDestroy(st2);
        if (evIn) (void)hipEventDestroy(evIn);
        if (evHash) (void)hipEventDestroy(evHash);
    }
    void reset()
    {
        used = 0;
        for (uint8_t* p : old) (void)hipFree(p);
        old.clear();
    }
    uint8_t* alloc(size_t n)
    {
        n = ((n ? n : 1) + 255) & ~(size_t)255;
        if (used + n > cap) {
            size_t ncap = cap * 2 > used + n ? cap * 2 : used + n + (64u << 20);
            uint8_t* nb = nullptr;
            if (hipMalloc(&nb, ncap) != hipSuccess) return nullptr;
            if (base) old.push_back(base);
            base = nb;
            cap = ncap;
            used = 0;
        }
        uint8_t* p = base + used;
        used += n;
        return p;
    }
};

namespace {

// a decoder that needs a stream a queued launch has not produced yet
constexpr int kBlocked = -1;

struct Decoder;

// Work queued by the decoders of one call (one frame or many): entropy items
// and constant chunks go out in shared launches once every decoder waits for
// one of their outputs; the launches' statuses are read once, at the end.
struct DQueue {
    ZL_DCtx_s* d;
    hipStream_t st;
    std::vector<ozl_ditem> items;
    std::vector<FillItem> fills;
    std::vector<std::pair<int*, uint32_t>> statuses;
    std::vector<std::vector<uint8_t>> keep;   // host bytes asynchronous copies read (when pinned staging fails)
    std::vector<DStream*> pend;                // streams produced by queued work
    std::vector<uint64_t> gSrc, gSize, gDst;  // queued copies (device addresses)
    int err = 0;

    uint8_t* alloc(size_t n)
    {
        uint8_t* p = d->alloc(n + 16);
        if (!p && !err) err = OZL_ERR_ALLOCATION;
        return p;
    }
    uint8_t* host_keep(const void* p, size_t n)
    {
        uint8_t* h = d->stage.put(p, n);
        if (h) return h;
        keep.emplace_back((const uint8_t*)p, (const uint8_t*)p + n);
        return keep.back().data();
    }
    int flush();
    // every launched item batch's status (one synchronisation)
    int check_statuses()
    {
        std::vector<std::vector<int>> h(statuses.size());
        for (size_t k = 0; k < statuses.size(); k++) {
            h[k].resize(statuses[k].second);
            OZL_HIP_CHECK(hipMemcpyAsync(h[k].data(), statuses[k].first, statuses[k].second * 4, hipMemcpyDeviceToHost, st));
        }
        OZL_HIP_CHECK(hipStreamSynchronize(st));
        statuses.clear();
        for (auto& v : h)
            for (int x : v)
                if (x) return x;
        return 0;
    }
};

struct Decoder {
    ZL_DCtx_s* d;
    DQueue* q;
    hipStream_t st;
    std::vector<DStream> S;
    int err = 0;

    uint8_t* alloc(size_t n)
    {
        uint8_t* p = d->alloc(n + 16);
        if (!p && !err) err = OZL_ERR_ALLOCATION;
        return p;
    }
    uint8_t* host_keep(const void* p, size_t n) { return q->host_keep(p, n); }

    // the frame's content (stream 0) goes straight to the caller's device
    // buffer when it has the size the frame header announces
    uint8_t* rootDst = nullptr;
    size_t rootSize = 0;
    // No regenerated stream of a hot-path frame is larger than the frame's
    // content plus the largest side stream (16-bit Huffman weights, 65536
    // bytes): a header that announces more is corrupt, and is rejected
    // before anything is allocated or filled.
    size_t limit = 0;
    uint8_t* out_alloc(const DNode& nd, size_t n)
    {
        if (n > limit) {
            if (!err) err = OZL_ERR_CORRUPTION;
            return nullptr;
        }
        if (nd.in == 0 && rootDst && n == rootSize) return rootDst;
        return alloc(n);
    }

    // copies of device ranges into dst at the given offsets, queued: they go
    // out with the other queued work (one launch for all queued copies)
    void gather(const std::vector<const uint8_t*>& srcs, const std::vector<uint64_t>& sizes,
                const std::vector<uint64_t>& offs, uint8_t* dst)
    {
        for (size_t k = 0; k < srcs.size(); k++) {
            if (!sizes[k]) continue;
            q->gSrc.push_back((uint64_t)(uintptr_t)srcs[k]);
            q->gSize.push_back(sizes[k]);
            q->gDst.push_back((uint64_t)(uintptr_t)dst + offs[k]);
        }
    }

    // make sure the given streams hold their final content
    int need(const DNode& nd)
    {
        for (uint32_t k = 0; k < nd.nbOut; k++) {
            const DStream& s = S[nd.firstOut + k];
            if (s.pending || s.waiting) return kBlocked;
            if (!s.ready) return OZL_ERR_CORRUPTION;
        }
        return 0;
    }

    int decode(const DNode& nd)
    {
        DStream& in = S[nd.in];
        const uint32_t o0 = nd.firstOut;
        switch (nd.trid) {
        case OZL_TR_FSE_NCOUNT: {
            // the normalized counts stay in their NCount form: the FSE item
            // that takes them as its side stream reads the header on the
            // device (no size has to come back to the host)
            int rc = need(nd);
            if (rc) return rc;
            if (S[o0].bytes == 0 || S[o0].bytes > 0x7FFFFFFF) return OZL_ERR_CORRUPTION;
            in.dev = S[o0].dev;
            in.bytes = S[o0].bytes;
            in.type = T_NUMERIC_;
            in.width = 2;
            in.ncount = true;
            in.ready = true;
            in.pending = false;
            return 0;
        }
        case OZL_TR_FSE_V2:
        case OZL_TR_HUFFMAN_V2: {
            int rc = need(nd);
            if (rc) return rc;
            if (nd.hdr.size() < 2 || nd.hdr.size() > 9) return OZL_ERR_CORRUPTION;
            uint64_t n = 0;
            for (size_t k = 1; k < nd.hdr.size(); k++) n |= (uint64_t)nd.hdr[k] << (8 * (k - 1));
            if (n > 0xFFFFFFFFull) return OZL_ERR_CORRUPTION;
            const DStream& side = S[o0];
            const DStream& bits = S[o0 + 1];
            in.dev = out_alloc(nd, n);
            if (err) return err;
            in.bytes = n;
            in.type = T_SERIAL_;
            in.width = 1;
            in.pending = true;
            q->pend.push_back(&in);
            ozl_ditem it;
            memset(&it, 0, sizeof(it));
            it.side = (uint64_t)(uintptr_t)side.dev;
            if (side.ncount) {
                if (nd.trid != OZL_TR_FSE_V2) return OZL_ERR_CORRUPTION;
                it.sideN = 0x80000000u | (uint32_t)side.bytes;   // NCount header of sideN & 0x7FFFFFFF bytes
            } else {
                it.sideN = (uint32_t)(nd.trid == OZL_TR_FSE_V2 ? side.bytes / 2 : side.bytes);
            }
            it.bits = (uint64_t)(uintptr_t)bits.dev;
            it.bitsLen = (uint32_t)bits.bytes;
            it.out = (uint64_t)(uintptr_t)in.dev;
            it.n = (uint32_t)n;
            it.huf = nd.trid == OZL_TR_HUFFMAN_V2;
            it.nbStates = nd.hdr[0];
            q->items.push_back(it);
            return 0;
        }
        case OZL_TR_CONV_NUM_TO_STRUCT_LE:
        case OZL_TR_CONV_STRUCT_TO_NUM_LE:
        case OZL_TR_CONV_STRUCT_TO_SERIAL:
        case OZL_TR_CONV_NUM_TO_SERIAL_LE: {
            // byte-identical: the regenerated stream aliases the decoded one,
            // pending or not (no wait for the batch that produces it)
            const DStream& s = S[o0];
            if (s.waiting) return kBlocked;
            if (!s.ready && !s.pending) return OZL_ERR_CORRUPTION;
            if (s.ncount) return OZL_ERR_CORRUPTION;
            int type = T_NUMERIC_, width = s.width;
            if (nd.trid == OZL_TR_CONV_STRUCT_TO_SERIAL) {
                uint64_t w = 0;
                for (size_t k = 0; k < nd.hdr.size() && k < 9; k++) w |= (uint64_t)(nd.hdr[k] & 127) << (7 * k);
                if (w == 0 || s.bytes % w) return OZL_ERR_CORRUPTION;
                type = T_STRUCT_;
                width = (int)w;
            } else if (nd.trid == OZL_TR_CONV_NUM_TO_SERIAL_LE) {
                if (nd.hdr.size() != 1 || nd.hdr[0] > 3) return OZL_ERR_CORRUPTION;
                const int w = 1 << nd.hdr[0];
                if (s.bytes % w) return OZL_ERR_CORRUPTION;
                width = w;
            } else if (nd.trid == OZL_TR_CONV_STRUCT_TO_NUM_LE) {   // numeric -> struct, same width
                if (s.type != T_NUMERIC_) return OZL_ERR_CORRUPTION;
                type = T_STRUCT_;
            }
            in = s;
            in.type = type;
            in.width = width;
            if (in.pending) q->pend.push_back(&in);
            return 0;
        }
        default: break;
        }
        int rc = need(nd);
        if (rc) return rc;
        switch (nd.trid) {
        case OZL_TR_DELTA_INT: {  // decode_delta_binding.c:10-62
            const DStream& dl = S[o0];
            int w = dl.width;
            if (dl.type != T_NUMERIC_) return OZL_ERR_CORRUPTION;
            size_t m = dl.bytes / w;
            size_t n = nd.hdr.empty() ? 0 : m + 1;
            if (!nd.hdr.empty() && nd.hdr.size() != (size_t)w) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, n * w);
            uint8_t* first = alloc(8);
            uint8_t* scratch = alloc(ozl_delta_decode_scratch_size(n, w));
            if (err) return err;
            if (n) {
                OZL_HIP_CHECK(hipMemcpyAsync(first, host_keep(nd.hdr.data(), w), w, hipMemcpyHostToDevice, st));
                rc = ozl_delta_decode(first, dl.dev, n, w, in.dev, scratch, st);
                if (rc) return rc;
            }
            in.bytes = n * w;
            in.type = T_NUMERIC_;
            in.width = w;
            break;
        }
        case OZL_TR_ZIGZAG: {
            const DStream& z = S[o0];
            if (z.type != T_NUMERIC_) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, z.bytes);
            if (err) return err;
            rc = ozl_zigzag_decode(z.dev, z.bytes / z.width, z.width, in.dev, st);
            if (rc) return rc;
            in.bytes = z.bytes;
            in.type = T_NUMERIC_;
            in.width = z.width;
            break;
        }
        case OZL_TR_TRANSPOSE_SPLIT: {
            uint32_t W = nd.nbVO;
            if (W == 0) return OZL_ERR_CORRUPTION;
            size_t n = S[o0].bytes;
            for (uint32_t p = 1; p < W; p++)
                if (S[o0 + p].bytes != n) return OZL_ERR_CORRUPTION;  // decode_transpose_binding.c:54-57
            in.dev = out_alloc(nd, n * W);
            if (err) return err;
            if (n && W <= 8) {   // straight from the planes
                const void* pl[8];
                for (uint32_t p = 0; p < W; p++) pl[p] = S[o0 + p].dev;
                rc = ozl_transpose_join_planes(pl, n, (int)W, in.dev, st);
                if (rc) return rc;
            } else if (n) {      // wider: planes copied together first
                uint8_t* planes = alloc(n * W);
                if (err) return err;
                std::vector<uint64_t> m(3 * (size_t)W);
                for (uint32_t p = 0; p < W; p++) {
                    m[p] = (uint64_t)(uintptr_t)S[o0 + p].dev;
                    m[W + p] = n;
                    m[2 * W + p] = (uint64_t)p * n;
                }
                uint64_t* dm = (uint64_t*)alloc(m.size() * 8);
                if (err) return err;
                OZL_HIP_CHECK(hipMemcpyAsync(dm, host_keep(m.data(), m.size() * 8), m.size() * 8, hipMemcpyHostToDevice, st));
                rc = ozl_gather_sized((const uint8_t* const*)dm, dm + W, dm + 2 * W, (int)W, n, planes, st);
                if (!rc) rc = ozl_transpose_join(planes, n, (int)W, in.dev, st);
                if (rc) return rc;
            }
            in.bytes = n * W;
            in.type = T_STRUCT_;
            in.width = (int)W;
            break;
        }
        case OZL_TR_BITPACK_SERIAL:
        case OZL_TR_BITPACK_INT: {  // decode_bitpack_binding.c:10-60
            const DStream& s = S[o0];
            if (nd.hdr.empty() || nd.hdr.size() > 2) return OZL_ERR_CORRUPTION;
            int w = 1 << (nd.hdr[0] >> 6);
            int nb = (nd.hdr[0] & 63) + 1;
            if (nb > 8 * w || (nd.trid == OZL_TR_BITPACK_SERIAL && w != 1)) return OZL_ERR_CORRUPTION;
            size_t extra = nd.hdr.size() == 2 ? nd.hdr[1] : 0;
            size_t maxElts = s.bytes * 8 / nb;
            if (extra > maxElts) return OZL_ERR_CORRUPTION;
            size_t n = maxElts - extra;
            if ((n * nb + 7) / 8 != s.bytes) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, n * w);
            if (err) return err;
            if (n) {
                rc = ozl_bitpack_decode(s.dev, n, w, nb, in.dev, st);
                if (rc) return rc;
            }
            in.bytes = n * w;
            in.type = nd.trid == OZL_TR_BITPACK_INT ? T_NUMERIC_ : T_SERIAL_;
            in.width = nd.trid == OZL_TR_BITPACK_INT ? w : 1;
            break;
        }
        case OZL_TR_FLATPACK: {
            const DStream& a = S[o0];
            const DStream& pk = S[o0 + 1];
            size_t card = a.bytes;
            if (card > 256) return OZL_ERR_CORRUPTION;
            size_t n = 0;
            uint32_t nb = card <= 1 ? (uint32_t)card : 1 + (uint32_t)ozl_hb32((uint32_t)card - 1);
            if (pk.bytes) {
                uint8_t last = 0;
                OZL_HIP_CHECK(hipMemcpyAsync(&last, pk.dev + pk.bytes - 1, 1, hipMemcpyDeviceToHost, st));
                OZL_HIP_CHECK(hipStreamSynchronize(st));
                if (last == 0 || nb == 0) return OZL_ERR_CORRUPTION;
                n = ((pk.bytes - 1) * 8 + ozl_hb32(last)) / nb;
            }
            in.dev = out_alloc(nd, n);
            if (err) return err;
            if (n)
                OZL_LAUNCH("k_flatpack_decode", k_flatpack_decode, dim3(ozl_blocks(n, kT, 4096)), dim3(kT), 0, st, a.dev,
                           (uint32_t)card, pk.dev, n, nb, in.dev);
            in.bytes = n;
            in.type = T_SERIAL_;
            in.width = 1;
            break;
        }
        case OZL_TR_CONSTANT_SERIAL: {  // decode_constant_binding.c:10
            const DStream& v = S[o0];
            uint64_t n = 0;
            for (size_t k = 0; k < nd.hdr.size() && k < 10; k++) n |= (uint64_t)(nd.hdr[k] & 127) << (7 * k);
            if (v.bytes != 1) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, n);
            if (err) return err;
            in.bytes = n;
            in.type = T_SERIAL_;
            in.width = 1;
            if (n) {
                // batched with the other constant chunks (flush)
                q->fills.push_back({ (uint64_t)(uintptr_t)in.dev, n, (uint64_t)(uintptr_t)v.dev });
                in.ready = false;
                in.pending = true;
                q->pend.push_back(&in);
                return 0;
            }
            break;
        }
        case OZL_TR_SPLITN:
        case OZL_TR_SPLITN_STRUCT:
        case OZL_TR_SPLITN_NUM: {  // decode_splitN_binding.c:16 -- concatenation
            // segments of one width; with none, the width is the header's
            // varint (encode_splitN_binding.c:122-130)
            int w = 1, type = T_SERIAL_;
            if (nd.trid != OZL_TR_SPLITN) {
                type = nd.trid == OZL_TR_SPLITN_STRUCT ? T_STRUCT_ : T_NUMERIC_;
                if (nd.nbVO) {
                    w = S[o0].width;
                    for (uint32_t k = 0; k < nd.nbVO; k++)
                        if (S[o0 + k].width != w || S[o0 + k].type != type) return OZL_ERR_CORRUPTION;
                } else {
                    uint64_t v = 0;
                    for (size_t k = 0; k < nd.hdr.size() && k < 9; k++) v |= (uint64_t)(nd.hdr[k] & 127) << (7 * k);
                    if (v == 0 || v > 0xFFFF) return OZL_ERR_CORRUPTION;
                    w = (int)v;
                }
            }
            size_t total = 0;
            for (uint32_t k = 0; k < nd.nbVO; k++) total += S[o0 + k].bytes;
            in.dev = out_alloc(nd, total);
            if (err) return err;
            std::vector<const uint8_t*> srcs;
            std::vector<uint64_t> sz, offs;
            size_t pos = 0;
            for (uint32_t k = 0; k < nd.nbVO; k++) {
                if (S[o0 + k].bytes) {
                    srcs.push_back(S[o0 + k].dev);
                    sz.push_back(S[o0 + k].bytes);
                    offs.push_back(pos);
                }
                pos += S[o0 + k].bytes;
            }
            gather(srcs, sz, offs, in.dev);
            in.bytes = total;
            in.type = type;
            in.width = w;
            in.ready = false;
            in.pending = true;   // the concatenation goes out with the queue
            q->pend.push_back(&in);
            return 0;
        }
        case OZL_TR_CONSTANT_FIXED: {  // decode_constant_binding.c:10-47
            const DStream& v = S[o0];
            uint64_t n = 0;
            for (size_t k = 0; k < nd.hdr.size() && k < 10; k++) n |= (uint64_t)(nd.hdr[k] & 127) << (7 * k);
            const int w = v.width;
            if (v.type != T_STRUCT_ || w < 1 || v.bytes != (size_t)w || n < 1 || n > limit) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, n * w);
            if (err) return err;
            rc = ozl_fill_fixed(in.dev, v.dev, w, n, st);
            if (rc) return rc;
            in.bytes = n * w;
            in.type = T_STRUCT_;
            in.width = w;
            break;
        }
        case OZL_TR_TOKENIZE_FIXED: {  // decode_tokenize_binding.c:12-37
            const DStream& a = S[o0];
            const DStream& ix = S[o0 + 1];
            const int w = a.width, iw = ix.width;
            if (a.type != T_STRUCT_ || ix.type != T_NUMERIC_ || w < 1 || (iw != 1 && iw != 2 && iw != 4))
                return OZL_ERR_CORRUPTION;
            const size_t card = a.bytes / w, n = ix.bytes / iw;
            if (n > limit / (size_t)w) return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, n * w);
            int* dst = (int*)alloc(4);
            if (err) return err;
            OZL_HIP_CHECK(hipMemsetAsync(dst, 0, 4, st));
            rc = ozl_tokenize_decode(a.dev, (uint32_t)card, w, ix.dev, iw, n, in.dev, dst, st);
            if (rc) return rc;
            q->statuses.push_back({ dst, 1 });
            in.bytes = n * w;
            in.type = T_STRUCT_;
            in.width = w;
            break;
        }
        case OZL_TR_HUFFMAN_STRUCT_V2: {  // decode_entropy_binding.c:415-472
            const DStream& wv = S[o0];
            const DStream& bt = S[o0 + 1];
            if (wv.type != T_NUMERIC_ || wv.width != 1 || nd.hdr.size() < 2 || nd.hdr.size() > 9)
                return OZL_ERR_CORRUPTION;
            uint64_t n = 0;
            for (size_t k = 1; k < nd.hdr.size(); k++) n |= (uint64_t)nd.hdr[k] << (8 * (k - 1));
            if (n < 2 || n > limit / 2 || n > 0xFFFFFFFFull || bt.bytes > 0xFFFFFFFFull || wv.bytes > 65536)
                return OZL_ERR_CORRUPTION;
            in.dev = out_alloc(nd, 2 * n);
            uint8_t* ws = alloc(2 * wv.bytes + 16);
            ozl_h16dec* dIt = (ozl_h16dec*)alloc(sizeof(ozl_h16dec));
            int* dst = (int*)alloc(4);
            if (err) return err;
            ozl_h16dec it;
            it.weights = (uint64_t)(uintptr_t)wv.dev;
            it.bits = (uint64_t)(uintptr_t)bt.dev;
            it.out = (uint64_t)(uintptr_t)in.dev;
            it.ws = (uint64_t)(uintptr_t)ws;
            it.nbWeights = (uint32_t)wv.bytes;
            it.bitsLen = (uint32_t)bt.bytes;
            it.n = (uint32_t)n;
            it.x4 = nd.hdr[0] & 1;
            OZL_HIP_CHECK(hipMemcpyAsync(dIt, host_keep(&it, sizeof(it)), sizeof(it), hipMemcpyHostToDevice, st));
            rc = ozl_e16_huf_decode(dIt, 1, dst, st);
            if (rc) return rc;
            q->statuses.push_back({ dst, 1 });
            in.bytes = 2 * n;
            in.type = T_STRUCT_;
            in.width = 2;
            break;
        }
        default: return OZL_ERR_CORRUPTION;
        }
        in.ready = true;
        in.pending = false;
        return 0;
    }

    static constexpr int T_SERIAL_ = ZL_Type_serial, T_STRUCT_ = ZL_Type_struct, T_NUMERIC_ = ZL_Type_numeric;
};

int DQueue::flush()
{
    if (!gSrc.empty()) {
        // destinations relative to the lowest one
        const size_t G = gSrc.size();
        uint64_t lo = gDst[0];
        for (uint64_t x : gDst) lo = x < lo ? x : lo;
        std::vector<uint64_t> m(3 * G);
        size_t mx = 0;
        for (size_t k = 0; k < G; k++) {
            m[k] = gSrc[k];
            m[G + k] = gSize[k];
            m[2 * G + k] = gDst[k] - lo;
            mx = gSize[k] > mx ? gSize[k] : mx;
        }
        uint64_t* dm = (uint64_t*)alloc(m.size() * 8);
        if (err) return err;
        OZL_HIP_CHECK(hipMemcpyAsync(dm, host_keep(m.data(), m.size() * 8), m.size() * 8, hipMemcpyHostToDevice, st));
        int rc = ozl_gather_sized((const uint8_t* const*)dm, dm + G, dm + 2 * G, (int)G, mx, (uint8_t*)(uintptr_t)lo, st);
        if (rc) return rc;
        gSrc.clear();
        gSize.clear();
        gDst.clear();
    }
    if (!fills.empty()) {
        const uint32_t B = (uint32_t)fills.size();
        uint64_t mx = 0;
        for (const FillItem& f : fills) mx = f.n > mx ? f.n : mx;
        FillItem* dF = (FillItem*)alloc(B * sizeof(FillItem));
        if (err) return err;
        OZL_HIP_CHECK(hipMemcpyAsync(dF, host_keep(fills.data(), B * sizeof(FillItem)), B * sizeof(FillItem),
                                     hipMemcpyHostToDevice, st));
        for (uint32_t y0 = 0; y0 < B; y0 += 65535) {
            const uint32_t cnt = B - y0 < 65535 ? B - y0 : 65535;
            OZL_LAUNCH("k_fill_items", k_fill_items, dim3((unsigned)((mx + 16383) / 16384), cnt), dim3(kT), 0, st,
                       dF + y0);
        }
        fills.clear();
    }
    if (!items.empty()) {
        // launched in stream order without waiting: later nodes that read
        // these outputs are enqueued after them. Items of one coder sit
        // together (each coder's kernel packs consecutive items into a
        // workgroup), longer ones first (by powers of two: a counting sort,
        // since a call may queue tens of thousands), so the long state chains
        // start first.
        {
            auto key = [](const ozl_ditem& it) {
                const uint32_t lg = it.n ? 31u - (uint32_t)__builtin_clz(it.n) : 0u;
                return (it.huf ? 32u : 0u) + (31u - lg);
            };
            uint32_t start[65] = { 0 };
            for (const ozl_ditem& it : items) start[key(it) + 1]++;
            for (int k = 0; k < 64; k++) start[k + 1] += start[k];
            std::vector<ozl_ditem> sorted(items.size());
            for (const ozl_ditem& it : items) sorted[start[key(it)]++] = it;
            items.swap(sorted);
        }
        const uint32_t B = (uint32_t)items.size();
        ozl_ditem* dItems = (ozl_ditem*)alloc(B * sizeof(ozl_ditem));
        int* dStatus = (int*)alloc(B * 4);
        const size_t wsz = ozl_entropy_items_workspace(B);
        uint8_t* ws = alloc(wsz);
        uint8_t* hItems = host_keep(items.data(), B * sizeof(ozl_ditem));
        if (err) return err;
        OZL_HIP_CHECK(hipMemcpyAsync(dItems, hItems, B * sizeof(ozl_ditem), hipMemcpyHostToDevice, st));
        int rc = ozl_entropy_decode_items(dItems, B, dStatus, ws, wsz, st);
        if (rc) return rc;
        statuses.push_back({ dStatus, B });
        items.clear();
    }
    // everything pending is now enqueued (conversions alias their producer's
    // buffer, so they were pending too)
    for (DStream* x : pend)
        if (x->pending) { x->pending = false; x->ready = true; }
    pend.clear();
    return 0;
}

// One chunk of a decompression call: its decoder and where its content goes.
struct ChunkJob {
    size_t frame;
    const ParsedChunk* C;
    std::unique_ptr<Decoder> D;
    std::vector<uint32_t> todo;   // nodes still to decode, in decode order (last node first)
    uint8_t tail[8];              // checksums
    uint64_t outOff = 0;          // in the destination
};

// One frame of a decompression call.
struct FrameJob {
    const uint8_t* src;    // the frame (host or device memory)
    size_t size;
    bool dev;              // src is device memory
    ParsedFrame P;
    uint8_t* dFrame = nullptr;
    size_t pos = 0;        // parse position
    bool done = false;
    size_t win = 0;        // window length of a device frame
    int wantType = 0;
};

// Decompresses n frames (host or device memory) into dst (host or device
// memory) back to back: the host reads every frame's headers (windows of
// device frames come over in batched copies); the decoders of all chunks of
// all frames run in lockstep, each until it needs a stream a queued launch
// still has to produce, then the queue goes out as shared launches;
// checksums are verified in batches; one synchronisation at the end.
int decompress_frames(ZL_DCtx_s* dc, const uint8_t* src, const size_t* frameSizes, size_t n, void* dst, size_t cap,
                      size_t* outSizes, ZL_OutputInfo* infos)
{
    if (!dc->st && hipStreamCreateWithFlags(&dc->st, hipStreamNonBlocking) != hipSuccess) return OZL_ERR_ALLOCATION;
    if (!dc->st2 && hipStreamCreateWithFlags(&dc->st2, hipStreamNonBlocking) != hipSuccess) return OZL_ERR_ALLOCATION;
    if (!dc->evIn && hipEventCreateWithFlags(&dc->evIn, hipEventDisableTiming) != hipSuccess) return OZL_ERR_ALLOCATION;
    if (!dc->evHash && hipEventCreateWithFlags(&dc->evHash, hipEventDisableTiming) != hipSuccess) return OZL_ERR_ALLOCATION;
    hipStream_t st = dc->st;
    // checksums are verified when their parameter resolves to enable
    // (decompress2.c:1492-1545, gdparams.c:8-45)
    const bool vK = ozl_dparam_on(dc->checkCompressed), vC = ozl_dparam_on(dc->checkContent);
    // a call that failed part-way may have left work running
    OZL_HIP_CHECK(hipStreamSynchronize(dc->st2));
    OZL_HIP_CHECK(hipStreamSynchronize(st));
    dc->reset();
    dc->stage.reset();
    // every return (errors included) leaves no work of this call running:
    // the kernels read the caller's buffers
    OzlStreamsDone done{ st, dc->st2 };
    std::vector<FrameJob> J(n);
    const bool devIn = n && ozl_is_device_ptr(src);
    size_t off = 0;
    for (size_t i = 0; i < n; i++) {
        J[i].src = src + off;
        J[i].size = frameSizes[i];
        J[i].dev = devIn;
        off += frameSizes[i];
    }
    // headers of every chunk
    if (!devIn) {
        std::vector<int> prc(n, 0);
        ozl_parallel_for(n, [&](size_t i) {
            bool longer = false;
            prc[i] = parse_window(J[i].src, 0, J[i].size, J[i].size, J[i].P, J[i].pos, J[i].done, longer);
        });
        for (int rc : prc)
            if (rc) return rc == kNeedMore ? OZL_ERR_CORRUPTION : rc;
    } else {
        // windows of the device frames (64 KiB, longer where a chunk's
        // headers need it), one gather and one copy per round
        constexpr size_t kWin = (size_t)64 << 10;
        for (FrameJob& j : J) j.win = j.size < kWin ? j.size : kWin;
        for (;;) {
            std::vector<size_t> L;
            for (size_t i = 0; i < n; i++)
                if (!J[i].done) L.push_back(i);
            if (L.empty()) break;
            std::vector<uint64_t> m(3 * L.size());
            std::vector<size_t> wOff(L.size()), wLen(L.size());
            size_t tot = 0, mx = 0;
            for (size_t k = 0; k < L.size(); k++) {
                FrameJob& j = J[L[k]];
                if (j.pos >= j.size) return OZL_ERR_CORRUPTION;   // no end marker
                const size_t len = j.size - j.pos < j.win ? j.size - j.pos : j.win;
                m[k] = (uint64_t)(uintptr_t)(j.src + j.pos);
                m[L.size() + k] = len;
                m[2 * L.size() + k] = tot;
                wOff[k] = tot;
                wLen[k] = len;
                tot += len;
                mx = len > mx ? len : mx;
            }
            // the windows land in pinned memory: straight copies for a few
            // frames (a frame of many chunks takes one round per chunk), a
            // gather and one copy for many
            uint8_t* hw = dc->stage.put(nullptr, tot + 1);
            if (!hw) return OZL_ERR_ALLOCATION;
            if (L.size() <= 4) {
                for (size_t k = 0; k < L.size(); k++)
                    OZL_HIP_CHECK(hipMemcpyAsync(hw + wOff[k], (const void*)(uintptr_t)m[k], wLen[k],
                                                 hipMemcpyDeviceToHost, st));
            } else {
                uint8_t* dWin = dc->alloc(tot + 16);
                uint64_t* dm = (uint64_t*)dc->alloc(m.size() * 8 + 16);
                const uint8_t* hm = dc->stage.put(m.data(), m.size() * 8);
                if (!dWin || !dm || !hm) return OZL_ERR_ALLOCATION;
                OZL_HIP_CHECK(hipMemcpyAsync(dm, hm, m.size() * 8, hipMemcpyHostToDevice, st));
                int rc = ozl_gather_sized((const uint8_t* const*)dm, dm + L.size(), dm + 2 * L.size(), (int)L.size(),
                                          mx, dWin, st);
                if (rc) return rc;
                OZL_HIP_CHECK(hipMemcpyAsync(hw, dWin, tot, hipMemcpyDeviceToHost, st));
            }
            OZL_HIP_CHECK(hipStreamSynchronize(st));
            std::vector<int> prc(L.size(), 0);
            std::vector<char> lg(L.size(), 0);
            ozl_parallel_for(L.size(), [&](size_t k) {
                FrameJob& j = J[L[k]];
                bool longer = false;
                prc[k] = parse_window(hw + wOff[k], j.pos, wLen[k], j.size, j.P, j.pos, j.done, longer, true);
                lg[k] = longer;
            });
            for (size_t k = 0; k < L.size(); k++) {
                FrameJob& j = J[L[k]];
                if (prc[k] == kNeedMore) {
                    if (lg[k]) {
                        if (wLen[k] >= j.size - j.pos) return OZL_ERR_CORRUPTION;
                        j.win *= 4;
                    }
                    continue;
                }
                if (prc[k]) return prc[k];
            }
        }
    }
    // chunks: decoders, stored streams, checksum tails
    std::vector<ChunkJob> U;
    for (size_t i = 0; i < n; i++)
        for (const ParsedChunk& C : J[i].P.chunks) {
            ChunkJob u;
            u.frame = i;
            u.C = &C;
            U.push_back(std::move(u));
        }
    DQueue Q;
    Q.d = dc;
    Q.st = st;
    for (FrameJob& j : J) {
        if (j.dev) {
            j.dFrame = const_cast<uint8_t*>(j.src);   // read in place
        } else {
            j.dFrame = dc->alloc(j.size + 16);
            if (!j.dFrame) return OZL_ERR_ALLOCATION;
            OZL_HIP_CHECK(hipMemcpyAsync(j.dFrame, j.src, j.size, hipMemcpyHostToDevice, st));
        }
    }
    // each chunk's decoder (host work only; one chunk per worker)
    ozl_parallel_for(U.size(), [&](size_t ui) {
        ChunkJob& u = U[ui];
        const ParsedChunk& C = *u.C;
        const FrameJob& j = J[u.frame];
        u.D.reset(new Decoder());
        Decoder& D = *u.D;
        D.d = dc;
        D.q = &Q;
        D.st = st;
        D.S.assign(C.nbStreams, DStream());
        D.limit = j.P.size + ((size_t)1 << 17);
        size_t pos = C.dataStart;
        for (size_t k = 0; k < C.storedId.size(); k++) {
            DStream& x = D.S[C.storedId[k]];
            x.dev = j.dFrame + pos;
            x.bytes = C.storedSize[k];
            x.type = ZL_Type_serial;
            x.width = 1;
            x.ready = true;
            pos += C.storedSize[k];
        }
        // stored numeric streams keep their type: side inputs of fse_v2
        // (int16) and huffman_v2 (u8) when stored raw
        for (auto& nd : C.nodes) {
            if (nd.trid == OZL_TR_FSE_V2 && D.S[nd.firstOut].ready) { D.S[nd.firstOut].type = ZL_Type_numeric; D.S[nd.firstOut].width = 2; }
            if (nd.trid == OZL_TR_HUFFMAN_V2 && D.S[nd.firstOut].ready) { D.S[nd.firstOut].type = ZL_Type_numeric; }
        }
        for (size_t k = C.nodes.size(); k-- > 0;) u.todo.push_back((uint32_t)k);
        if (!j.dev) memcpy(u.tail, j.src + C.csumPos, C.end - C.csumPos);
    });
    // a device frame's checksum tails, 16 bytes a chunk; after the last
    // chunk of a frame whose end marker the parse did not read, that byte too
    std::vector<uint8_t> tails;
    uint8_t* dTails = nullptr;
    if (devIn && !U.empty()) {
        const size_t nu = U.size();
        std::vector<uint64_t> m(3 * nu);
        for (size_t k = 0; k < nu; k++) {
            const FrameJob& j = J[U[k].frame];
            const bool withEnd = j.P.endDeferred && U[k].C == &j.P.chunks.back();
            m[k] = (uint64_t)(uintptr_t)(j.src + U[k].C->csumPos);
            m[nu + k] = U[k].C->end - U[k].C->csumPos + (withEnd ? 1 : 0);
            m[2 * nu + k] = k * 16;
        }
        dTails = dc->alloc(nu * 16 + 16);
        uint64_t* dm = (uint64_t*)dc->alloc(m.size() * 8 + 16);
        if (!dTails || !dm) return OZL_ERR_ALLOCATION;
        OZL_HIP_CHECK(hipMemcpyAsync(dm, Q.host_keep(m.data(), m.size() * 8), m.size() * 8, hipMemcpyHostToDevice, st));
        int rc = ozl_gather_sized((const uint8_t* const*)dm, dm + nu, dm + 2 * nu, (int)nu, 9, dTails, st);
        if (rc) return rc;
        tails.resize(nu * 16);
    }
    // compressed checksums (they read only the frames) on the second stream
    std::vector<uint64_t> kJobs;
    std::vector<size_t> kIdx;
    size_t kSum = 0, kMax = 0;
    for (size_t k = 0; k < U.size(); k++) {
        const FrameJob& j = J[U[k].frame];
        const ParsedChunk& C = *U[k].C;
        if (!(j.P.flags & 2) || !vK) continue;
        const size_t clen = C.csumPos + ((j.P.flags & 1) ? 4 : 0) - C.start;
        kJobs.insert(kJobs.end(), { (uint64_t)(uintptr_t)(j.dFrame + C.start), clen, kSum });
        kIdx.push_back(k);
        kSum += ozl_xxh3_batch_sums(clen);
        kMax = clen > kMax ? clen : kMax;
    }
    uint64_t* dHK = nullptr;
    if (!kIdx.empty()) {
        uint64_t* dJobs = (uint64_t*)dc->alloc(kJobs.size() * 8 + 16);
        uint64_t* dSums = (uint64_t*)dc->alloc(kSum * 8 + 16);
        dHK = (uint64_t*)dc->alloc(kIdx.size() * 8 + 16);
        if (!dJobs || !dSums || !dHK) return OZL_ERR_ALLOCATION;
        OZL_HIP_CHECK(hipMemcpyAsync(dJobs, Q.host_keep(kJobs.data(), kJobs.size() * 8), kJobs.size() * 8,
                                     hipMemcpyHostToDevice, st));
        OZL_HIP_CHECK(hipEventRecord(dc->evIn, st));
        OZL_HIP_CHECK(hipStreamWaitEvent(dc->st2, dc->evIn, 0));
        int rc = ozl_xxh3_device_batch((const ozl_xjob*)dJobs, (uint32_t)kIdx.size(), kMax, dSums, dHK, dc->st2);
        if (rc) return rc;
        OZL_HIP_CHECK(hipEventRecord(dc->evHash, dc->st2));
    }
    // output places: frames back to back; a frame of one chunk whose place
    // in a device destination is 16-byte aligned (as the library's own
    // buffers are) is decoded in place
    std::vector<uint64_t> frameOut(n + 1, 0);
    for (size_t i = 0; i < n; i++) frameOut[i + 1] = frameOut[i] + J[i].P.size;
    const bool devOut = frameOut[n] && frameOut[n] <= cap && ozl_is_device_ptr(dst);
    if (devOut)
        for (ChunkJob& u : U)
            if (J[u.frame].P.chunks.size() == 1 && (((uintptr_t)dst + frameOut[u.frame]) & 15) == 0) {
                u.D->rootDst = (uint8_t*)dst + frameOut[u.frame];
                u.D->rootSize = J[u.frame].P.size;
            }
    // the decoders in lockstep: every node runs as soon as the streams it
    // reads are ready or queued; a node that waits for a queued launch (or
    // for such a node) is retried once the queue has gone out
    for (;;) {
        bool left = false;
        for (ChunkJob& u : U) {
            Decoder& D = *u.D;
            std::vector<uint32_t> rest;
            for (uint32_t k : u.todo) {
                const DNode& nd = u.C->nodes[k];
                int rc = D.decode(nd);
                if (rc == kBlocked) {
                    D.S[nd.in].waiting = true;
                    rest.push_back(k);
                    continue;
                }
                if (rc) return rc;
                if (D.err) return D.err;
            }
            for (uint32_t k : rest) D.S[u.C->nodes[k].in].waiting = false;
            u.todo.swap(rest);
            if (!u.todo.empty()) left = true;
        }
        int rc = Q.flush();
        if (!rc) rc = Q.err;
        if (rc) return rc;
        if (!left) break;
    }
    // outputs, content checksum jobs
    std::vector<uint64_t> got(n, 0);
    std::vector<int> width(n, 0);
    std::vector<uint64_t> cJobs;
    std::vector<size_t> cIdx;
    size_t cSum = 0, cMax = 0;
    for (size_t i = 0; i < n; i++)
        J[i].wantType = J[i].P.type == OZL_T_SERIAL ? ZL_Type_serial
                        : J[i].P.type == OZL_T_STRUCT ? ZL_Type_struct : ZL_Type_numeric;
    for (size_t k = 0; k < U.size(); k++) {
        ChunkJob& u = U[k];
        FrameJob& j = J[u.frame];
        DStream& out = u.D->S[0];
        if (u.C->nodes.empty()) out.type = j.wantType;   // stored input
        // a stored input keeps the frame's type; a regenerated one must match
        if (!u.C->nodes.empty() && out.type != j.wantType && j.wantType != ZL_Type_serial) return OZL_ERR_CORRUPTION;
        if (!width[u.frame]) width[u.frame] = out.width;
        u.outOff = frameOut[u.frame] + got[u.frame];
        got[u.frame] += out.bytes;
        if (got[u.frame] > j.P.size) return OZL_ERR_CORRUPTION;
        if ((j.P.flags & 1) && vC) {
            const uint8_t* p = out.dev ? out.dev : j.dFrame;
            cJobs.insert(cJobs.end(), { (uint64_t)(uintptr_t)p, out.bytes, cSum });
            cIdx.push_back(k);
            cSum += ozl_xxh3_batch_sums(out.bytes);
            cMax = out.bytes > cMax ? out.bytes : cMax;
        }
    }
    for (size_t i = 0; i < n; i++)
        if (got[i] != J[i].P.size) return OZL_ERR_CORRUPTION;
    if (frameOut[n] > cap) return OZL_ERR_DSTCAPACITY_TOOSMALL;
    std::vector<uint64_t> hC(cIdx.size()), hK(kIdx.size());
    if (!cIdx.empty()) {
        uint64_t* dJobs = (uint64_t*)dc->alloc(cJobs.size() * 8 + 16);
        uint64_t* dSums = (uint64_t*)dc->alloc(cSum * 8 + 16);
        uint64_t* dHC = (uint64_t*)dc->alloc(cIdx.size() * 8 + 16);
        if (!dJobs || !dSums || !dHC) return OZL_ERR_ALLOCATION;
        OZL_HIP_CHECK(hipMemcpyAsync(dJobs, Q.host_keep(cJobs.data(), cJobs.size() * 8), cJobs.size() * 8,
                                     hipMemcpyHostToDevice, st));
        int rc = ozl_xxh3_device_batch((const ozl_xjob*)dJobs, (uint32_t)cIdx.size(), cMax, dSums, dHC, st);
        if (rc) return rc;
        OZL_HIP_CHECK(hipMemcpyAsync(hC.data(), dHC, cIdx.size() * 8, hipMemcpyDeviceToHost, st));
    }
    if (!kIdx.empty()) {
        OZL_HIP_CHECK(hipStreamWaitEvent(st, dc->evHash, 0));
        OZL_HIP_CHECK(hipMemcpyAsync(hK.data(), dHK, kIdx.size() * 8, hipMemcpyDeviceToHost, st));
    }
    const bool dstDev = frameOut[n] && ozl_is_device_ptr(dst);
    for (ChunkJob& u : U) {
        const DStream& out = u.D->S[0];
        if (!out.bytes || out.dev == (uint8_t*)dst + u.outOff) continue;
        OZL_HIP_CHECK(hipMemcpyAsync((uint8_t*)dst + u.outOff, out.dev, out.bytes,
                                     dstDev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
    }
    if (devIn && !U.empty()) OZL_HIP_CHECK(hipMemcpyAsync(tails.data(), dTails, U.size() * 16, hipMemcpyDeviceToHost, st));
    int rc = Q.check_statuses();   // synchronises the stream
    if (rc) return rc;
    if (devIn)
        for (size_t k = 0; k < U.size(); k++) {
            const size_t tl = U[k].C->end - U[k].C->csumPos;
            memcpy(U[k].tail, tails.data() + k * 16, tl);
            const FrameJob& j = J[U[k].frame];
            if (j.P.endDeferred && U[k].C == &j.P.chunks.back() && tails[k * 16 + tl] != 0) return OZL_ERR_CORRUPTION;
        }
    for (size_t k = 0; k < kIdx.size(); k++) {
        const ChunkJob& u = U[kIdx[k]];
        if ((uint32_t)hK[k] != ozl_le32(u.tail + ((J[u.frame].P.flags & 1) ? 4 : 0))) return OZL_ERR_COMPRESSED_CHECKSUM;
    }
    for (size_t k = 0; k < cIdx.size(); k++)
        if ((uint32_t)hC[k] != ozl_le32(U[cIdx[k]].tail)) return OZL_ERR_CONTENT_CHECKSUM;
    for (size_t i = 0; i < n; i++) {
        outSizes[i] = J[i].P.size;
        if (infos) {
            infos[i].type = (ZL_Type)J[i].wantType;
            infos[i].fixedWidth = (uint32_t)(J[i].wantType == ZL_Type_serial || !width[i] ? 1 : width[i]);
            infos[i].decompressedByteSize = J[i].P.size;
            infos[i].numElts = J[i].P.size / infos[i].fixedWidth;
        }
    }
    return 0;
}

int decompress_impl(ZL_DCtx_s* dc, const uint8_t* src, size_t srcSize, void* dst, size_t cap, ZL_OutputInfo* info)
{
    size_t outSize = 0;
    return decompress_frames(dc, src, &srcSize, 1, dst, cap, &outSize, info);
}

// Parameters requested on the context are dropped at the end of every
// decompression unless they are sticky (decompress2.c:1764-1767).
void dctx_end(ZL_DCtx_s* d)
{
    if (!d->sticky) d->checkCompressed = d->checkContent = 0;
}

// Frame header of a host frame: type, size, format version.
int host_frame_info(const void* src, size_t srcSize, ParsedFrame& P)
{
    if (!src) return ZL_ErrorCode_parameter_invalid;
    HostReader r = { (const uint8_t*)src, 0, 0, srcSize, false };
    return parse_frame_header(r, P);
}

ZL_Type frame_ztype(int t) { return t == OZL_T_SERIAL ? ZL_Type_serial : t == OZL_T_STRUCT ? ZL_Type_struct : ZL_Type_numeric; }

}  // namespace

// An output buffer of ZL_DCtx_decompressTBuffer: either its own host memory,
// sized from the frame header, or a caller's buffer of a declared type.
struct ZL_TypedBuffer_s {
    std::vector<uint8_t> own;
    void* wrap = nullptr;
    size_t cap = 0;
    int wrapType = 0;      // 0: not wrapped
    size_t wrapWidth = 0;
    ZL_Type type = ZL_Type_serial;
    size_t width = 1, bytes = 0;
};

struct ZL_FrameInfo {
    ZL_Type type;
    uint64_t size;
};

extern "C" {

ZL_DCtx* ZL_DCtx_create(void) { return new (std::nothrow) ZL_DCtx_s(); }
void ZL_DCtx_free(ZL_DCtx* d) { delete d; }

// gdparams.c:15-33: unknown parameters are compressionParameter_invalid.
ZL_Report ZL_DCtx_setParameter(ZL_DCtx* d, ZL_DParam p, int v)
{
    if (!d) return rep_err(ZL_ErrorCode_parameter_invalid);
    switch (p) {
    case ZL_DParam_stickyParameters: d->sticky = v != 0; break;
    case ZL_DParam_checkCompressedChecksum: d->checkCompressed = v; break;
    case ZL_DParam_checkContentChecksum: d->checkContent = v; break;
    default: return rep_err(ZL_ErrorCode_compressionParameter_invalid);
    }
    return rep_ok(0);
}

int ZL_DCtx_getParameter(const ZL_DCtx* d, ZL_DParam p)
{
    if (!d) return 0;
    switch (p) {
    case ZL_DParam_stickyParameters: return d->sticky;
    case ZL_DParam_checkCompressedChecksum: return d->checkCompressed;
    case ZL_DParam_checkContentChecksum: return d->checkContent;
    default: return 0;
    }
}

ZL_Report ZL_DCtx_resetParameters(ZL_DCtx* d)
{
    if (!d) return rep_err(ZL_ErrorCode_parameter_invalid);
    d->sticky = d->checkCompressed = d->checkContent = 0;
    return rep_ok(0);
}

const char* ZL_DCtx_getErrorContextString(const ZL_DCtx*, ZL_Report r) { return ZL_ErrorCode_toString(r._code); }
const char* ZL_DCtx_getErrorContextString_fromError(const ZL_DCtx*, ZL_Error e) { return ZL_ErrorCode_toString(e._code); }

ZL_Report ZL_DCtx_decompressTyped(ZL_DCtx* dctx, ZL_OutputInfo* info, void* dst, size_t cap, const void* src,
                                  size_t srcSize)
{
    if (!dctx || !src) return rep_err(ZL_ErrorCode_parameter_invalid);
    ZL_OutputInfo tmp;
    int rc = decompress_impl(dctx, (const uint8_t*)src, srcSize, dst, cap, &tmp);
    dctx_end(dctx);
    if (rc) return rep_err(rc);
    if (info) *info = tmp;
    return rep_ok(tmp.decompressedByteSize);
}

// decompress2.c:1845-1866: only frames of serial content.
ZL_Report ZL_DCtx_decompress(ZL_DCtx* dctx, void* dst, size_t cap, const void* src, size_t srcSize)
{
    ZL_OutputInfo info;
    ZL_Report r = ZL_DCtx_decompressTyped(dctx, &info, dst, cap, src, srcSize);
    if (!ZL_isError(r) && info.type != ZL_Type_serial) return rep_err(ZL_ErrorCode_GENERIC);
    return r;
}

ZL_Report ozl_DCtx_decompressBatch(ZL_DCtx* dctx, void* dst, size_t dstCapacity, size_t* outSizes, const void* src,
                                   const size_t* frameSizes, size_t nbFrames)
{
    if (!dctx || !outSizes || (nbFrames && (!src || !frameSizes))) return rep_err(ZL_ErrorCode_parameter_invalid);
    if (nbFrames == 0) return rep_ok(0);
    int rc = decompress_frames(dctx, (const uint8_t*)src, frameSizes, nbFrames, dst, dstCapacity, outSizes, nullptr);
    dctx_end(dctx);
    if (rc) return rep_err(rc);
    size_t total = 0;
    for (size_t i = 0; i < nbFrames; i++) total += outSizes[i];
    return rep_ok(total);
}

// decompress2.c:1868-1882
ZL_Report ZL_decompress(void* dst, size_t cap, const void* src, size_t srcSize)
{
    ZL_DCtx_s d;
    return ZL_DCtx_decompress(&d, dst, cap, src, srcSize);
}

ZL_TypedBuffer* ZL_TypedBuffer_create(void) { return new (std::nothrow) ZL_TypedBuffer_s(); }
void ZL_TypedBuffer_free(ZL_TypedBuffer* t) { delete t; }

static ZL_TypedBuffer* wrap_buffer(void* p, size_t cap, int type, size_t width)
{
    if (!p && cap) return nullptr;
    ZL_TypedBuffer_s* t = new (std::nothrow) ZL_TypedBuffer_s();
    if (!t) return nullptr;
    t->wrap = p;
    t->cap = cap;
    t->wrapType = type;
    t->wrapWidth = width;
    return t;
}
ZL_TypedBuffer* ZL_TypedBuffer_createWrapSerial(void* b, size_t cap) { return wrap_buffer(b, cap, ZL_Type_serial, 1); }
ZL_TypedBuffer* ZL_TypedBuffer_createWrapStruct(void* b, size_t w, size_t n)
{
    return w ? wrap_buffer(b, w * n, ZL_Type_struct, w) : nullptr;
}
ZL_TypedBuffer* ZL_TypedBuffer_createWrapNumeric(void* b, size_t w, size_t n)
{
    return (w == 1 || w == 2 || w == 4 || w == 8) ? wrap_buffer(b, w * n, ZL_Type_numeric, w) : nullptr;
}

// decompress2.c:1581-1700: a wrapped buffer must be large enough and, once
// the frame's output type and width are known, of that type and width
// (stream.c:146-160 reports a mismatch as corruption).
ZL_Report ZL_DCtx_decompressTBuffer(ZL_DCtx* dctx, ZL_TypedBuffer* t, const void* src, size_t srcSize)
{
    if (!dctx || !t || !src) return rep_err(ZL_ErrorCode_parameter_invalid);
    void* dst = t->wrap;
    size_t cap = t->cap;
    if (!t->wrapType) {
        size_t sz = 0;
        if (ozl_is_device_ptr(src)) {
            ZL_Report r = ZL_getDecompressedSize(src, srcSize);
            if (ZL_isError(r)) { dctx_end(dctx); return r; }
            sz = ZL_validResult(r);
        } else {
            ParsedFrame P;
            int rc = host_frame_info(src, srcSize, P);
            if (rc) { dctx_end(dctx); return rep_err(rc); }
            sz = P.size;
        }
        t->own.assign(sz ? sz : 1, 0);
        dst = t->own.data();
        cap = sz;
    }
    ZL_OutputInfo info;
    ZL_Report r = ZL_DCtx_decompressTyped(dctx, &info, dst, cap, src, srcSize);
    if (ZL_isError(r)) return r;
    if (t->wrapType && (info.type != (ZL_Type)t->wrapType || info.fixedWidth != t->wrapWidth))
        return rep_err(ZL_ErrorCode_corruption);
    t->type = info.type;
    t->width = info.fixedWidth;
    t->bytes = info.decompressedByteSize;
    return rep_ok(t->bytes);
}

ZL_Report ZL_DCtx_decompressMultiTBuffer(ZL_DCtx* dctx, ZL_TypedBuffer* outputs[], size_t nbOutputs,
                                         const void* src, size_t srcSize)
{
    if (!dctx || !outputs) return rep_err(ZL_ErrorCode_parameter_invalid);
    // decompress2.c:839-846: the exact number of outputs of the frame
    if (nbOutputs != 1) {
        dctx_end(dctx);
        return rep_err(ZL_ErrorCode_userBuffers_invalidNum);
    }
    ZL_Report r = ZL_DCtx_decompressTBuffer(dctx, outputs[0], src, srcSize);
    return ZL_isError(r) ? r : rep_ok(1);
}

ZL_Type ZL_TypedBuffer_type(const ZL_TypedBuffer* t) { return t ? t->type : ZL_Type_serial; }
size_t ZL_TypedBuffer_byteSize(const ZL_TypedBuffer* t) { return t ? t->bytes : 0; }
const void* ZL_TypedBuffer_rPtr(const ZL_TypedBuffer* t)
{
    if (!t) return nullptr;
    return t->wrapType ? t->wrap : (const void*)t->own.data();
}
size_t ZL_TypedBuffer_numElts(const ZL_TypedBuffer* t) { return t && t->width ? t->bytes / t->width : 0; }
size_t ZL_TypedBuffer_eltWidth(const ZL_TypedBuffer* t) { return t ? t->width : 0; }

ZL_Report ZL_getDecompressedSize(const void* src, size_t srcSize)
{
    const uint8_t* f = (const uint8_t*)src;
    uint8_t head[16];
    if (src && srcSize && ozl_is_device_ptr(src)) {   // a frame in device memory: its first bytes
        srcSize = srcSize < sizeof(head) ? srcSize : sizeof(head);
        if (hipMemcpy(head, src, srcSize, hipMemcpyDeviceToHost) != hipSuccess) return rep_err(ZL_ErrorCode_GENERIC);
        f = head;
    }
    HostReader r = { f, 0, 0, srcSize, false };
    uint32_t magic = 0;
    for (int k = 0; k < 4; k++) magic |= r.u8() << (8 * k);
    if (r.bad || magic != OZL_MAGIC_BASE + OZL_FORMAT_VERSION) return rep_err(ZL_ErrorCode_header_unknown);
    r.u8();
    uint32_t tok = r.u8();
    if ((tok & 15) != 1) return rep_err(ZL_ErrorCode_frameParameter_unsupported);
    uint64_t s1 = r.varint();
    if (r.bad || s1 == 0) return rep_err(ZL_ErrorCode_corruption);
    return rep_ok(s1 - 1);
}

// decode_frameheader.c:375-470, 570-599 (frames of this library: one output)
ZL_FrameInfo* ZL_FrameInfo_create(const void* src, size_t srcSize)
{
    ParsedFrame P;
    if (host_frame_info(src, srcSize, P)) return nullptr;
    return new (std::nothrow) ZL_FrameInfo{ frame_ztype(P.type), P.size };
}
void ZL_FrameInfo_free(ZL_FrameInfo* fi) { delete fi; }
ZL_Report ZL_FrameInfo_getFormatVersion(const ZL_FrameInfo* fi)
{
    return fi ? rep_ok(OZL_FORMAT_VERSION) : rep_err(ZL_ErrorCode_GENERIC);
}
ZL_Report ZL_FrameInfo_getNumOutputs(const ZL_FrameInfo* fi) { return fi ? rep_ok(1) : rep_err(ZL_ErrorCode_GENERIC); }
ZL_Report ZL_FrameInfo_getOutputType(const ZL_FrameInfo* fi, int id)
{
    if (!fi) return rep_err(ZL_ErrorCode_GENERIC);
    return id < 1 ? rep_ok((size_t)fi->type) : rep_err(ZL_ErrorCode_outputID_invalid);
}
ZL_Report ZL_FrameInfo_getDecompressedSize(const ZL_FrameInfo* fi, int id)
{
    if (!fi) return rep_err(ZL_ErrorCode_GENERIC);
    return id < 1 ? rep_ok(fi->size) : rep_err(ZL_ErrorCode_outputID_invalid);
}
// decode_frameheader.c:439-470: serial outputs only
ZL_Report ZL_FrameInfo_getNumElts(const ZL_FrameInfo* fi, int id)
{
    if (!fi) return rep_err(ZL_ErrorCode_GENERIC);
    if (id >= 1) return rep_err(ZL_ErrorCode_outputID_invalid);
    if (fi->type != ZL_Type_serial) return rep_err(ZL_ErrorCode_temporaryLibraryLimitation);
    return rep_ok(fi->size);
}
ZL_Report ZL_getNumOutputs(const void* src, size_t srcSize)
{
    ZL_FrameInfo* fi = ZL_FrameInfo_create(src, srcSize);
    ZL_Report r = ZL_FrameInfo_getNumOutputs(fi);
    ZL_FrameInfo_free(fi);
    return r;
}
ZL_Report ZL_getOutputType(ZL_Type* stPtr, const void* src, size_t srcSize)
{
    if (!stPtr) return rep_err(ZL_ErrorCode_parameter_invalid);
    ParsedFrame P;
    int rc = host_frame_info(src, srcSize, P);
    if (rc) return rep_err(rc);
    *stPtr = frame_ztype(P.type);
    return rep_ok(0);
}

}  // extern "C"
