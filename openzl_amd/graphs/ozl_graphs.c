/* ozl_graphs.c -- user graphs written against the public ZL_* API only.
 *
 * These are the graphs a reference user writes in C (include/openzl/
 * zl_compressor.h, zl_graph_api.h); the same source compiles against the
 * reference headers. They are built into libozl_graphs.so, which links
 * libozl_hip.so, so every codec they select runs on the GPU.
 */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "openzl_amd.h"

/* Model-weight graph: transpose_split of 2-byte structs into byte planes;
 * plane 0 (low byte, mantissa bits of bf16) -> ZL_GRAPH_FSE, plane 1 (sign
 * and exponent) -> ZL_GRAPH_HUFFMAN. Same shape as a function graph built
 * with ZL_Edge_runTransposeSplit + ZL_Edge_setDestination
 * (include/openzl/zl_graph_api.h). */
static ZL_Report weights_split_fn(ZL_Graph* graph, ZL_Edge* inputs[], size_t nbInputs)
{
    ZL_Result_ZL_EdgeList r;
    ZL_Report rep;
    if (nbInputs != 1) {
        rep._code = ZL_ErrorCode_graph_invalid;
        return rep;
    }
    r = ZL_Edge_runTransposeSplit(inputs[0], graph);
    if (r._code != ZL_ErrorCode_no_error) {
        rep._code = r._code;
        return rep;
    }
    if (r._value._value.nbEdges != 2) {
        rep._code = ZL_ErrorCode_node_invalid_input;
        return rep;
    }
    rep = ZL_Edge_setDestination(r._value._value.edges[0], ZL_GRAPH_FSE);
    if (rep._code != ZL_ErrorCode_no_error) return rep;
    return ZL_Edge_setDestination(r._value._value.edges[1], ZL_GRAPH_HUFFMAN);
}

static const ZL_Type kStructMask[1] = { ZL_Type_struct };

ZL_GraphID ozl_graph_weights_split(ZL_Compressor* compressor)
{
    ZL_FunctionGraphDesc d = { 0 };
    d.name = "weights_split_fse_huffman";
    d.graph_f = weights_split_fn;
    d.inputTypeMasks = kStructMask;
    d.nbInputs = 1;
    return ZL_Compressor_registerFunctionGraph(compressor, &d);
}

/* Fixed-size segmenter (include/openzl/zl_segmenter.h): the input is cut
 * into chunks of chunkElts elements (the last one shorter), each compressed
 * by `inner` into its own chunk of the frame. The chunk size travels in the
 * segmenter's opaque pointer, freed with the compressor. */
static ZL_Report fixed_chunks_fn(ZL_Segmenter* sctx)
{
    const size_t per = *(const size_t*)ZL_Segmenter_getOpaquePtr(sctx);
    ZL_GraphIDList gl = ZL_Segmenter_getCustomGraphs(sctx);
    ZL_Report rep;
    size_t left = 0;
    if (gl.nbGraphIDs != 1 || per == 0) {
        rep._code = ZL_ErrorCode_graph_invalid;
        return rep;
    }
    rep = ZL_Segmenter_getNumElts(sctx, &left, 1);
    if (rep._code != ZL_ErrorCode_no_error) return rep;
    while (left > 0) {
        size_t n = left < per ? left : per;
        rep = ZL_Segmenter_processChunk(sctx, &n, 1, gl.graphids[0], NULL);
        if (rep._code != ZL_ErrorCode_no_error) return rep;
        left -= n;
    }
    rep._code = ZL_ErrorCode_no_error;
    rep._value._value = 0;
    return rep;
}

static void free_chunk_size(void* unused, void* p)
{
    (void)unused;
    free(p);
}

static const ZL_Type kAnyMask[1] = { (ZL_Type)(ZL_Type_serial | ZL_Type_struct | ZL_Type_numeric) };

ZL_GraphID ozl_graph_fixed_chunks(ZL_Compressor* compressor, ZL_GraphID inner, size_t chunkElts)
{
    ZL_SegmenterDesc d;
    size_t* per = (size_t*)malloc(sizeof(size_t));
    ZL_GraphID bad = { 0 };
    if (!per) return bad;
    *per = chunkElts;
    memset(&d, 0, sizeof(d));
    d.name = "fixed_chunks";
    d.segmenterFn = fixed_chunks_fn;
    d.inputTypeMasks = kAnyMask;
    d.numInputs = 1;
    d.customGraphs = &inner;
    d.numCustomGraphs = 1;
    d.opaque.ptr = per;
    d.opaque.freeOpaquePtr = NULL;
    d.opaque.freeFn = free_chunk_size;
    return ZL_Compressor_registerSegmenter(compressor, &d);
}
