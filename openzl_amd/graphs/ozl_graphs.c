/* ozl_graphs.c -- user graphs written against the public ZL_* API only.
 *
 * These are the graphs a reference user writes in C (include/openzl/
 * zl_compressor.h, zl_graph_api.h); the same source compiles against the
 * reference headers. They are built into libozl_graphs.so, which links
 * libozl_hip.so, so every codec they select runs on the GPU.
 */
#include <stddef.h>

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
