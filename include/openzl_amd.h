/*
 * openzl_amd.h -- C ABI of libozl_hip.so, the MI355X (gfx950) implementation
 * of OpenZL's data-parallel codec hot path.
 *
 * Two layers:
 *
 * 1. The reference's own public API, same names, same argument meaning, same
 *    numeric node / graph / transform IDs and the same ZL_Report layout, for
 *    the graphs of the hot path (ZL_NODE_DELTA_INT, ZL_NODE_ZIGZAG,
 *    ZL_NODE_TRANSPOSE_SPLIT, ZL_GRAPH_BITPACK, ZL_GRAPH_FSE, ZL_GRAPH_HUFFMAN,
 *    ZL_GRAPH_STORE, ZL_GRAPH_CONSTANT, ZL_GRAPH_FLATPACK, function graphs and
 *    the brute-force selector). A program written against the include/openzl/ headers
 *    for these graphs links against this library instead; the codec work runs
 *    on the GPU and frames are byte-identical to the reference's. Each
 *    declaration cites the reference declaration it replaces.
 *
 * 2. Device-resident batch entry points (ozl_*): device pointers in and out,
 *    many independent blocks per call. They have no counterpart in the
 *    reference (whose contexts are single-threaded, src/openzl/compress/
 *    cctx.c:1160-1168); a binding reaches them the same way as layer 1.
 *
 * All functions are thread-compatible: one ZL_CCtx / ZL_DCtx per thread.
 */
#ifndef OPENZL_AMD_H
#define OPENZL_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ errors */
/* include/openzl/zl_errors_types.h:41-99 (subset, same values) */
typedef enum {
    ZL_ErrorCode_no_error = 0,
    ZL_ErrorCode_GENERIC = 1,
    ZL_ErrorCode_srcSize_tooSmall = 3,
    ZL_ErrorCode_srcSize_tooLarge = 4,
    ZL_ErrorCode_dstCapacity_tooSmall = 5,
    ZL_ErrorCode_userBuffers_invalidNum = 8,
    ZL_ErrorCode_header_unknown = 10,
    ZL_ErrorCode_frameParameter_unsupported = 11,
    ZL_ErrorCode_corruption = 12,
    ZL_ErrorCode_compressedChecksumWrong = 13,
    ZL_ErrorCode_contentChecksumWrong = 14,
    ZL_ErrorCode_outputs_tooNumerous = 15,
    ZL_ErrorCode_compressionParameter_invalid = 20,
    ZL_ErrorCode_parameter_invalid = 21,
    ZL_ErrorCode_outputID_invalid = 22,
    ZL_ErrorCode_segmenter_inputNotConsumed = 26,
    ZL_ErrorCode_graph_invalid = 30,
    ZL_ErrorCode_graph_invalidNumInputs = 33,
    ZL_ErrorCode_successor_invalid = 40,
    ZL_ErrorCode_inputType_unsupported = 43,
    ZL_ErrorCode_node_invalid_input = 55,
    ZL_ErrorCode_node_invalid = 56,
    ZL_ErrorCode_formatVersion_unsupported = 60,
    ZL_ErrorCode_formatVersion_notSet = 61,
    ZL_ErrorCode_allocation = 70,
    ZL_ErrorCode_internalBuffer_tooSmall = 71,
    ZL_ErrorCode_logicError = 80,
    ZL_ErrorCode_temporaryLibraryLimitation = 81,
} ZL_ErrorCode;

/* include/openzl/zl_errors.h:206-257: ZL_Report is a 16-byte union whose
 * first member is the error code; on success the value follows. */
typedef union {
    ZL_ErrorCode _code;
    struct {
        ZL_ErrorCode _code;
        size_t _value;
    } _value;
    struct {
        ZL_ErrorCode _code;
        void* _info;
    } _error;
} ZL_Report;

static inline int ZL_isError(ZL_Report r) { return r._code != ZL_ErrorCode_no_error; }
static inline size_t ZL_validResult(ZL_Report r) { return r._value._value; }
static inline ZL_ErrorCode ZL_errorCode(ZL_Report r) { return r._code; }
/* include/openzl/zl_errors.h: ZL_ErrorCode_toString */
const char* ZL_ErrorCode_toString(ZL_ErrorCode code);
/* include/openzl/detail/zl_errors_detail.h:217-220: an error is its code and
 * an info pointer (16 bytes, the error member of ZL_Report) */
typedef struct {
    ZL_ErrorCode _code;
    void* _info;
} ZL_Error;

/* ------------------------------------------------------------ types and IDs */
/* include/openzl/zl_data.h: ZL_Type bit flags */
typedef enum { ZL_Type_serial = 1, ZL_Type_struct = 2, ZL_Type_numeric = 4, ZL_Type_string = 8 } ZL_Type;

typedef struct { unsigned nid; } ZL_NodeID;   /* include/openzl/zl_opaque_types.h */
typedef struct { unsigned gid; } ZL_GraphID;

/* include/openzl/zl_nodes.h:13-88 */
#define ZL_NODE_ILLEGAL ((ZL_NodeID){ 0 })
#define ZL_NODE_DELTA_INT ((ZL_NodeID){ 2 })
#define ZL_NODE_TRANSPOSE_SPLIT ((ZL_NodeID){ 3 })
#define ZL_NODE_ZIGZAG ((ZL_NodeID){ 4 })
#define ZL_NODE_CONVERT_STRUCT_TO_SERIAL ((ZL_NodeID){ 10 })
#define ZL_NODE_CONVERT_NUM_TO_STRUCT_LE ((ZL_NodeID){ 11 })
#define ZL_NODE_CONVERT_NUM_TO_SERIAL_LE ((ZL_NodeID){ 12 })

/* include/openzl/zl_graphs.h:11-42 */
#define ZL_GRAPH_ILLEGAL ((ZL_GraphID){ 0 })
#define ZL_GRAPH_STORE ((ZL_GraphID){ 2 })
#define ZL_GRAPH_FSE ((ZL_GraphID){ 3 })
#define ZL_GRAPH_HUFFMAN ((ZL_GraphID){ 4 })
#define ZL_GRAPH_CONSTANT ((ZL_GraphID){ 6 })
#define ZL_GRAPH_BITPACK ((ZL_GraphID){ 8 })
#define ZL_GRAPH_FLATPACK ((ZL_GraphID){ 9 })

/* include/openzl/zl_compress.h:64-131 */
typedef enum {
    ZL_CParam_stickyParameters = 1,
    ZL_CParam_compressionLevel = 2,
    ZL_CParam_decompressionLevel = 3,
    ZL_CParam_formatVersion = 4,
    ZL_CParam_permissiveCompression = 5,
    ZL_CParam_compressedChecksum = 6,
    ZL_CParam_contentChecksum = 7,
    ZL_CParam_minStreamSize = 11,
} ZL_CParam;
typedef enum { ZL_TernaryParam_auto = 0, ZL_TernaryParam_enable = 1, ZL_TernaryParam_disable = 2 } ZL_TernaryParam;
#define ZL_MIN_FORMAT_VERSION 8  /* include/openzl/zl_version.h:49 */
#define ZL_MAX_FORMAT_VERSION 23 /* include/openzl/zl_version.h:57 */
#define ZL_MINSTREAMSIZE_DEFAULT 10 /* include/openzl/zl_compress.h:137 */
/* include/openzl/zl_decompress.h:91-127 */
typedef enum {
    ZL_DParam_stickyParameters = 1,
    ZL_DParam_checkCompressedChecksum = 2,
    ZL_DParam_checkContentChecksum = 3,
} ZL_DParam;

typedef struct ZL_Compressor_s ZL_Compressor;
typedef struct ZL_CCtx_s ZL_CCtx;
typedef struct ZL_DCtx_s ZL_DCtx;
typedef struct ZL_TypedRef_s ZL_TypedRef;
typedef struct ZL_Graph_s ZL_Graph;
typedef struct ZL_Edge_s ZL_Edge;
typedef struct ZL_Input_s ZL_Input;

/* ------------------------------------------------------------ compressor */
/* include/openzl/zl_compressor.h:69 */
ZL_Compressor* ZL_Compressor_create(void);
void ZL_Compressor_free(ZL_Compressor* compressor);
/* include/openzl/zl_compressor.h:151 */
ZL_Report ZL_Compressor_setParameter(ZL_Compressor* compressor, ZL_CParam param, int value);
/* include/openzl/zl_compressor.h:232 */
ZL_GraphID ZL_Compressor_registerStaticGraph_fromNode1o(ZL_Compressor* compressor, ZL_NodeID headNode,
                                                        ZL_GraphID dstGraph);
/* include/openzl/zl_compressor.h:281 */
ZL_GraphID ZL_Compressor_registerStaticGraph_fromNode(ZL_Compressor* compressor, ZL_NodeID headNode,
                                                      const ZL_GraphID* dstGraphs, size_t nbDstGraphs);
/* include/openzl/codecs/zl_transpose.h:43 */
ZL_GraphID ZL_Compressor_registerTransposeSplitGraph(ZL_Compressor* compressor, ZL_GraphID successor);
/* src/openzl/compress/selectors/selector_brute_force.c:52-66 (declared in
 * include/openzl/zl_compressor.h) */
ZL_GraphID ZL_Compressor_registerBruteForceSelectorGraph(ZL_Compressor* compressor, const ZL_GraphID* successors,
                                                         size_t numSuccessors);
/* include/openzl/zl_compressor.h:555 */
ZL_Report ZL_Compressor_selectStartingGraphID(ZL_Compressor* compressor, ZL_GraphID graph);
/* include/openzl/zl_compressor.h:162-164 */
int ZL_Compressor_getParameter(const ZL_Compressor* compressor, ZL_CParam gcparam);
const char* ZL_Compressor_getErrorContextString_fromError(const ZL_Compressor* compressor, ZL_Error error);

/* Function graphs: include/openzl/zl_graph_api.h:58-120 (subset). */
typedef ZL_Report (*ZL_FunctionGraphFn)(ZL_Graph* graph, ZL_Edge* inputs[], size_t nbInputs);
typedef struct {
    const char* name;
    ZL_FunctionGraphFn graph_f;
    void* validate_f;                 /* unused */
    const ZL_Type* inputTypeMasks;
    size_t nbInputs;
    int lastInputIsVariable;
    const ZL_GraphID* customGraphs;
    size_t nbCustomGraphs;
    const ZL_NodeID* customNodes;
    size_t nbCustomNodes;
    char localParams[48];             /* layout placeholder, unused */
    char opaque[24];                  /* layout placeholder, unused */
} ZL_FunctionGraphDesc;
/* include/openzl/zl_graph_api.h:114 */
ZL_GraphID ZL_Compressor_registerFunctionGraph(ZL_Compressor* compressor, const ZL_FunctionGraphDesc* desc);

/* Segmenters: include/openzl/zl_segmenter.h:69-108, 127-330 (subset: one
 * input; processChunk takes no runtime parameters). A segmenter cuts its
 * input into chunks, each compressed by a graph into a chunk of the frame
 * (format version 21+). */
typedef struct {
    const ZL_GraphID* graphids;
    size_t nbGraphIDs;
} ZL_GraphIDList;                     /* include/openzl/zl_common_types.h:50-53 */
typedef struct {
    void* ptr;
    void* freeOpaquePtr;
    void (*freeFn)(void* freeOpaquePtr, void* ptr);
} ZL_OpaquePtr;                       /* include/openzl/zl_common_types.h:25-49 */
typedef struct ZL_Segmenter_s ZL_Segmenter;
typedef ZL_Report (*ZL_SegmenterFn)(ZL_Segmenter* sctx);
typedef struct {
    const char* name;
    ZL_SegmenterFn segmenterFn;
    const ZL_Type* inputTypeMasks;
    size_t numInputs;
    bool lastInputIsVariable;         /* 1 input only here */
    const ZL_GraphID* customGraphs;
    size_t numCustomGraphs;
    char localParams[48];             /* layout placeholder, unused */
    ZL_OpaquePtr opaque;
} ZL_SegmenterDesc;
/* include/openzl/zl_segmenter.h:97 */
ZL_GraphID ZL_Compressor_registerSegmenter(ZL_Compressor* compressor, const ZL_SegmenterDesc* segDesc);
/* include/openzl/zl_segmenter.h:127 */
const void* ZL_Segmenter_getOpaquePtr(const ZL_Segmenter* segCtx);
/* include/openzl/zl_segmenter.h:140 */
int ZL_Segmenter_getCParam(const ZL_Segmenter* segCtx, ZL_CParam gparam);
/* include/openzl/zl_segmenter.h:187 */
ZL_GraphIDList ZL_Segmenter_getCustomGraphs(const ZL_Segmenter* segCtx);
/* include/openzl/zl_segmenter.h:200 */
size_t ZL_Segmenter_numInputs(const ZL_Segmenter* segCtx);
/* include/openzl/zl_segmenter.h:222 (the remaining input; sizes and type only:
 * its bytes are in device memory) */
const ZL_Input* ZL_Segmenter_getInput(const ZL_Segmenter* segCtx, size_t inputID);
/* include/openzl/zl_segmenter.h:253 */
ZL_Report ZL_Segmenter_getNumElts(const ZL_Segmenter* segCtx, size_t numElts[], size_t numInputs);
/* include/openzl/zl_segmenter.h:277 (host memory, freed when the call ends) */
void* ZL_Segmenter_getScratchSpace(ZL_Segmenter* segCtx, size_t size);
/* include/openzl/zl_segmenter.h:300-330; rGraphParams must be NULL */
ZL_Report ZL_Segmenter_processChunk(ZL_Segmenter* segCtx, const size_t numElts[], size_t numInputs,
                                    ZL_GraphID startingGraphID, const void* rGraphParams);
/* Selector graphs: include/openzl/zl_selector.h:156-222. The selector
 * function runs on the host during the graph walk and returns the successor
 * of its whole input (any graph of the compressor accepting the input's type);
 * customGraphs are handed back to it. The input's bytes are in device memory
 * (ZL_Input_ptr). localParams are not supported (must be zero). */
typedef struct ZL_Selector_s ZL_Selector;
typedef ZL_GraphID (*ZL_SelectorFn)(const ZL_Selector* selectorAPI, const ZL_Input* input,
                                    const ZL_GraphID* customGraphs, size_t nbCustomGraphs);
typedef struct {
    ZL_SelectorFn selector_f;
    ZL_Type inStreamType;
    const ZL_GraphID* customGraphs;
    size_t nbCustomGraphs;
    char localParams[48];             /* layout of ZL_LocalParams; must be zero */
    const char* name;
    ZL_OpaquePtr opaque;
} ZL_SelectorDesc;
/* include/openzl/zl_selector.h:222 (the compressor owns desc->opaque from
 * here on, even when registration fails) */
ZL_GraphID ZL_Compressor_registerSelectorGraph(ZL_Compressor* compressor, const ZL_SelectorDesc* desc);
/* include/openzl/zl_selector.h:235 */
const void* ZL_Selector_getOpaquePtr(const ZL_Selector* selector);
/* include/openzl/zl_input.h:107: the first byte of the input (device memory here) */
const void* ZL_Input_ptr(const ZL_Input* input);

/* The learned selector as a selector function (the counterpart of the
 * reference's ML selector, src/openzl/compress/selectors/ml/mlselector.c:
 * ZL_Compressor_registerMLSelectorGraph with a model of its own): the 40
 * features and the bf16 MLP of ozl_select_features / ozl_select_mlp run on
 * the device over the input, and the argmax picks customGraphs[choice]. The
 * selector's opaque pointer is a model from ozl_learned_model_create, and
 * customGraphs are the 8 candidates in the order of ozl_select_compress
 * (STORE, FSE, BITPACK, DELTA_INT -> FSE, DELTA_INT -> ZIGZAG -> BITPACK,
 * transpose_split -> FSE, transpose_split -> HUFFMAN, DELTA_INT ->
 * transpose_split -> FSE). With no model or fewer candidates than the choice
 * it returns customGraphs[0]. No counterpart in the reference. */
ZL_GraphID ozl_learned_selector(const ZL_Selector* selectorAPI, const ZL_Input* input, const ZL_GraphID* customGraphs,
                                size_t nbCustomGraphs);
/* A device copy of the MLP parameters (ozl_select_mlp_params_bytes() bytes
 * at hostParams), or NULL; ozl_learned_model_free is its ZL_OpaquePtr free
 * function. No counterpart in the reference. */
void* ozl_learned_model_create(const void* hostParams, size_t bytes);
void ozl_learned_model_free(void* freeOpaquePtr, void* model);

typedef struct {
    ZL_Edge** edges;
    size_t nbEdges;
} ZL_EdgeList;
typedef union {
    ZL_ErrorCode _code;
    struct {
        ZL_ErrorCode _code;
        ZL_EdgeList _value;
    } _value;
} ZL_Result_ZL_EdgeList;
/* include/openzl/zl_graph_api.h:254 */
ZL_Result_ZL_EdgeList ZL_Edge_runNode(ZL_Edge* input, ZL_NodeID node);
/* include/openzl/codecs/zl_transpose.h:56 */
ZL_Result_ZL_EdgeList ZL_Edge_runTransposeSplit(ZL_Edge* edge, const ZL_Graph* graph);
/* include/openzl/zl_graph_api.h:283 */
ZL_Report ZL_Edge_setDestination(ZL_Edge* edge, ZL_GraphID destination);
/* include/openzl/zl_graph_api.h:156 */
const ZL_Input* ZL_Edge_getData(const ZL_Edge* edge);
/* include/openzl/zl_input.h:54-99 (inline there, exported functions here) */
size_t ZL_Input_numElts(const ZL_Input* input);
size_t ZL_Input_eltWidth(const ZL_Input* input);
size_t ZL_Input_contentSize(const ZL_Input* input);
ZL_Type ZL_Input_type(const ZL_Input* input);

/* ------------------------------------------------------------ compression */
/* include/openzl/zl_compress.h:50 */
size_t ZL_compressBound(size_t totalSrcSize);
/* include/openzl/zl_compress.h:60-61 */
ZL_CCtx* ZL_CCtx_create(void);
void ZL_CCtx_free(ZL_CCtx* cctx);
/* include/openzl/zl_compressor.h:597 */
ZL_Report ZL_CCtx_refCompressor(ZL_CCtx* cctx, const ZL_Compressor* compressor);
/* include/openzl/zl_compress.h:152. Every ZL_CParam is accepted (others:
 * compressionParameter_invalid; formatVersion must be 0 or 8..23). A CCtx
 * value that is not 0 overrides the compressor's, which overrides the
 * default (src/openzl/compress/cctx.c:350-364, gcparams.c:124-137). Unless
 * ZL_CParam_stickyParameters is 1, every compression ends with
 * ZL_CCtx_resetParameters (compress2.c:141-145), which also drops the
 * compressor reference. minStreamSize 0 means 10; a negative value turns the
 * automatic store off. */
ZL_Report ZL_CCtx_setParameter(ZL_CCtx* cctx, ZL_CParam param, int value);
/* include/openzl/zl_compress.h:161 */
int ZL_CCtx_getParameter(const ZL_CCtx* cctx, ZL_CParam param);
/* include/openzl/zl_compress.h:168 */
ZL_Report ZL_CCtx_resetParameters(ZL_CCtx* cctx);
/* include/openzl/zl_compressor.h:411-422; only NULL is accepted by
 * ZL_CCtx_selectStartingGraphID here (runtime graph parameters are outside
 * the hot path) */
typedef struct ZL_GraphParameters_s ZL_GraphParameters;
/* include/openzl/zl_compressor.h:616-620 (cctx.c:269-279): the starting graph
 * of the next compression, reset with the other requested parameters unless
 * sticky; a non-NULL compressor becomes the referenced one; rgp != NULL is
 * temporaryLibraryLimitation. */
ZL_Report ZL_CCtx_selectStartingGraphID(ZL_CCtx* cctx, const ZL_Compressor* compressor, ZL_GraphID graphID,
                                        const ZL_GraphParameters* rgp);
/* include/openzl/zl_compress.h:213-223, zl_decompress.h:188-204,
 * zl_compressor.h:115-117: the message of an error (here the error code's
 * name: no dynamic context is recorded). */
const char* ZL_CCtx_getErrorContextString(const ZL_CCtx* cctx, ZL_Report report);
const char* ZL_CCtx_getErrorContextString_fromError(const ZL_CCtx* cctx, ZL_Error error);
/* include/openzl/zl_compress.h:313-349 */
ZL_TypedRef* ZL_TypedRef_createSerial(const void* src, size_t srcSize);
ZL_TypedRef* ZL_TypedRef_createStruct(const void* start, size_t structWidth, size_t structCount);
ZL_TypedRef* ZL_TypedRef_createNumeric(const void* start, size_t numWidth, size_t numCount);
void ZL_TypedRef_free(ZL_TypedRef* tref);
/* include/openzl/zl_compress.h:274 */
ZL_Report ZL_CCtx_compressTypedRef(ZL_CCtx* cctx, void* dst, size_t dstCapacity, const ZL_TypedRef* input);
/* include/openzl/zl_compress.h:295. One input writes the frame
 * ZL_CCtx_compressTypedRef writes; frames of several inputs are outside the
 * hot path and are rejected with temporaryLibraryLimitation. */
ZL_Report ZL_CCtx_compressMultiTypedRef(ZL_CCtx* cctx, void* dst, size_t dstCapacity, const ZL_TypedRef* inputs[],
                                        size_t nbInputs);
/* include/openzl/zl_compress.h:194. Without a compressor the reference runs
 * its generic LZ graph (compress2.c:371-380), which is not part of the hot
 * path: that call returns graph_invalid here. */
ZL_Report ZL_CCtx_compress(ZL_CCtx* cctx, void* dst, size_t dstCapacity, const void* src, size_t srcSize);
/* nbInputs calls of ZL_CCtx_compressTypedRef (include/openzl/zl_compress.h:274)
 * in one: frame i (the same bytes as that call writes for inputs[i]) lands in
 * dst after frames 0..i-1, frameSizes[i] = its size; returns the total. The
 * entropy stages of all frames run as one launch per coder. Inputs and dst
 * may be host or device memory (device inputs are read in place); the
 * single-frame calls accept device memory the same way. No counterpart in the
 * reference: its callers loop over ZL_CCtx_compressTypedRef. */
ZL_Report ozl_CCtx_compressBatch(ZL_CCtx* cctx, void* dst, size_t dstCapacity, size_t* frameSizes,
                                 const ZL_TypedRef* const* inputs, size_t nbInputs);

/* ------------------------------------------------------------ decompression */
/* include/openzl/zl_decompress.h:355-366 */
typedef struct {
    ZL_Type type;
    uint32_t fixedWidth;
    uint64_t decompressedByteSize;
    uint64_t numElts;
} ZL_OutputInfo;
/* include/openzl/zl_decompress.h:32 */
ZL_Report ZL_decompress(void* dst, size_t dstCapacity, const void* src, size_t srcSize);
/* include/openzl/zl_decompress.h:47 */
ZL_Report ZL_getDecompressedSize(const void* compressed, size_t cSize);
/* include/openzl/zl_decompress.h:80, 87 */
ZL_DCtx* ZL_DCtx_create(void);
void ZL_DCtx_free(ZL_DCtx* dctx);
/* include/openzl/zl_decompress.h:140, 149, 159. Parameters reset at the end
 * of every decompression unless ZL_DParam_stickyParameters is 1
 * (decompress2.c); a checksum is verified unless its parameter is
 * ZL_TernaryParam_disable. Unknown parameters: parameter_invalid. */
ZL_Report ZL_DCtx_setParameter(ZL_DCtx* dctx, ZL_DParam param, int value);
int ZL_DCtx_getParameter(const ZL_DCtx* dctx, ZL_DParam param);
ZL_Report ZL_DCtx_resetParameters(ZL_DCtx* dctx);
const char* ZL_DCtx_getErrorContextString(const ZL_DCtx* dctx, ZL_Report report);
const char* ZL_DCtx_getErrorContextString_fromError(const ZL_DCtx* dctx, ZL_Error error);
/* include/openzl/zl_decompress.h:230 */
ZL_Report ZL_DCtx_decompress(ZL_DCtx* dctx, void* dst, size_t dstCapacity, const void* compressed, size_t cSize);
/* include/openzl/zl_decompress.h:388 */
ZL_Report ZL_DCtx_decompressTyped(ZL_DCtx* dctx, ZL_OutputInfo* outputInfo, void* dst, size_t dstByteCapacity,
                                  const void* compressed, size_t cSize);

/* Typed buffers: include/openzl/zl_decompress.h:396-613 (String type
 * excluded: string outputs are not part of the hot path). An empty buffer
 * (ZL_TypedBuffer_create) owns host memory sized by the frame; a wrapped one
 * writes into the caller's buffer (host or device memory). */
typedef struct ZL_TypedBuffer_s ZL_TypedBuffer;
ZL_TypedBuffer* ZL_TypedBuffer_create(void);
void ZL_TypedBuffer_free(ZL_TypedBuffer* tbuffer);
ZL_TypedBuffer* ZL_TypedBuffer_createWrapSerial(void* buffer, size_t bufferCapacity);
ZL_TypedBuffer* ZL_TypedBuffer_createWrapStruct(void* structBuffer, size_t structWidth, size_t numStructs);
ZL_TypedBuffer* ZL_TypedBuffer_createWrapNumeric(void* numArray, size_t numWidth, size_t numCapacity);
/* include/openzl/zl_decompress.h:505 */
ZL_Report ZL_DCtx_decompressTBuffer(ZL_DCtx* dctx, ZL_TypedBuffer* output, const void* compressed, size_t cSize);
/* include/openzl/zl_decompress.h:527 (frames of this library have one
 * output: nbOutputs must be 1, else userBuffers_invalidNum) */
ZL_Report ZL_DCtx_decompressMultiTBuffer(ZL_DCtx* dctx, ZL_TypedBuffer* outputs[], size_t nbOutputs,
                                         const void* compressed, size_t cSize);
/* include/openzl/zl_decompress.h:550-600 */
ZL_Type ZL_TypedBuffer_type(const ZL_TypedBuffer* tbuffer);
size_t ZL_TypedBuffer_byteSize(const ZL_TypedBuffer* tbuffer);
const void* ZL_TypedBuffer_rPtr(const ZL_TypedBuffer* tbuffer);
size_t ZL_TypedBuffer_numElts(const ZL_TypedBuffer* tbuffer);
size_t ZL_TypedBuffer_eltWidth(const ZL_TypedBuffer* tbuffer);

/* Frame queries: include/openzl/zl_decompress.h:245-342 (host memory). */
ZL_Report ZL_getNumOutputs(const void* compressed, size_t cSize);
ZL_Report ZL_getOutputType(ZL_Type* stPtr, const void* compressed, size_t cSize);
typedef struct ZL_FrameInfo ZL_FrameInfo;
ZL_FrameInfo* ZL_FrameInfo_create(const void* compressed, size_t cSize);
void ZL_FrameInfo_free(ZL_FrameInfo* fi);
ZL_Report ZL_FrameInfo_getFormatVersion(const ZL_FrameInfo* fi);
ZL_Report ZL_FrameInfo_getNumOutputs(const ZL_FrameInfo* fi);
ZL_Report ZL_FrameInfo_getOutputType(const ZL_FrameInfo* fi, int outputID);
ZL_Report ZL_FrameInfo_getDecompressedSize(const ZL_FrameInfo* fi, int outputID);
ZL_Report ZL_FrameInfo_getNumElts(const ZL_FrameInfo* fi, int outputID);
/* nbFrames calls of ZL_DCtx_decompressTyped (include/openzl/zl_decompress.h:388)
 * in one: frame i is src[sum of frameSizes[0..i), + frameSizes[i]); its
 * content lands in dst after those of frames 0..i-1, outSizes[i] = its size;
 * returns the total. The entropy stages of all frames run as shared launches.
 * src and dst may be host or device memory. No counterpart in the reference:
 * its callers loop over ZL_DCtx_decompressTyped. */
ZL_Report ozl_DCtx_decompressBatch(ZL_DCtx* dctx, void* dst, size_t dstCapacity, size_t* outSizes, const void* src,
                                   const size_t* frameSizes, size_t nbFrames);

/* ------------------------------------------------------------ device batch entry points */
/* Numeric transforms on device buffers (no reference counterpart: the
 * reference kernels ZS_deltaEncode etc. take host pointers,
 * src/openzl/codecs/delta/encode_delta_kernel.c:80). `stream` is a hipStream_t
 * (0 = default). Return 0 or a ZL_ErrorCode. */
int ozl_delta_encode(const void* src, size_t n, int width, void* deltas, void* first, void* stream);
size_t ozl_delta_decode_scratch_size(size_t n, int width);
int ozl_delta_decode(const void* first, const void* deltas, size_t n, int width, void* dst, void* scratch,
                     void* stream);
int ozl_zigzag_encode(const void* src, size_t n, int width, void* dst, void* stream);
int ozl_zigzag_decode(const void* src, size_t n, int width, void* dst, void* stream);
int ozl_transpose_split(const void* src, size_t n, int width, void* planes, void* stream);
int ozl_transpose_join(const void* planes, size_t n, int width, void* dst, void* stream);
int ozl_max_async(const void* src, size_t n, int width, unsigned long long* dmax, void* stream);
int ozl_bitpack_encode(const void* src, size_t n, int width, int nbBits, void* dst, void* stream);
int ozl_bitpack_decode(const void* src, size_t n, int width, int nbBits, void* dst, void* stream);

/* Batched ZL_GRAPH_FSE (hufMode 0) / ZL_GRAPH_HUFFMAN (hufMode 1) frames, one
 * per block of < 100000 bytes: the same frame as ZL_CCtx_compress
 * (include/openzl/zl_compress.h:194) of each block with a compressor whose
 * starting graph is ZL_GRAPH_FSE / ZL_GRAPH_HUFFMAN (src/openzl/codecs/
 * entropy/encode_entropy_binding.c:814-851), format version 23, both
 * checksums on. All pointers are device pointers; offsets has nbBlocks+1
 * entries; frame b lands at dst + frameOffsets[b]; status[b] is 0 or the
 * block's error code. */
size_t ozl_entropy_compress_workspace(uint32_t nbBlocks, size_t maxBlock);
size_t ozl_entropy_frame_bound(size_t blockSize);
int ozl_entropy_compress_batch(int hufMode, const void* src, const uint64_t* offsets, uint32_t nbBlocks,
                               size_t maxBlock, void* dst, size_t dstCapacity, uint64_t* frameOffsets,
                               int* status, void* workspace, size_t workspaceSize, void* stream);
/* ZL_decompress (include/openzl/zl_decompress.h:32) of many such frames in
 * one call; frames of any other layout are reported as corruption in status
 * and are left to ZL_DCtx_decompressTyped. */
size_t ozl_entropy_decompress_workspace(uint32_t nbBlocks);
int ozl_entropy_decompress_batch(const void* frames, const uint64_t* frameOffsets, uint32_t nbFrames, void* dst,
                                 const uint64_t* dstOffsets, int* status, void* workspace, size_t workspaceSize,
                                 void* stream);
/* No counterpart in the reference: ozl_entropy_decompress_batch for callers
 * that cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc). The workspace is the library's own device array of
 * 9 MiB per GPU (held until the process ends), so at most
 * ozl_entropy_decompress_noalloc_max_frames() = 16384 frames per call; more
 * is internalBuffer_tooSmall. Calls that overlap on one GPU share that array
 * and race. */
int ozl_entropy_decompress_batch_noalloc(const void* frames, const uint64_t* frameOffsets, uint32_t nbFrames,
                                         void* dst, const uint64_t* dstOffsets, int* status, void* stream);
size_t ozl_entropy_decompress_noalloc_max_frames(void);

/* Batched numeric columns, one frame per column, with the graph
 * ZL_NODE_DELTA_INT -> ZL_NODE_ZIGZAG -> ZL_GRAPH_BITPACK: the frame
 * ZL_CCtx_compressTypedRef (include/openzl/zl_compress.h:274) writes for a
 * numeric ZL_TypedRef of each column with that graph
 * (ZL_Compressor_registerStaticGraph_fromNode1o, zl_compressor.h), format
 * version 23, both checksums on. Column c is
 * src[colSpans[2c], colSpans[2c] + colSpans[2c+1]) of element width
 * widths[c] (1, 2, 4 or 8; start aligned to the width). `tile` is the number
 * of deltas per work item (multiple of 32, at most 4096), wgPerCu the
 * persistent workgroups per CU (0 = 4). Device pointers; no host
 * synchronisation and no memset, so a call can be captured in a hipGraph. */
size_t ozl_columns_workspace(uint32_t nbColumns, size_t totalBytes, uint32_t tile);
size_t ozl_columns_frame_bound(size_t columnBytes);
int ozl_columns_compress_batch(const void* src, const uint64_t* colSpans, const uint8_t* widths, uint32_t nbColumns,
                               void* dst, size_t dstCapacity, uint64_t* frameOffsets, int* status, void* workspace,
                               size_t workspaceSize, uint32_t tile, uint32_t wgPerCu, void* stream);
/* ZL_decompress (include/openzl/zl_decompress.h:32) of such frames into
 * dst[dstSpans[2c], +dstSpans[2c+1]); status[c] 0 or the frame's error code
 * (corruption, checksum, a frame of another layout). */
int ozl_columns_decompress_batch(const void* frames, const uint64_t* frameOffsets, uint32_t nbColumns, void* dst,
                                 const uint64_t* dstSpans, int* status, void* workspace, size_t workspaceSize,
                                 uint32_t tile, uint32_t wgPerCu, void* stream);
/* Times compress calls over tiles {1024, 2048, 4096} x workgroups per CU
 * {2, 4, 8} on the caller's batch (HIP events) and returns the fastest; the
 * result is cached per (nbColumns, totalBytes). No reference counterpart. */
int ozl_columns_autotune(const void* src, const uint64_t* colSpans, const uint8_t* widths, uint32_t nbColumns,
                         uint64_t totalBytes, void* dst, size_t dstCapacity, uint64_t* frameOffsets, int* status,
                         void* workspace, size_t workspaceSize, uint32_t* tileOut, uint32_t* wgPerCuOut,
                         void* stream);
/* A hipGraph holding one compress (dir 0: a = src, aOff = colSpans, b = dst,
 * bOff = frameOffsets) or decompress (dir 1: a = frames, aOff =
 * frameOffsets, b = dst, bOff = dstSpans) call with fixed arguments;
 * ozl_graph_launch replays it on `stream`. No reference counterpart. */
void* ozl_columns_graph_create(int dir, const void* a, const uint64_t* aOff, const uint8_t* widths,
                               uint32_t nbColumns, void* b, size_t bCap, uint64_t* bOff, int* status, void* workspace,
                               size_t workspaceSize, uint32_t tile, uint32_t wgPerCu);
int ozl_graph_launch(void* graph, void* stream);
void ozl_graph_destroy(void* graph);

/* Learned graph selection over eight candidate graphs for numeric chunks
 * (< 100000 bytes each): 0 ZL_GRAPH_STORE, 1 ZL_GRAPH_FSE, 2 ZL_GRAPH_BITPACK,
 * 3 DELTA_INT -> FSE, 4 DELTA_INT -> ZIGZAG -> BITPACK, 5 transpose_split ->
 * FSE, 6 transpose_split -> HUFFMAN, 7 DELTA_INT -> transpose_split -> FSE.
 * ozl_select_trial is the brute-force selector over them
 * (src/openzl/compress/selectors/selector_brute_force.c:9-50,
 * ZL_Compressor_registerBruteForceSelectorGraph :52-66): each candidate's
 * frame size with checksums off (sizes, K x 8) and the selected candidate
 * (choice). ozl_select_compress writes each chunk's frame with candidate
 * choice[c] (format version 23, checksums on): the reference's frame for that
 * graph, and the brute-force selector's frame when choice comes from the
 * trial. ozl_select_features / ozl_select_mlp are the learned selector (the
 * counterpart of the GBT selector, src/openzl/compress/selectors/ml/
 * mlselector.c): 40 features per chunk and a 40-64-64-8 MLP in bf16 whose
 * argmax is the choice. Device pointers; trial and compress synchronise the
 * host once per call. */
size_t ozl_select_workspace(uint32_t nbChunks, size_t maxChunk);
size_t ozl_select_frame_bound(size_t chunkBytes);
int ozl_select_trial(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t nbChunks,
                     size_t maxChunk, uint32_t* sizes, uint8_t* choice, void* workspace, size_t workspaceSize,
                     void* stream);
int ozl_select_compress(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t nbChunks,
                        size_t maxChunk, const uint8_t* choice, void* dst, size_t dstCapacity,
                        uint64_t* frameOffsets, int* status, void* workspace, size_t workspaceSize, void* stream);
size_t ozl_select_mlp_params_bytes(void);
int ozl_select_features(const void* src, const uint64_t* spans, const uint8_t* widths, uint32_t nbChunks,
                        float* features, void* stream);
int ozl_select_mlp(const float* features, uint32_t nbChunks, const void* params, uint8_t* choice, float* logits,
                   void* stream);

/* Per-kernel timing (HIP events around every launch while enabled; no
 * reference counterpart). ozl_profile_end writes "<tag><kernel> <launches>
 * <total ms>" lines; ozl_profile_tag sets the tag of later launches. */
void ozl_profile_begin(void);
void ozl_profile_tag(const char* tag);
size_t ozl_profile_end(char* out, size_t capacity);

#ifdef __cplusplus
}
#endif
#endif
