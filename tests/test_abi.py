"""The C ABI: libozl_hip.so loads without a GPU and exports every function
that include/openzl_amd.h declares; libozl_graphs.so (user graphs in C against
that header) links against it. No compute call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "openzl_amd.h")
LIB = os.path.join(ROOT, "openzl_amd", "libozl_hip.so")
GRAPHS = os.path.join(ROOT, "openzl_amd", "libozl_graphs.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set()
    for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
        line = text[m.start():text.find("\n", m.start())]
        if line.startswith(("static", "typedef", "#")) or "(*" in line:
            continue
        names.add(m.group(1))
    return sorted(names)


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("ZL_Compressor_create", "ZL_CCtx_compressTypedRef", "ZL_DCtx_decompressTyped",
                 "ZL_Compressor_registerFunctionGraph", "ozl_entropy_compress_batch", "ozl_bitpack_decode"):
        assert must in names
    assert len(names) >= 50


# The public functions a replacement of the reference library must export for
# the hot path and its callers: the drop-in list of the boundary (compressor,
# contexts, typed refs, universal decompressor, frame queries) plus what the
# reference C++ wrapper's contexts call (cpp/src/openzl/cpp/CCtx.cpp:29-67,
# DCtx.cpp:13-44, FrameInfo.cpp).
DROP_IN = [
    "ZL_Compressor_create", "ZL_Compressor_free", "ZL_Compressor_setParameter",
    "ZL_Compressor_registerStaticGraph_fromNode", "ZL_Compressor_registerStaticGraph_fromNode1o",
    "ZL_Compressor_registerFunctionGraph", "ZL_Compressor_selectStartingGraphID",
    "ZL_Compressor_registerTransposeSplitGraph", "ZL_Compressor_registerBruteForceSelectorGraph",
    "ZL_Compressor_registerSegmenter",
    "ZL_CCtx_create", "ZL_CCtx_free", "ZL_CCtx_setParameter", "ZL_CCtx_getParameter", "ZL_CCtx_resetParameters",
    "ZL_CCtx_refCompressor", "ZL_CCtx_compressTypedRef", "ZL_CCtx_compressMultiTypedRef", "ZL_CCtx_compress",
    "ZL_TypedRef_createNumeric", "ZL_TypedRef_createStruct", "ZL_TypedRef_createSerial", "ZL_TypedRef_free",
    "ZL_DCtx_create", "ZL_DCtx_free", "ZL_DCtx_setParameter", "ZL_DCtx_getParameter", "ZL_DCtx_resetParameters",
    "ZL_DCtx_decompress", "ZL_DCtx_decompressTyped", "ZL_DCtx_decompressTBuffer", "ZL_DCtx_decompressMultiTBuffer",
    "ZL_TypedBuffer_create", "ZL_TypedBuffer_free", "ZL_TypedBuffer_createWrapSerial",
    "ZL_TypedBuffer_createWrapStruct", "ZL_TypedBuffer_createWrapNumeric", "ZL_TypedBuffer_type",
    "ZL_TypedBuffer_byteSize", "ZL_TypedBuffer_rPtr", "ZL_TypedBuffer_numElts", "ZL_TypedBuffer_eltWidth",
    "ZL_FrameInfo_create", "ZL_FrameInfo_free", "ZL_FrameInfo_getFormatVersion", "ZL_FrameInfo_getNumOutputs",
    "ZL_FrameInfo_getOutputType", "ZL_FrameInfo_getDecompressedSize", "ZL_FrameInfo_getNumElts",
    "ZL_decompress", "ZL_getDecompressedSize", "ZL_getNumOutputs", "ZL_getOutputType", "ZL_compressBound",
    "ZL_ErrorCode_toString", "ZL_Edge_runNode", "ZL_Edge_setDestination", "ZL_Edge_getData",
    "ZL_CCtx_selectStartingGraphID", "ZL_Compressor_getParameter", "ZL_CCtx_getErrorContextString",
    "ZL_CCtx_getErrorContextString_fromError", "ZL_DCtx_getErrorContextString",
    "ZL_DCtx_getErrorContextString_fromError", "ZL_Compressor_getErrorContextString_fromError",
    "ZL_Compressor_registerSelectorGraph", "ZL_Selector_getOpaquePtr", "ZL_Input_ptr",
]


def test_header_declares_the_drop_in_list():
    names = set(declared_functions())
    missing = [n for n in DROP_IN if n not in names]
    assert not missing, missing


def test_descriptor_layouts_match_the_reference():
    """Offsets of the descriptor fields a caller fills in: the reference's
    declarations (include/openzl/zl_graph_api.h:74-95, zl_segmenter.h:70-108,
    zl_selector.h:174-193) compiled with the same compiler."""
    import subprocess
    import tempfile
    src = r'''
#include <stddef.h>
#include <stdio.h>
#include "openzl_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(ZL_SegmenterDesc), offsetof(ZL_SegmenterDesc, lastInputIsVariable),
         offsetof(ZL_SegmenterDesc, customGraphs), offsetof(ZL_SegmenterDesc, opaque), sizeof(bool));
  printf("%zu %zu %zu %zu %zu %zu %zu\n", sizeof(ZL_SelectorDesc), offsetof(ZL_SelectorDesc, inStreamType),
         offsetof(ZL_SelectorDesc, customGraphs), offsetof(ZL_SelectorDesc, nbCustomGraphs),
         offsetof(ZL_SelectorDesc, localParams), offsetof(ZL_SelectorDesc, name), offsetof(ZL_SelectorDesc, opaque));
  return 0; }
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-std=gnu11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        lines = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    size, lv, cg, op, bsz = map(int, lines[0].split())
    # name, segmenterFn, inputTypeMasks, numInputs (8 bytes each), then the
    # bool, padded to the next pointer
    assert bsz == 1 and lv == 32 and cg == 40 and op == 40 + 8 + 8 + 48 and size == op + 24
    # ZL_SelectorDesc (zl_selector.h:174-193, as the reference's header
    # compiles with gcc): selector_f, the ZL_Type enum padded to 8, the graph
    # list, ZL_LocalParams (3 x pointer + size), name, ZL_OpaquePtr
    assert list(map(int, lines[1].split())) == [112, 8, 16, 24, 32, 80, 88]


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(GRAPHS), reason="graphs library not built")
def test_c_user_graphs_link_against_the_library():
    ctypes.CDLL(LIB, mode=ctypes.RTLD_GLOBAL)
    g = ctypes.CDLL(GRAPHS)
    assert hasattr(g, "ozl_graph_weights_split")


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_context_parameters_follow_the_reference():
    """ZL_CCtx / ZL_DCtx parameter handling without any compression call (no
    GPU work): every ZL_CParam is accepted on both the compressor and the
    context (gcparams.c:62-100), unknown ones are compressionParameter_invalid,
    format versions outside 8..23 are formatVersion_unsupported, a reset
    clears every parameter (cctx.c:293-314), and decompression parameters
    behave the same way (gdparams.c:15-33)."""
    import sys
    sys.path.insert(0, ROOT)
    from openzl_amd import api as A
    L = A._api()
    comp = L.ZL_Compressor_create()
    cctx = L.ZL_CCtx_create()
    try:
        for p in (A.CP_STICKY, A.CP_LEVEL, A.CP_DLEVEL, A.CP_FORMAT, A.CP_PERMISSIVE, A.CP_KSUM, A.CP_CSUM,
                  A.CP_MINSTREAM):
            v = 23 if p == A.CP_FORMAT else 1
            assert L.ZL_Compressor_setParameter(comp, p, v).code == 0, p
            assert L.ZL_CCtx_setParameter(cctx, p, v).code == 0, p
            assert L.ZL_CCtx_getParameter(cctx, p) == v
        assert L.ZL_CCtx_setParameter(cctx, 99, 1).code == 20
        assert L.ZL_Compressor_setParameter(comp, 99, 1).code == 20
        assert L.ZL_CCtx_setParameter(cctx, A.CP_FORMAT, 7).code == 60
        assert L.ZL_CCtx_setParameter(cctx, A.CP_FORMAT, 24).code == 60
        assert L.ZL_CCtx_setParameter(cctx, A.CP_MINSTREAM, -1).code == 0
        assert L.ZL_CCtx_getParameter(cctx, A.CP_MINSTREAM) == -1
        assert L.ZL_CCtx_setParameter(cctx, A.CP_STICKY, 5).code == 0
        assert L.ZL_CCtx_getParameter(cctx, A.CP_STICKY) == 1
        assert L.ZL_CCtx_resetParameters(cctx).code == 0
        assert all(L.ZL_CCtx_getParameter(cctx, p) == 0 for p in (1, 2, 3, 4, 5, 6, 7, 11))
        # no compressor referenced: compressionParameter_invalid (compress2.c:349)
        src = (ctypes.c_uint8 * 16)()
        dst = (ctypes.c_uint8 * 1024)()
        tr = L.ZL_TypedRef_createSerial(src, 16)
        assert L.ZL_CCtx_compressTypedRef(cctx, dst, 1024, tr).code == 20
        L.ZL_TypedRef_free(tr)
    finally:
        L.ZL_CCtx_free(cctx)
        L.ZL_Compressor_free(comp)
    d = L.ZL_DCtx_create()
    try:
        for p in (1, 2, 3):
            assert L.ZL_DCtx_setParameter(d, p, 2).code == 0
        assert L.ZL_DCtx_getParameter(d, 1) == 1 and L.ZL_DCtx_getParameter(d, 3) == 2
        assert L.ZL_DCtx_setParameter(d, 4, 1).code == 20
        assert L.ZL_DCtx_resetParameters(d).code == 0
        assert all(L.ZL_DCtx_getParameter(d, p) == 0 for p in (1, 2, 3))
    finally:
        L.ZL_DCtx_free(d)


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_frame_queries_on_golden_frames():
    """ZL_FrameInfo_* / ZL_getOutputType / ZL_getNumOutputs read host frame
    headers only (decode_frameheader.c:375-470, 570-599)."""
    import sys
    sys.path.insert(0, ROOT)
    from openzl_amd import api as A
    import golden_cases as GC
    L = A._api()
    types = {0: 1, 1: 2, 2: 4}
    for e in GC.small_entries():
        fr = GC.small_frame(e["name"])
        buf = ctypes.create_string_buffer(fr, len(fr))
        fi = L.ZL_FrameInfo_create(buf, len(fr))
        assert fi, e["name"]
        try:
            assert L.ZL_FrameInfo_getFormatVersion(fi).value == 23
            assert L.ZL_FrameInfo_getNumOutputs(fi).value == 1
            assert L.ZL_FrameInfo_getOutputType(fi, 0).value == types[e["type"]]
            assert L.ZL_FrameInfo_getOutputType(fi, 1).code == 22
            assert L.ZL_FrameInfo_getDecompressedSize(fi, 0).value == e["input_bytes"]
            ne = L.ZL_FrameInfo_getNumElts(fi, 0)
            if e["type"] == 0:
                assert ne.code == 0 and ne.value == e["input_bytes"]
            else:
                assert ne.code == 81
        finally:
            L.ZL_FrameInfo_free(fi)
        t = ctypes.c_int(0)
        assert L.ZL_getOutputType(ctypes.byref(t), buf, len(fr)).code == 0 and t.value == types[e["type"]]
        assert L.ZL_getNumOutputs(buf, len(fr)).value == 1
    bad = ctypes.create_string_buffer(b"\x00" * 16, 16)
    assert not L.ZL_FrameInfo_create(bad, 16)
    assert L.ZL_getNumOutputs(bad, 16).code != 0
