"""Python view of the ZL_* C API of libozl_hip.so (the same functions a C
program calls). Builds the graphs of the benchmark configurations exactly as a
reference user would (include/openzl/zl_compressor.h), compresses host
buffers and decompresses frames; every codec runs on the GPU."""
import ctypes
import os

import numpy as np

from ._lib import lib

P = ctypes.c_void_p
SZ = ctypes.c_size_t


class ZLReport(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("value", ctypes.c_size_t)]


class GraphID(ctypes.Structure):
    _fields_ = [("gid", ctypes.c_uint)]


class NodeID(ctypes.Structure):
    _fields_ = [("nid", ctypes.c_uint)]


class OpaquePtr(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("freeOpaquePtr", ctypes.c_void_p), ("freeFn", ctypes.c_void_p)]


class SelectorDesc(ctypes.Structure):
    """ZL_SelectorDesc (include/openzl/zl_selector.h:174-193)."""
    _fields_ = [("selector_f", ctypes.c_void_p), ("inStreamType", ctypes.c_int),
                ("customGraphs", ctypes.c_void_p), ("nbCustomGraphs", ctypes.c_size_t),
                ("localParams", ctypes.c_char * 48), ("name", ctypes.c_char_p), ("opaque", OpaquePtr)]


class OutputInfo(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("fixedWidth", ctypes.c_uint32), ("decompressedByteSize", ctypes.c_uint64),
                ("numElts", ctypes.c_uint64)]


SERIAL, STRUCT, NUMERIC = 1, 2, 4
ZL_GRAPH_INVALID, ZL_ALLOCATION = 30, 70
LEARNED_MODEL = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector_mlp.bin")
# ZL_CParam values (include/openzl/zl_compress.h:64-131)
CP_STICKY, CP_LEVEL, CP_DLEVEL, CP_FORMAT, CP_PERMISSIVE, CP_KSUM, CP_CSUM, CP_MINSTREAM = 1, 2, 3, 4, 5, 6, 7, 11
G_STORE, G_FSE, G_HUFFMAN, G_BITPACK = GraphID(2), GraphID(3), GraphID(4), GraphID(8)
N_DELTA, N_TRANSPOSE_SPLIT, N_ZIGZAG = NodeID(2), NodeID(3), NodeID(4)

# graph kinds, same numbering as the reference harness (oracle/ref_harness.c)
STORE, FSE, HUFFMAN, BITPACK, DELTA_TS_FSE, DELTA_ZZ_BP, TS_HUF_FSE, DELTA_FSE, TS_FSE, TS_HUF, ZZ_BP, DELTA_BP, SELECT8, \
    SEGMENTED, LEARNED8 = range(15)
# the learned selector's candidates, in the model's output order
SELECT_CANDIDATES = (STORE, FSE, BITPACK, DELTA_FSE, DELTA_ZZ_BP, TS_FSE, TS_HUF, DELTA_TS_FSE)

_setup = False


def _api():
    global _setup
    L = lib()
    if _setup:
        return L
    sig = {
        "ZL_Compressor_create": (P, []),
        "ZL_Compressor_free": (None, [P]),
        "ZL_Compressor_setParameter": (ZLReport, [P, ctypes.c_int, ctypes.c_int]),
        "ZL_Compressor_registerStaticGraph_fromNode1o": (GraphID, [P, NodeID, GraphID]),
        "ZL_Compressor_registerTransposeSplitGraph": (GraphID, [P, GraphID]),
        "ZL_Compressor_registerBruteForceSelectorGraph": (GraphID, [P, ctypes.POINTER(GraphID), SZ]),
        "ZL_Compressor_selectStartingGraphID": (ZLReport, [P, GraphID]),
        "ZL_CCtx_create": (P, []),
        "ZL_CCtx_free": (None, [P]),
        "ZL_CCtx_refCompressor": (ZLReport, [P, P]),
        "ZL_CCtx_setParameter": (ZLReport, [P, ctypes.c_int, ctypes.c_int]),
        "ZL_CCtx_getParameter": (ctypes.c_int, [P, ctypes.c_int]),
        "ZL_CCtx_resetParameters": (ZLReport, [P]),
        "ZL_CCtx_selectStartingGraphID": (ZLReport, [P, P, GraphID, P]),
        "ZL_Compressor_registerSelectorGraph": (GraphID, [P, P]),
        "ozl_learned_model_create": (P, [P, SZ]),
        "ZL_CCtx_getErrorContextString": (ctypes.c_char_p, [P, ZLReport]),
        "ZL_DCtx_getErrorContextString": (ctypes.c_char_p, [P, ZLReport]),
        "ZL_Compressor_getParameter": (ctypes.c_int, [P, ctypes.c_int]),
        "ZL_CCtx_compressMultiTypedRef": (ZLReport, [P, P, SZ, ctypes.POINTER(P), SZ]),
        "ZL_DCtx_setParameter": (ZLReport, [P, ctypes.c_int, ctypes.c_int]),
        "ZL_DCtx_getParameter": (ctypes.c_int, [P, ctypes.c_int]),
        "ZL_DCtx_resetParameters": (ZLReport, [P]),
        "ZL_DCtx_decompress": (ZLReport, [P, P, SZ, P, SZ]),
        "ZL_decompress": (ZLReport, [P, SZ, P, SZ]),
        "ZL_TypedBuffer_create": (P, []),
        "ZL_TypedBuffer_free": (None, [P]),
        "ZL_TypedBuffer_createWrapSerial": (P, [P, SZ]),
        "ZL_TypedBuffer_createWrapStruct": (P, [P, SZ, SZ]),
        "ZL_TypedBuffer_createWrapNumeric": (P, [P, SZ, SZ]),
        "ZL_DCtx_decompressTBuffer": (ZLReport, [P, P, P, SZ]),
        "ZL_DCtx_decompressMultiTBuffer": (ZLReport, [P, ctypes.POINTER(P), SZ, P, SZ]),
        "ZL_TypedBuffer_type": (ctypes.c_int, [P]),
        "ZL_TypedBuffer_byteSize": (SZ, [P]),
        "ZL_TypedBuffer_rPtr": (P, [P]),
        "ZL_TypedBuffer_numElts": (SZ, [P]),
        "ZL_TypedBuffer_eltWidth": (SZ, [P]),
        "ZL_getNumOutputs": (ZLReport, [P, SZ]),
        "ZL_getOutputType": (ZLReport, [ctypes.POINTER(ctypes.c_int), P, SZ]),
        "ZL_FrameInfo_create": (P, [P, SZ]),
        "ZL_FrameInfo_free": (None, [P]),
        "ZL_FrameInfo_getFormatVersion": (ZLReport, [P]),
        "ZL_FrameInfo_getNumOutputs": (ZLReport, [P]),
        "ZL_FrameInfo_getOutputType": (ZLReport, [P, ctypes.c_int]),
        "ZL_FrameInfo_getDecompressedSize": (ZLReport, [P, ctypes.c_int]),
        "ZL_FrameInfo_getNumElts": (ZLReport, [P, ctypes.c_int]),
        "ZL_TypedRef_createSerial": (P, [P, SZ]),
        "ZL_TypedRef_createStruct": (P, [P, SZ, SZ]),
        "ZL_TypedRef_createNumeric": (P, [P, SZ, SZ]),
        "ZL_TypedRef_free": (None, [P]),
        "ZL_CCtx_compressTypedRef": (ZLReport, [P, P, SZ, P]),
        "ZL_compressBound": (SZ, [SZ]),
        "ZL_DCtx_create": (P, []),
        "ZL_DCtx_free": (None, [P]),
        "ZL_DCtx_decompressTyped": (ZLReport, [P, ctypes.POINTER(OutputInfo), P, SZ, P, SZ]),
        "ZL_ErrorCode_toString": (ctypes.c_char_p, [ctypes.c_int]),
        "ZL_getDecompressedSize": (ZLReport, [P, SZ]),
        "ozl_CCtx_compressBatch": (ZLReport, [P, P, SZ, ctypes.POINTER(SZ), ctypes.POINTER(P), SZ]),
        "ozl_DCtx_decompressBatch": (ZLReport, [P, P, SZ, ctypes.POINTER(SZ), P, ctypes.POINTER(SZ), SZ]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _setup = True
    return L


class ZLError(RuntimeError):
    def __init__(self, code, what):
        L = _api()
        super().__init__(f"{what}: {L.ZL_ErrorCode_toString(code).decode()} ({code})")
        self.code = code


def _chk(r, what):
    if r.code:
        raise ZLError(r.code, what)
    return r.value


_graphs = None


def _graphs_lib():
    """libozl_graphs.so: the user graphs written in C against the public API
    (openzl_amd/graphs/ozl_graphs.c)."""
    global _graphs
    if _graphs is None:
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libozl_graphs.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built: run openzl_amd/build.py")
        _api()
        g = ctypes.CDLL(path)
        g.ozl_graph_weights_split.restype = GraphID
        g.ozl_graph_weights_split.argtypes = [P]
        g.ozl_graph_fixed_chunks.restype = GraphID
        g.ozl_graph_fixed_chunks.argtypes = [P, GraphID, SZ]
        _graphs = g
    return _graphs


def build_graph(L, c, kind, segment=None):
    """GraphID of a graph kind; SEGMENTED takes segment = (inner kind, chunk elements)."""
    if kind == SEGMENTED:
        inner, elts = segment
        return _graphs_lib().ozl_graph_fixed_chunks(c, build_graph(L, c, inner), elts)
    if kind == STORE:
        return G_STORE
    if kind == FSE:
        return G_FSE
    if kind == HUFFMAN:
        return G_HUFFMAN
    if kind == BITPACK:
        return G_BITPACK
    if kind == DELTA_TS_FSE:
        ts = L.ZL_Compressor_registerTransposeSplitGraph(c, G_FSE)
        return L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_DELTA, ts)
    if kind == DELTA_ZZ_BP:
        zz = L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_ZIGZAG, G_BITPACK)
        return L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_DELTA, zz)
    if kind == TS_HUF_FSE:
        # transpose_split, plane 0 -> FSE, plane 1 -> Huffman (a C function graph)
        return _graphs_lib().ozl_graph_weights_split(c)
    if kind == DELTA_FSE:
        return L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_DELTA, G_FSE)
    if kind == TS_FSE:
        return L.ZL_Compressor_registerTransposeSplitGraph(c, G_FSE)
    if kind == TS_HUF:
        return L.ZL_Compressor_registerTransposeSplitGraph(c, G_HUFFMAN)
    if kind == ZZ_BP:
        return L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_ZIGZAG, G_BITPACK)
    if kind == DELTA_BP:
        return L.ZL_Compressor_registerStaticGraph_fromNode1o(c, N_DELTA, G_BITPACK)
    raise ValueError(kind)


class Compressor:
    """A ZL_Compressor + ZL_CCtx pair for one graph kind (format version 23)."""

    PARAMS = {"minStreamSize": CP_MINSTREAM, "compressionLevel": CP_LEVEL, "contentChecksum": CP_CSUM,
              "compressedChecksum": CP_KSUM}

    def __init__(self, kind, candidates=None, content_checksum=True, compressed_checksum=True, segment=None,
                 params=None, model=None):
        """params: further compressor parameters by name (PARAMS); model: the
        learned selector's parameter bytes for LEARNED8 (default: the
        packaged model)."""
        L = _api()
        self.L = L
        self.c = L.ZL_Compressor_create()
        _chk(L.ZL_Compressor_setParameter(self.c, CP_FORMAT, 23), "formatVersion")
        for k, v in (params or {}).items():
            _chk(L.ZL_Compressor_setParameter(self.c, self.PARAMS[k], int(v)), k)
        if not content_checksum:
            _chk(L.ZL_Compressor_setParameter(self.c, 7, 2), "contentChecksum")
        if not compressed_checksum:
            _chk(L.ZL_Compressor_setParameter(self.c, 6, 2), "compressedChecksum")
        if kind == LEARNED8:
            # the learned selector over the 8 candidates, the model owned by
            # the compressor (freed with it)
            self._gids = (GraphID * len(SELECT_CANDIDATES))(*[build_graph(L, self.c, k) for k in SELECT_CANDIDATES])
            raw = np.fromfile(LEARNED_MODEL, dtype=np.uint8) if model is None else np.ascontiguousarray(model)
            m = L.ozl_learned_model_create(raw.ctypes.data, raw.nbytes)
            if not m:
                raise ZLError(ZL_ALLOCATION, "ozl_learned_model_create")
            d = SelectorDesc()
            d.selector_f = ctypes.cast(L.ozl_learned_selector, ctypes.c_void_p).value
            d.inStreamType = SERIAL | NUMERIC
            d.customGraphs = ctypes.cast(self._gids, ctypes.c_void_p).value
            d.nbCustomGraphs = len(SELECT_CANDIDATES)
            d.opaque = OpaquePtr(m, None, ctypes.cast(L.ozl_learned_model_free, ctypes.c_void_p).value)
            g = L.ZL_Compressor_registerSelectorGraph(self.c, ctypes.byref(d))
            if g.gid == 0:
                raise ZLError(ZL_GRAPH_INVALID, "ZL_Compressor_registerSelectorGraph")
        elif candidates is not None:
            gids = (GraphID * len(candidates))(*[build_graph(L, self.c, k) for k in candidates])
            g = L.ZL_Compressor_registerBruteForceSelectorGraph(self.c, gids, len(candidates))
        else:
            g = build_graph(L, self.c, kind, segment)
        _chk(L.ZL_Compressor_selectStartingGraphID(self.c, g), "selectStartingGraphID")
        self.cctx = L.ZL_CCtx_create()
        # a context drops its parameters and its compressor after every call
        # unless they are sticky (as the reference's does)
        _chk(L.ZL_CCtx_setParameter(self.cctx, CP_STICKY, 1), "stickyParameters")
        _chk(L.ZL_CCtx_refCompressor(self.cctx, self.c), "refCompressor")

    def compress(self, arr, typ, width):
        L = self.L
        arr = np.ascontiguousarray(arr)
        nb = arr.nbytes
        n = nb // width
        if typ == SERIAL:
            tr = L.ZL_TypedRef_createSerial(arr.ctypes.data, nb)
        elif typ == STRUCT:
            tr = L.ZL_TypedRef_createStruct(arr.ctypes.data, width, n)
        else:
            tr = L.ZL_TypedRef_createNumeric(arr.ctypes.data, width, n)
        cap = L.ZL_compressBound(nb)
        out = np.empty(cap, np.uint8)
        try:
            size = _chk(L.ZL_CCtx_compressTypedRef(self.cctx, out.ctypes.data, cap, tr), "compress")
        finally:
            L.ZL_TypedRef_free(tr)
        return out[:size].tobytes()

    def _typed_ref(self, ptr, nbytes, typ, width):
        L = self.L
        if typ == SERIAL:
            return L.ZL_TypedRef_createSerial(ptr, nbytes)
        if typ == STRUCT:
            return L.ZL_TypedRef_createStruct(ptr, width, nbytes // width)
        return L.ZL_TypedRef_createNumeric(ptr, width, nbytes // width)

    def compress_ptrs(self, ptrs, nbytes, typ, width, dst, cap):
        """ozl_CCtx_compressBatch over inputs at ptrs (host or device
        addresses, nbytes[i] bytes each) into dst (host or device address,
        cap bytes); returns the frame sizes."""
        L = self.L
        n = len(ptrs)
        refs = [self._typed_ref(p, b, typ, width) for p, b in zip(ptrs, nbytes)]
        sizes = (SZ * max(n, 1))()
        try:
            _chk(L.ozl_CCtx_compressBatch(self.cctx, dst, cap, sizes, (P * max(n, 1))(*refs), n), "compressBatch")
        finally:
            for r in refs:
                L.ZL_TypedRef_free(r)
        return [sizes[i] for i in range(n)]

    def compress_batch(self, arrs, typ, width):
        """One frame per host array, in one ozl_CCtx_compressBatch call."""
        arrs = [np.ascontiguousarray(a) for a in arrs]
        cap = sum(self.L.ZL_compressBound(a.nbytes) for a in arrs) + 64
        out = np.empty(cap, np.uint8)
        sizes = self.compress_ptrs([a.ctypes.data for a in arrs], [a.nbytes for a in arrs], typ, width,
                                   out.ctypes.data, cap)
        frames, pos = [], 0
        for z in sizes:
            frames.append(out[pos:pos + z].tobytes())
            pos += z
        return frames

    def __del__(self):
        try:
            self.L.ZL_CCtx_free(self.cctx)
            self.L.ZL_Compressor_free(self.c)
        except Exception:
            pass


class Decompressor:
    def __init__(self):
        L = _api()
        self.L = L
        self.d = L.ZL_DCtx_create()

    def decompress(self, frame, capacity):
        """Returns (bytes, type, width)."""
        src = np.frombuffer(frame, np.uint8)
        out = np.empty(max(capacity, 1), np.uint8)
        info = OutputInfo()
        size = _chk(self.L.ZL_DCtx_decompressTyped(self.d, ctypes.byref(info), out.ctypes.data, out.size,
                                                   src.ctypes.data, src.size), "decompress")
        return out[:size].tobytes(), info.type, info.fixedWidth

    def decompress_ptrs(self, src, frame_sizes, dst, cap):
        """ozl_DCtx_decompressBatch of the frames back to back at src (host or
        device address) into dst (host or device address, cap bytes); returns
        the content sizes."""
        n = len(frame_sizes)
        fs = (SZ * max(n, 1))(*frame_sizes)
        outs = (SZ * max(n, 1))()
        _chk(self.L.ozl_DCtx_decompressBatch(self.d, dst, cap, outs, src, fs, n), "decompressBatch")
        return [outs[i] for i in range(n)]

    def decompress_batch(self, frames, capacity):
        """Contents of host frames, decoded in one ozl_DCtx_decompressBatch call."""
        src = np.frombuffer(b"".join(frames) or b"\0", np.uint8)
        out = np.empty(max(capacity, 1), np.uint8)
        sizes = self.decompress_ptrs(src.ctypes.data, [len(f) for f in frames], out.ctypes.data, out.size)
        res, pos = [], 0
        for z in sizes:
            res.append(out[pos:pos + z].tobytes())
            pos += z
        return res

    def __del__(self):
        try:
            self.L.ZL_DCtx_free(self.d)
        except Exception:
            pass
