"""ctypes access to the two checkers the tests compare against.

* ``oracle()``  -- oracle/liboracle.so, the CPU restatement (always built).
* ``ref()``     -- oracle/_ref/libopenzl_ref.so, the reference library compiled
                   from its own sources; ``refh()`` is the flat frame-level
                   harness around it. Both are None when they were not built
                   (no reference checkout), and tests that need them skip.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libopenzl_ref.so")
REFH_SO = os.path.join(ROOT, "oracle", "_ref", "libref_harness.so")

P = ctypes.c_void_p
SZ = ctypes.c_size_t
U = ctypes.c_uint
I = ctypes.c_int

_cache = {}


def _load(path, setup):
    if path in _cache:
        return _cache[path]
    lib = None
    if os.path.exists(path):
        lib = ctypes.CDLL(path)
        setup(lib)
    _cache[path] = lib
    return lib


def _setup_oracle(L):
    L.orc_varint_write.restype = SZ
    L.orc_varint_write.argtypes = [ctypes.c_uint64, P]
    L.orc_next_pow2.restype = I
    L.orc_next_pow2.argtypes = [ctypes.c_uint64]
    for f in ("orc_delta_encode", "orc_delta_decode"):
        getattr(L, f).argtypes = [P, P, P, SZ, I]
    for f in ("orc_zigzag_encode", "orc_zigzag_decode"):
        getattr(L, f).argtypes = [P, P, SZ, I]
    for f in ("orc_transpose_split", "orc_transpose_join"):
        getattr(L, f).argtypes = [P, P, SZ, I]
    L.orc_bitpack_nbbits.restype = I
    L.orc_bitpack_nbbits.argtypes = [P, SZ, I]
    L.orc_bitpack_encode.restype = SZ
    L.orc_bitpack_encode.argtypes = [P, P, SZ, I, I]
    L.orc_bitpack_header.restype = SZ
    L.orc_bitpack_header.argtypes = [P, SZ, I, I, SZ]
    L.orc_bitpack_decode.restype = SZ
    L.orc_bitpack_decode.argtypes = [P, SZ, P, SZ, P, SZ, P]
    L.orc_flatpack_encode.restype = SZ
    L.orc_flatpack_encode.argtypes = [P, P, P, P, SZ]
    L.orc_flatpack_decode.restype = SZ
    L.orc_flatpack_decode.argtypes = [P, SZ, P, SZ, P, SZ]
    L.orc_histogram.argtypes = [P, P, P, P, SZ]
    L.orc_entropy.restype = ctypes.c_double
    L.orc_entropy.argtypes = [P, SZ, SZ]
    L.orc_entropy_decide.restype = I
    L.orc_entropy_decide.argtypes = [P, SZ, I]
    L.orc_fse_optimal_tablelog.restype = U
    L.orc_fse_optimal_tablelog.argtypes = [U, SZ, U, U]
    L.orc_fse_normalize.restype = I
    L.orc_fse_normalize.argtypes = [P, U, P, SZ, U, I]
    L.orc_fse_build_ctable.restype = I
    L.orc_fse_build_ctable.argtypes = [P, P, U, U]
    L.orc_fse_encode.restype = SZ
    L.orc_fse_encode.argtypes = [P, SZ, P, SZ, P, U]
    L.orc_fse_write_ncount.restype = SZ
    L.orc_fse_write_ncount.argtypes = [P, SZ, P, U, U]
    L.orc_fse_read_ncount.restype = ctypes.c_long
    L.orc_fse_read_ncount.argtypes = [P, P, P, P, SZ]
    L.orc_fse_decode.restype = I
    L.orc_fse_decode.argtypes = [P, SZ, P, SZ, P, U, U, U]
    L.orc_huf_optimal_tablelog.restype = U
    L.orc_huf_optimal_tablelog.argtypes = [U, SZ, U]
    L.orc_histogram16.argtypes = [P, P, P, P, SZ]
    L.orc_entropy2_decide.restype = I
    L.orc_entropy2_decide.argtypes = [P, SZ, P]
    L.orc_tokenize2.restype = SZ
    L.orc_tokenize2.argtypes = [P, P, P, SZ]
    L.orc_huf16_build.restype = I
    L.orc_huf16_build.argtypes = [P, P, P, U]
    L.orc_huf16_encode.restype = SZ
    L.orc_huf16_encode.argtypes = [P, SZ, P, SZ, P, P, I]
    L.orc_huf16_decode.restype = I
    L.orc_huf16_decode.argtypes = [P, SZ, P, SZ, P, SZ, I]
    L.orc_huf_build_ctable.restype = I
    L.orc_huf_build_ctable.argtypes = [P, P, P, U, U]
    for f in ("orc_huf_encode_1x", "orc_huf_encode_4x"):
        getattr(L, f).restype = SZ
        getattr(L, f).argtypes = [P, SZ, P, SZ, P, P]
    L.orc_huf_decode.restype = I
    L.orc_huf_decode.argtypes = [P, SZ, P, SZ, P, SZ, I]
    L.orc_xxh3_64.restype = ctypes.c_uint64
    L.orc_xxh3_64.argtypes = [P, SZ]


def _setup_ref(L):
    L.ZL_calculateEntropy.restype = ctypes.c_double
    L.ZL_calculateEntropy.argtypes = [P, SZ, SZ]
    L.ZS_FSE_optimalTableLog.restype = U
    L.ZS_FSE_optimalTableLog.argtypes = [U, SZ, U]
    L.ZS_FSE_normalizeCount.restype = SZ
    L.ZS_FSE_normalizeCount.argtypes = [P, U, P, SZ, U, U]
    L.ZS_FSE_buildCTable.restype = SZ
    L.ZS_FSE_buildCTable.argtypes = [P, P, U, U]
    L.ZS_FSE_compress_usingCTable.restype = SZ
    L.ZS_FSE_compress_usingCTable.argtypes = [P, SZ, P, SZ, P, U]
    L.ZS_FSE_writeNCount.restype = SZ
    L.ZS_FSE_writeNCount.argtypes = [P, SZ, P, U, U]
    L.ZS_FSE_readNCount.restype = SZ
    L.ZS_FSE_readNCount.argtypes = [P, P, P, P, SZ]
    L.ZS_FSE_isError.restype = U
    L.ZS_FSE_isError.argtypes = [SZ]
    L.ZS_HUF_optimalTableLog.restype = U
    L.ZS_HUF_optimalTableLog.argtypes = [U, SZ, U]
    L.ZS_HUF_buildCTable.restype = SZ
    L.ZS_HUF_buildCTable.argtypes = [P, P, U, U]
    for f in ("ZS_HUF_compress1X_usingCTable", "ZS_HUF_compress4X_usingCTable"):
        getattr(L, f).restype = SZ
        getattr(L, f).argtypes = [P, SZ, P, SZ, P]
    L.ZS_HUF_isError.restype = U
    L.ZS_HUF_isError.argtypes = [SZ]
    L.ZS_deltaEncode.argtypes = [P, P, P, SZ, SZ]
    L.ZS_deltaDecode.argtypes = [P, P, P, SZ, SZ]
    L.ZS_splitTransposeEncode.argtypes = [P, P, SZ, SZ]
    L.ZS_splitTransposeDecode.argtypes = [P, P, SZ, SZ]
    L.ZS_bitpackEncode.restype = SZ
    L.ZS_bitpackEncode.argtypes = [P, SZ, P, SZ, SZ, I]
    L.ZS_bitpackDecode.restype = SZ
    L.ZS_bitpackDecode.argtypes = [P, SZ, SZ, P, SZ, I]
    L.ZS_flatpackEncode.restype = SZ
    L.ZS_flatpackEncode.argtypes = [P, SZ, P, SZ, P, SZ]
    for w in (8, 16, 32, 64):
        getattr(L, f"ZL_zigzagEncode{w}").argtypes = [P, P, SZ]
        getattr(L, f"ZL_zigzagDecode{w}").argtypes = [P, P, SZ]


def _setup_refh(L):
    L.refh_huf16_build.restype = I
    L.refh_huf16_build.argtypes = [P, P, P, U]
    L.refh_huf16_encode.restype = SZ
    L.refh_huf16_encode.argtypes = [P, SZ, P, SZ, P, P, U, I, I]
    L.refh_huf16_decode.restype = ctypes.c_longlong
    L.refh_huf16_decode.argtypes = [P, SZ, P, SZ, P, SZ, I, I]
    L.refh_compress.restype = ctypes.c_longlong
    L.refh_compress.argtypes = [I, I, I, P, SZ, P, SZ, I, I]
    L.refh_decompress.restype = ctypes.c_longlong
    L.refh_decompress.argtypes = [P, SZ, P, SZ, P, P]
    L.refh_compressBound.restype = SZ
    L.refh_compressBound.argtypes = [SZ]


def oracle():
    return _load(ORACLE_SO, _setup_oracle)


def ref():
    return _load(REF_SO, _setup_ref)


def refh():
    return _load(REFH_SO, _setup_refh)


def ptr(a):
    return a.ctypes.data


# graph kinds understood by refh_compress (oracle/ref_harness.c)
STORE, FSE, HUFFMAN, BITPACK, DELTA_TS_FSE, DELTA_ZZ_BP, TS_HUF_FSE, DELTA_FSE, TS_FSE, TS_HUF, ZZ_BP, DELTA_BP, SELECT8, \
    SEGMENTED = range(14)


def ref_set_min_stream_size(v):
    """ZL_CParam_minStreamSize of the reference harness's later compressions (0 = default)."""
    h = refh()
    h.refh_set_min_stream_size.argtypes = [ctypes.c_int]
    h.refh_set_min_stream_size(int(v))


def ref_set_segmenter(kind, chunk_elts):
    """The graph and chunk size (elements) of the reference harness's SEGMENTED kind."""
    h = refh()
    h.refh_set_segmenter.argtypes = [ctypes.c_int, ctypes.c_size_t]
    h.refh_set_segmenter(kind, chunk_elts)
SERIAL, STRUCT, NUMERIC = 0, 1, 2


def ref_compress(kind, arr, typ, width, content_checksum=True, compressed_checksum=True):
    """Compress ``arr`` with the reference library; returns the frame bytes."""
    h = refh()
    arr = np.ascontiguousarray(arr)
    nb = arr.nbytes
    n = nb // width
    cap = h.refh_compressBound(nb) + 4096
    out = np.empty(cap, np.uint8)
    r = h.refh_compress(kind, typ, width, ptr(arr) if nb else None, n, ptr(out), cap,
                        int(content_checksum), int(compressed_checksum))
    if r < 0:
        raise RuntimeError(f"reference compression failed, error code {-r}")
    return out[:r].tobytes()


def ref_decompress(frame, capacity):
    h = refh()
    src = np.frombuffer(frame, np.uint8)
    out = np.empty(max(capacity, 1), np.uint8)
    t = ctypes.c_int(0)
    w = ctypes.c_int(0)
    r = h.refh_decompress(ptr(src), src.size, ptr(out), out.size, ctypes.byref(t), ctypes.byref(w))
    if r < 0:
        raise RuntimeError(f"reference decompression failed, error code {-r}")
    return out[:r].tobytes(), t.value, w.value
