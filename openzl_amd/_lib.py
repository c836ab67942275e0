"""Loads libozl_hip.so (the C ABI) and declares its entry points for ctypes.

The product path is the HIP library: if it is missing this module raises, it
never falls back to any CPU implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libozl_hip.so")

P = ctypes.c_void_p
SZ = ctypes.c_size_t
I = ctypes.c_int

_lib = None


class OzlError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: OpenZL-AMD error code {code}")
        self.code = code


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build() or python openzl_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    sigs = {
        "ozl_delta_encode": (I, [P, SZ, I, P, P, P]),
        "ozl_delta_decode": (I, [P, P, SZ, I, P, P, P]),
        "ozl_delta_decode_scratch_size": (SZ, [SZ, I]),
        "ozl_zigzag_encode": (I, [P, SZ, I, P, P]),
        "ozl_zigzag_decode": (I, [P, SZ, I, P, P]),
        "ozl_transpose_split": (I, [P, SZ, I, P, P]),
        "ozl_transpose_join": (I, [P, SZ, I, P, P]),
        "ozl_max_async": (I, [P, SZ, I, P, P]),
        "ozl_bitpack_encode": (I, [P, SZ, I, I, P, P]),
        "ozl_bitpack_decode": (I, [P, SZ, I, I, P, P]),
        "ozl_entropy_compress_workspace": (SZ, [ctypes.c_uint32, SZ]),
        "ozl_entropy_frame_bound": (SZ, [SZ]),
        "ozl_entropy_compress_batch": (I, [I, P, P, ctypes.c_uint32, SZ, P, SZ, P, P, P, SZ, P]),
        "ozl_entropy_decompress_workspace": (SZ, [ctypes.c_uint32]),
        "ozl_entropy_decompress_batch": (I, [P, P, ctypes.c_uint32, P, P, P, P, SZ, P]),
        "ozl_columns_workspace": (SZ, [ctypes.c_uint32, SZ, ctypes.c_uint32]),
        "ozl_columns_frame_bound": (SZ, [SZ]),
        "ozl_columns_compress_batch": (I, [P, P, P, ctypes.c_uint32, P, SZ, P, P, P, SZ, ctypes.c_uint32,
                                           ctypes.c_uint32, P]),
        "ozl_columns_decompress_batch": (I, [P, P, ctypes.c_uint32, P, P, P, P, SZ, ctypes.c_uint32,
                                             ctypes.c_uint32, P]),
        "ozl_columns_autotune": (I, [P, P, P, ctypes.c_uint32, ctypes.c_uint64, P, SZ, P, P, P, SZ, P, P, P]),
        "ozl_columns_graph_create": (P, [I, P, P, P, ctypes.c_uint32, P, SZ, P, P, P, SZ, ctypes.c_uint32,
                                         ctypes.c_uint32]),
        "ozl_select_workspace": (SZ, [ctypes.c_uint32, SZ]),
        "ozl_select_frame_bound": (SZ, [SZ]),
        "ozl_select_trial": (I, [P, P, P, ctypes.c_uint32, SZ, P, P, P, SZ, P]),
        "ozl_select_compress": (I, [P, P, P, ctypes.c_uint32, SZ, P, P, SZ, P, P, P, SZ, P]),
        "ozl_select_mlp_params_bytes": (SZ, []),
        "ozl_select_features": (I, [P, P, P, ctypes.c_uint32, P, P]),
        "ozl_select_mlp": (I, [P, ctypes.c_uint32, P, P, P, P]),
        "ozl_graph_launch": (I, [P, P]),
        "ozl_graph_destroy": (None, [P]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(code, what):
    if code != 0:
        raise OzlError(code, what)
