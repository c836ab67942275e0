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
