"""The host reader of frame and chunk headers (openzl_amd/csrc/ozl_hparse.h,
CPU unit build) on the reference's own frames: frames of several chunks made
by a segmenter (include/openzl/zl_segmenter.h) and single-chunk frames, read
whole and through small windows (as device frames are read)."""
import ctypes
import os

import numpy as np
import pytest

import ozl_checkers as C
from test_oracle import zipf_bytes

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "cpu_unit", "libozl_core_cpu.so")


@pytest.fixture(scope="module")
def shim():
    if not os.path.exists(SO):
        pytest.skip("CPU unit build missing")
    if C.refh() is None:
        pytest.skip("reference library not built")
    L = ctypes.CDLL(SO)
    L.cu_parse_frame.restype = ctypes.c_int
    L.cu_parse_frame.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t]
    return L


def parse(L, frame, window):
    f = np.frombuffer(frame, np.uint8)
    out = np.zeros(1 << 14, np.uint64)
    rc = L.cu_parse_frame(f.ctypes.data, f.size, window, out.ctypes.data, out.size)
    return rc, out


def frames():
    rng = np.random.default_rng(31)
    x = zipf_bytes(rng, 180000, 1.1)
    w = rng.integers(0, 6, 90000).astype(np.uint8)
    out = []
    C.ref_set_segmenter(C.FSE, 50000)
    out.append((C.ref_compress(C.SEGMENTED, x, C.SERIAL, 1), 4))
    C.ref_set_segmenter(C.HUFFMAN, 7000)
    out.append((C.ref_compress(C.SEGMENTED, w, C.SERIAL, 1), 13))
    out.append((C.ref_compress(C.FSE, x, C.SERIAL, 1), 1))
    out.append((C.ref_compress(C.HUFFMAN, w[:5000], C.SERIAL, 1), 1))
    return out


def test_chunks_tile_the_frame(shim):
    for frame, nchunks in frames():
        rc, out = parse(shim, frame, 0)
        assert rc == 0
        n, hdr_end, end = int(out[0]), int(out[1]), int(out[2])
        assert n == nchunks
        assert end == len(frame) - 1 and frame[end] == 0       # end-of-frame marker is the last byte
        prev = hdr_end
        for k in range(n):
            start, data, csum, cend = (int(v) for v in out[3 + 6 * k: 3 + 6 * k + 4])
            assert start == prev and start < data <= csum < cend  # chunks follow each other
            assert cend - csum == 8                             # both checksums
            prev = cend
        assert prev == end


@pytest.mark.parametrize("window", [16, 64, 300, 4096])
def test_windows_read_the_same_headers(shim, window):
    for frame, _ in frames():
        rc0, whole = parse(shim, frame, 0)
        rc, win = parse(shim, frame, window)
        assert rc0 == 0 and rc == 0
        assert np.array_equal(whole, win)


def test_truncated_frame_is_corruption(shim):
    frame, _ = frames()[0]
    for cut in (len(frame) - 1, len(frame) // 2, 40):
        rc, _ = parse(shim, frame[:cut], 0)
        assert rc != 0
