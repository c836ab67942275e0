"""Frames from the ZL_* API of libozl_hip.so (graph engine + GPU codecs) are
byte-identical to the reference library's frames for the same graphs and
inputs, and the GPU universal decompressor decodes both.

Graphs: the timestamps pipeline (delta_int -> transpose_split -> FSE), the
column pipeline (delta_int -> zigzag -> BITPACK), the weight-tensor function
graph (transpose_split, plane 0 FSE, plane 1 Huffman), and the other graphs of
the selector candidates, at sizes that exercise 32 KiB chunking (>= 100000
elements), minStreamSize auto-store and tiny / empty inputs.
"""
import ctypes

import numpy as np
import pytest

import ozl_checkers as C
from test_oracle import zipf_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def api():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if C.refh() is None:
        pytest.skip("reference library not built")
    from openzl_amd import api as A
    return A


def cases():
    from openzl_amd import datagen as G
    rng = np.random.default_rng(17)
    out = []
    for n in (0, 1, 5, 100, 1000, 30000, 130000, 262144):
        ts = G.timestamps_u32(max(n, 1))[:n]
        out.append(("cfg1", 4, ts, 2, 4))              # DELTA_TS_FSE on numeric u32
    cols = G.random_walk_columns(n_cols=12, min_len=16, max_len=300000)
    for c in cols:
        out.append(("cfg3", 5, c, 2, c.dtype.itemsize))  # DELTA_ZZ_BP
    w = G.bf16_normal(200000)
    out.append(("cfg4", 6, w, 1, 2))                    # TS_HUF_FSE on struct bf16
    out.append(("cfg4s", 6, w[:3000], 1, 2))
    for n in (10, 500, 5000, 99999, 100000, 150000):
        x = zipf_bytes(rng, n, 1.1)
        out.append(("fse", 1, x, 0, 1))
        out.append(("huf", 2, x, 0, 1))
        out.append(("tsfse", 8, rng.integers(0, 60, n // 4 * 4).astype(np.uint8).view(np.uint32), 1, 4))
        out.append(("tshuf", 9, zipf_bytes(rng, n // 2 * 2, 1.3).view(np.uint16), 1, 2))
    out.append(("fse_const", 1, np.full(200000, 3, np.uint8), 0, 1))
    out.append(("fse_rand", 1, rng.integers(0, 256, 120000).astype(np.uint8), 0, 1))
    out.append(("huf_2sym", 2, rng.integers(0, 2, 70000).astype(np.uint8), 0, 1))
    out.append(("bp_num", 3, rng.integers(0, 1000, 5000).astype(np.uint16), 2, 2))
    out.append(("bp_ser", 3, rng.integers(0, 7, 5000).astype(np.uint8), 0, 1))
    out.append(("dfse", 7, np.cumsum(rng.integers(0, 9, 150000)).astype(np.uint32), 2, 4))
    out.append(("zzbp", 10, rng.integers(-50, 50, 9000).astype(np.int16).view(np.uint16), 2, 2))
    out.append(("dbp", 11, np.cumsum(rng.integers(0, 5, 9000)).astype(np.uint64), 2, 8))
    return out


TYPES = {0: C.SERIAL, 1: C.STRUCT, 2: C.NUMERIC}


def test_engine_frames_match_reference(api):
    A = api
    bad = []
    for name, kind, arr, typ, width in cases():
        print(f"frame {name} kind={kind} bytes={arr.nbytes}", flush=True)
        ref = C.ref_compress(kind, arr, TYPES[typ], width)
        comp = A.Compressor(kind)
        got = comp.compress(arr, {0: A.SERIAL, 1: A.STRUCT, 2: A.NUMERIC}[typ], width)
        if got != ref:
            bad.append((name, arr.nbytes, len(got), len(ref)))
    assert not bad, bad


def test_gpu_decompressor_roundtrips(api):
    A = api
    dec = A.Decompressor()
    bad = []
    for name, kind, arr, typ, width in cases():
        ref = C.ref_compress(kind, arr, TYPES[typ], width)
        out, t, w = dec.decompress(ref, arr.nbytes + 64)
        if out != arr.tobytes():
            bad.append((name, arr.nbytes))
    assert not bad, bad


def batch_cases():
    """Groups of inputs for one batch call each: (kind, type, width, arrays)."""
    from openzl_amd import datagen as G
    rng = np.random.default_rng(23)
    ts = G.timestamps_u32(600000)
    w = G.bf16_normal(1 << 20)
    return [
        ("cfg1", 4, 2, 4, [ts[:262144], ts[262144:262144 + 5], ts[:0], ts[300000:600000]]),
        ("cfg4", 6, 1, 2, [w[:300000], w[300000:300001], w[300001:]]),
        ("fse", 1, 0, 1, [zipf_bytes(rng, n, 1.1) for n in (10, 5000, 99999, 150000)]),
        ("huf", 2, 0, 1, [zipf_bytes(rng, n, 1.2) for n in (150000, 0, 70000)]),
        ("zzbp", 5, 2, 2, [c for c in G.random_walk_columns(n_cols=5, min_len=16, max_len=50000)
                           if c.dtype.itemsize == 2] or [np.arange(100, dtype=np.uint16)]),
    ]


def test_batch_frames_match_reference(api):
    """ozl_CCtx_compressBatch writes, per input, the reference's frame, and
    ozl_DCtx_decompressBatch returns the inputs, for host buffers."""
    A = api
    dec = A.Decompressor()
    for name, kind, typ, width, arrs in batch_cases():
        print(f"batch {name} inputs={len(arrs)}", flush=True)
        comp = A.Compressor(kind)
        at = {0: A.SERIAL, 1: A.STRUCT, 2: A.NUMERIC}[typ]
        frames = comp.compress_batch(arrs, at, width)
        for a, f in zip(arrs, frames):
            assert f == C.ref_compress(kind, a, TYPES[typ], width), (name, a.nbytes)
        outs = dec.decompress_batch(frames, sum(a.nbytes for a in arrs) + 64)
        assert outs == [a.tobytes() for a in arrs], name


def test_batch_device_buffers(api):
    """Inputs, frames and outputs in device memory: the same frames as from
    host buffers, decoded back in place; single calls take device memory too."""
    import torch
    A = api
    dec = A.Decompressor()
    for name, kind, typ, width, arrs in batch_cases()[:3]:
        print(f"device batch {name}", flush=True)
        comp = A.Compressor(kind)
        at = {0: A.SERIAL, 1: A.STRUCT, 2: A.NUMERIC}[typ]
        want = comp.compress_batch(arrs, at, width)
        dins = [torch.from_numpy(a.view(np.uint8).copy()).cuda() for a in arrs]
        cap = sum(len(f) for f in want) + 4096
        dfr = torch.zeros(cap, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        sizes = comp.compress_ptrs([t.data_ptr() for t in dins], [a.nbytes for a in arrs], at, width,
                                   dfr.data_ptr(), cap)
        got = dfr[:sum(sizes)].cpu().numpy().tobytes()
        assert sizes == [len(f) for f in want] and got == b"".join(want), name
        total = sum(a.nbytes for a in arrs)
        dout = torch.zeros(total + 64, dtype=torch.uint8, device="cuda")
        outs = dec.decompress_ptrs(dfr.data_ptr(), sizes, dout.data_ptr(), total + 64)
        assert outs == [a.nbytes for a in arrs], name
        assert dout[:total].cpu().numpy().tobytes() == b"".join(a.tobytes() for a in arrs), name
        # one frame through the single-frame calls, device in and out
        L = comp.L
        tr = comp._typed_ref(dins[-1].data_ptr(), arrs[-1].nbytes, at, width)
        try:
            z = A._chk(L.ZL_CCtx_compressTypedRef(comp.cctx, dfr.data_ptr(), cap, tr), "compress")
        finally:
            L.ZL_TypedRef_free(tr)
        assert dfr[:z].cpu().numpy().tobytes() == want[-1], name
        info = A.OutputInfo()
        m = A._chk(dec.L.ZL_DCtx_decompressTyped(dec.d, ctypes.byref(info), dout.data_ptr(), total + 64,
                                                 dfr.data_ptr(), z), "decompress")
        assert dout[:m].cpu().numpy().tobytes() == arrs[-1].tobytes(), name


def test_device_frames_end_marker(api):
    """Device-resident frames: the end-of-frame byte after the last chunk is
    read with the checksum tails (not in a header round of its own); a
    nonzero one is rejected as corruption, in any frame of a batch."""
    import torch
    A = api
    dec = A.Decompressor()
    for name, kind, typ, width, arrs in batch_cases()[:3]:
        comp = A.Compressor(kind)
        at = {0: A.SERIAL, 1: A.STRUCT, 2: A.NUMERIC}[typ]
        frames = comp.compress_batch(arrs, at, width)
        sizes = [len(f) for f in frames]
        total = sum(a.nbytes for a in arrs)
        for bad in range(len(frames)):
            print(f"end marker {name} frame {bad}", flush=True)
            blob = bytearray(b"".join(frames))
            end = sum(sizes[:bad + 1]) - 1
            assert blob[end] == 0
            blob[end] = 1
            dfr = torch.from_numpy(np.frombuffer(bytes(blob), np.uint8).copy()).cuda()
            dout = torch.zeros(total + 64, dtype=torch.uint8, device="cuda")
            with pytest.raises(A.ZLError):
                dec.decompress_ptrs(dfr.data_ptr(), sizes, dout.data_ptr(), total + 64)
        dfr = torch.from_numpy(np.frombuffer(b"".join(frames), np.uint8).copy()).cuda()
        dout = torch.zeros(total + 64, dtype=torch.uint8, device="cuda")
        assert dec.decompress_ptrs(dfr.data_ptr(), sizes, dout.data_ptr(), total + 64) == [a.nbytes for a in arrs]
        assert dout[:total].cpu().numpy().tobytes() == b"".join(a.tobytes() for a in arrs), name


def segmented_cases():
    """(inner kind, type, width, chunk elements, array): frames of several chunks."""
    from openzl_amd import datagen as G
    rng = np.random.default_rng(29)
    ts = G.timestamps_u32(300000)
    w = G.bf16_normal(400000)
    return [
        (1, 0, 1, 50000, zipf_bytes(rng, 180000, 1.1)),       # FSE, 4 chunks (last shorter)
        (2, 0, 1, 100000, zipf_bytes(rng, 100000, 1.2)),      # Huffman, exactly one chunk
        (4, 2, 4, 120000, ts),                                # DELTA_TS_FSE on u32, 3 chunks
        (6, 1, 2, 150000, w),                                 # transpose_split HUF/FSE on bf16
        (1, 0, 1, 4096, rng.integers(0, 3, 10000).astype(np.uint8)),
    ]


def test_segmented_frames_match_reference(api):
    """A segmenter (ZL_Compressor_registerSegmenter) cutting fixed-size
    chunks: frames of several chunks equal the reference's, and the universal
    decompressor reads them (single calls, batch calls, device buffers)."""
    import torch
    A = api
    dec = A.Decompressor()
    frames, arrs = [], []
    for inner, typ, width, elts, arr in segmented_cases():
        print(f"segmented inner={inner} elts={elts} bytes={arr.nbytes}", flush=True)
        C.ref_set_segmenter(inner, elts)
        ref = C.ref_compress(C.SEGMENTED, arr, TYPES[typ], width)
        comp = A.Compressor(A.SEGMENTED, segment=(inner, elts))
        got = comp.compress(arr, {0: A.SERIAL, 1: A.STRUCT, 2: A.NUMERIC}[typ], width)
        assert got == ref, (inner, elts, arr.nbytes, len(got), len(ref))
        out, _, _ = dec.decompress(ref, arr.nbytes + 64)
        assert out == arr.tobytes(), (inner, elts)
        frames.append(ref)
        arrs.append(arr)
    total = sum(a.nbytes for a in arrs)
    outs = dec.decompress_batch(frames, total + 64)
    assert outs == [a.tobytes() for a in arrs]
    # ZL_getDecompressedSize reads the frame header (host and device frames)
    L = A._api()
    for f, a in zip(frames, arrs):
        src = np.frombuffer(f, np.uint8)
        assert A._chk(L.ZL_getDecompressedSize(src.ctypes.data, src.size), "size") == a.nbytes
        d = torch.from_numpy(src.copy()).cuda()
        torch.cuda.synchronize()
        assert A._chk(L.ZL_getDecompressedSize(d.data_ptr(), src.size), "size") == a.nbytes
    dfr = torch.from_numpy(np.frombuffer(b"".join(frames), np.uint8).copy()).cuda()
    dout = torch.zeros(total + 64, dtype=torch.uint8, device="cuda")
    sizes = dec.decompress_ptrs(dfr.data_ptr(), [len(f) for f in frames], dout.data_ptr(), total + 64)
    assert sizes == [a.nbytes for a in arrs]
    assert dout[:total].cpu().numpy().tobytes() == b"".join(a.tobytes() for a in arrs)
