"""GPU path against the committed golden fixtures (reference frames written by
tests/golden/make_golden.py); needs no reference library at run time.

Small cases (every graph of the hot path, chunking, minStreamSize stores,
empty-ish inputs): the ZL_* engine writes byte-identical frames, the GPU
universal decompressor returns the inputs from the reference frames, and the
oracle's CPU decoder returns the inputs from the GPU frames.

Full-size workloads: per-frame digests of the reference frames for the
timestamps (16 MiB, one frame), the entropy stress (16,384 x 64 KiB, FSE and
Huffman frames), the batched columns (4,096 frames) and the bf16 weights (64
frames of 8 Mi elements); the GPU frames must match every digest and decode
back to the input.
"""
import numpy as np
import pytest

import golden_cases as GC
import oracle_frames as OF

pytestmark = pytest.mark.gpu

ENTRIES = GC.small_entries()
HUF16 = GC.huf16_entries()
TYPES = {0: 1, 1: 2, 2: 4}  # frame type -> ZL_Type


@pytest.fixture(scope="module")
def api():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from openzl_amd import api as A
    return A


def test_small_engine_frames_are_golden(api):
    bad = []
    for e in ENTRIES:
        if e["name"] in HUF16:
            continue  # test_small_engine_huffman_struct_v2_frames
        x = GC.make_input(e["gen"], e["kwargs"])
        print(f"golden engine {e['name']} kind={e['kind']} bytes={x.nbytes}", flush=True)
        got = api.Compressor(e["kind"], params=e.get("params")).compress(x, TYPES[e["type"]], e["width"])
        if GC.sha256(got) != e["frame_sha256"]:
            bad.append((e["name"], len(got), e["frame_bytes"]))
            continue
        out, _, _ = OF.decode_frame(got)
        assert out == x.tobytes(), e["name"]
    assert not bad, bad


@pytest.mark.skip(reason=GC.HUF16_ENCODE_OFF)
def test_small_engine_huffman_struct_v2_frames(api):
    for e in ENTRIES:
        if e["name"] not in HUF16:
            continue
        x = GC.make_input(e["gen"], e["kwargs"])
        got = api.Compressor(e["kind"], params=e.get("params")).compress(x, TYPES[e["type"]], e["width"])
        assert GC.sha256(got) == e["frame_sha256"], e["name"]


def test_small_gpu_decompressor_on_golden_frames(api):
    dec = api.Decompressor()
    bad = []
    for e in ENTRIES:
        x = GC.make_input(e["gen"], e["kwargs"])
        print(f"golden decode {e['name']}", flush=True)
        try:
            out, typ, width = dec.decompress(GC.small_frame(e["name"]), x.nbytes + 64)
        except api.ZLError as err:
            bad.append((e["name"], str(err)))
            continue
        if out != x.tobytes() or typ != TYPES[e["type"]]:
            bad.append((e["name"], "mismatch", typ))
    assert not bad, bad


def test_decompressor_rejects_corrupted_golden_frame(api):
    dec = api.Decompressor()
    e = next(e for e in ENTRIES if e["name"] == "ts_delta_ts_fse_30000")
    fr = bytearray(GC.small_frame(e["name"]))
    fr[len(fr) // 2] ^= 0x04
    with pytest.raises(api.ZLError):
        dec.decompress(bytes(fr), e["input_bytes"] + 64)


# ------------------------------------------------------------------ full size
META = GC.full_meta()
needs_full = pytest.mark.skipif(META is None, reason="full-size digests not generated")


def _digests(frames):
    return np.array([[len(f), GC.frame_digest(f)] for f in frames], np.uint64)


@needs_full
def test_full_timestamps_frame(api):
    from openzl_amd import datagen as G
    m = META["timestamps_delta_ts_fse"]
    x = G.timestamps_u32()
    assert GC.sha256(x.tobytes()) == m["input_sha256"]
    print("full timestamps: compress 16 MiB", flush=True)
    fr = api.Compressor(m["kind"]).compress(x, api.NUMERIC, 4)
    assert np.array_equal(_digests([fr]), GC.full_table("timestamps_delta_ts_fse"))
    out, _, _ = api.Decompressor().decompress(fr, x.nbytes + 64)
    assert out == x.tobytes()


@needs_full
@pytest.mark.parametrize("mode,name", [(0, "zipf_blocks_fse"), (1, "zipf_blocks_huf")])
def test_full_entropy_stress_frames(api, mode, name):
    import torch
    from openzl_amd import batch, datagen as G
    m = META[name]
    nb = m["frames"]
    print(f"full entropy stress {name}: {nb} blocks", flush=True)
    src = G.zipf_blocks_torch(torch, nb)
    off = torch.arange(0, (nb + 1) * 65536, 65536, dtype=torch.int64, device="cuda")
    comp = batch.EntropyBatchCompressor(nb, 65536)
    dst, foff, status = comp.run(mode, src, off)
    torch.cuda.synchronize()
    assert not status.any().item()
    fo = foff.cpu().numpy()
    host = dst[: int(fo[-1])].cpu().numpy()
    frames = [host[fo[i]:fo[i + 1]].tobytes() for i in range(nb)]
    ref = GC.full_table(name)
    got = _digests(frames)
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, bad[:10]
    back = torch.empty_like(src)
    dec = batch.EntropyBatchDecompressor(nb)
    st = dec.run(dst, foff, back, off)
    torch.cuda.synchronize()
    assert not st.any().item()
    assert torch.equal(back, src)


@needs_full
def test_full_columns_frames(api):
    from openzl_amd import datagen as G
    m = META["columns_delta_zz_bp"]
    cols = G.random_walk_columns()
    print(f"full columns: {len(cols)} frames", flush=True)
    comp = api.Compressor(m["kind"])
    frames = [comp.compress(c, api.NUMERIC, c.dtype.itemsize) for c in cols]
    got = _digests(frames)
    ref = GC.full_table("columns_delta_zz_bp")
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, bad[:10]
    dec = api.Decompressor()
    for i in range(0, len(cols), 97):
        out, _, _ = dec.decompress(frames[i], cols[i].nbytes + 64)
        assert out == cols[i].tobytes(), i


@needs_full
def test_full_bf16_frames(api):
    from openzl_amd import datagen as G
    m = META["bf16_ts_huf_fse"]
    w = G.bf16_normal()
    seg = m["segment_elements"]
    print(f"full bf16: {m['frames']} segments", flush=True)
    comp = api.Compressor(m["kind"])
    frames = [comp.compress(w[a:a + seg], api.STRUCT, 2) for a in range(0, w.size, seg)]
    got = _digests(frames)
    ref = GC.full_table("bf16_ts_huf_fse")
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, bad[:10]
    dec = api.Decompressor()
    for k in (0, len(frames) - 1):
        out, _, _ = dec.decompress(frames[k], seg * 2 + 64)
        assert out == w[k * seg:(k + 1) * seg].tobytes()
