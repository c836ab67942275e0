"""Batched ZL_GRAPH_FSE / ZL_GRAPH_HUFFMAN frames from the HIP path are
byte-identical to the reference library's frames for the same blocks.

Blocks cover every backend the per-chunk decision can take
(encode_entropy_binding.c:601-753): constant, flatpack, bitpack, store, FSE
with 2 and 4 states, Huffman 1X and 4X, nested weight graphs, inputs under
minStreamSize, and the Zipf(1.1) 64 KiB blocks of the entropy benchmark.
"""
import numpy as np
import pytest

import ozl_checkers as C
from test_oracle import zipf_bytes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    if C.refh() is None:
        pytest.skip("reference library not built")
    from openzl_amd import batch
    return torch, batch


def sample_blocks(seed=7):
    rng = np.random.default_rng(seed)
    blocks = []
    for n in (1, 2, 5, 9, 10, 11, 17, 64, 200, 999, 1000, 1001, 4096, 20000, 65536, 99999):
        blocks.append(zipf_bytes(rng, n, 1.1))
        blocks.append(np.full(n, 42, np.uint8))
        blocks.append(rng.choice(np.array([3, 250], np.uint8), n))
        blocks.append(rng.integers(0, 4, n).astype(np.uint8))
        blocks.append(rng.integers(0, 256, n).astype(np.uint8))
        blocks.append(zipf_bytes(rng, n, 0.5, card=int(rng.integers(5, 60))))
        blocks.append(zipf_bytes(rng, n, 2.5, card=int(rng.integers(2, 256))))
        g = np.minimum(rng.geometric(0.3, n) - 1, 255).astype(np.uint8)
        blocks.append(g)
    return blocks


def run_batch(torch, batch, blocks, mode):
    sizes = np.array([b.size for b in blocks], np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    data = np.concatenate(blocks).astype(np.uint8)
    src = torch.from_numpy(data).cuda()
    off = torch.from_numpy(offs).cuda()
    comp = batch.EntropyBatchCompressor(len(blocks), int(sizes.max()))
    dst, foff, status = comp.run(mode, src, off)
    torch.cuda.synchronize()
    assert not status.cpu().numpy().any(), status.cpu().numpy()
    foff = foff.cpu().numpy()
    out = dst[: int(foff[-1])].cpu().numpy()
    return [out[foff[i]:foff[i + 1]].tobytes() for i in range(len(blocks))]


@pytest.mark.parametrize("mode", [0, 1], ids=["fse", "huffman"])
def test_batch_frames_match_reference(env, mode):
    torch, batch = env
    blocks = sample_blocks()
    print(f"entropy batch mode={mode} blocks={len(blocks)}", flush=True)
    frames = run_batch(torch, batch, blocks, mode)
    kind = C.FSE if mode == 0 else C.HUFFMAN
    bad = []
    for i, (b, f) in enumerate(zip(blocks, frames)):
        ref = C.ref_compress(kind, b, C.SERIAL, 1)
        if f != ref:
            bad.append((i, b.size, len(f), len(ref)))
    assert not bad, bad[:10]


@pytest.mark.parametrize("mode", [0, 1], ids=["fse", "huffman"])
def test_batch_frames_decode_with_reference(env, mode):
    torch, batch = env
    rng = np.random.default_rng(99)
    blocks = [zipf_bytes(rng, 65536, 1.1) for _ in range(64)]
    frames = run_batch(torch, batch, blocks, mode)
    for b, f in zip(blocks, frames):
        out, typ, width = C.ref_decompress(f, b.size + 64)
        assert out == b.tobytes()


@pytest.mark.parametrize("mode", [0, 1], ids=["fse", "huffman"])
def test_batch_frames_of_aligned_blocks_match_reference(env, mode):
    """Blocks that start 16-byte aligned take the histogram + content-checksum
    pass that hashes 1 KiB blocks across a workgroup's waves: sizes around the
    short-input limit (240), one XXH3 block (1024) and the tail stripes."""
    torch, batch = env
    rng = np.random.default_rng(11)
    sizes = [16, 240, 256, 1008, 1024, 1040, 2048, 3072, 4096, 16384, 65536, 65552, 99984, 32, 1029]
    blocks = [zipf_bytes(rng, n, 1.1) if i % 3 else rng.integers(0, 7, n).astype(np.uint8)
              for i, n in enumerate(sizes)]
    print(f"aligned entropy batch mode={mode} blocks={len(blocks)}", flush=True)
    frames = run_batch(torch, batch, blocks, mode)
    kind = C.FSE if mode == 0 else C.HUFFMAN
    bad = [(i, b.size) for i, (b, f) in enumerate(zip(blocks, frames)) if f != C.ref_compress(kind, b, C.SERIAL, 1)]
    assert not bad, bad


def gpu_decode(torch, batch, frames, sizes):
    foffs = np.concatenate([[0], np.cumsum([len(f) for f in frames])]).astype(np.int64)
    doffs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    fr = torch.from_numpy(np.frombuffer(b"".join(frames), np.uint8).copy()).cuda()
    dst = torch.zeros(max(int(doffs[-1]), 1), dtype=torch.uint8, device="cuda")
    dec = batch.EntropyBatchDecompressor(len(frames))
    status = dec.run(fr, torch.from_numpy(foffs).cuda(), dst, torch.from_numpy(doffs).cuda())
    torch.cuda.synchronize()
    out = dst.cpu().numpy()
    return status.cpu().numpy(), [out[doffs[i]:doffs[i + 1]].tobytes() for i in range(len(frames))]


@pytest.mark.parametrize("mode", [0, 1], ids=["fse", "huffman"])
def test_gpu_decodes_reference_frames(env, mode):
    torch, batch = env
    blocks = sample_blocks(seed=21)
    kind = C.FSE if mode == 0 else C.HUFFMAN
    frames = [C.ref_compress(kind, b, C.SERIAL, 1) for b in blocks]
    print(f"decode reference frames mode={mode} n={len(frames)}", flush=True)
    status, outs = gpu_decode(torch, batch, frames, [b.size for b in blocks])
    bad = [(i, blocks[i].size, int(status[i])) for i in range(len(blocks))
           if status[i] != 0 or outs[i] != blocks[i].tobytes()]
    assert not bad, bad[:10]


def test_gpu_decoder_rejects_corruption(env):
    torch, batch = env
    rng = np.random.default_rng(4)
    blocks = [zipf_bytes(rng, 5000, 1.1) for _ in range(4)]
    frames = [bytearray(C.ref_compress(C.FSE, b, C.SERIAL, 1)) for b in blocks]
    frames[1][len(frames[1]) // 2] ^= 0x10   # payload bit flip -> compressed checksum
    frames[2][-6] ^= 0x01                    # stored compressed checksum
    status, outs = gpu_decode(torch, batch, [bytes(f) for f in frames], [b.size for b in blocks])
    assert status[0] == 0 and outs[0] == blocks[0].tobytes()
    assert status[1] != 0 and status[2] != 0
    assert status[3] == 0
