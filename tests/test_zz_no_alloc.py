"""ozl_entropy_decompress_batch_noalloc (the allocation-free variant of
ozl_entropy_decompress_batch, scratch in the library's own device array)
through the C ABI: on two calls in a row, the second with fewer and smaller
frames, its outputs and statuses equal the workspace entry point's, byte for
byte; more frames than its scratch holds is internalBuffer_tooSmall; a call
leaves the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from openzl_amd import _lib, batch
    L = _lib.lib()
    P, U32, SZ = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_size_t
    L.ozl_entropy_decompress_batch_noalloc.restype = ctypes.c_int
    L.ozl_entropy_decompress_batch_noalloc.argtypes = [P, P, U32, P, P, P, P]
    L.ozl_entropy_decompress_noalloc_max_frames.restype = SZ
    return torch, L, batch


def _frames(torch, batch, mode, sizes, seed):
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, 257) ** 1.1
    p /= p.sum()
    data = np.concatenate([rng.choice(256, n, p=p).astype(np.uint8) for n in sizes])
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    src = torch.from_numpy(data).cuda()
    doff = torch.from_numpy(off).cuda()
    enc = batch.EntropyBatchCompressor(len(sizes), max(sizes))
    dst, foff, st = enc.run(mode, src, doff)
    torch.cuda.synchronize()
    assert int(st.abs().sum()) == 0
    n = int(foff[-1].item())
    return src, doff, dst[:n].clone(), foff.clone()


@pytest.mark.parametrize("mode", [0, 1], ids=["fse", "huf"])
def test_noalloc_matches_workspace_entry_point(env, mode):
    torch, L, batch = env
    st = torch.cuda.current_stream().cuda_stream
    calls = [([65536] * 300 + [40000, 1000, 17, 1, 99999], 1), ([5000] * 40 + [123, 64], 2)]
    for k, (sizes, seed) in enumerate(calls):
        print(f"noalloc mode={mode} call={k} frames={len(sizes)}", flush=True)
        src, doff, frames, foff = _frames(torch, batch, mode, sizes, seed + 10 * mode)
        B = len(sizes)
        ref_out = torch.zeros_like(src)
        ref_st = batch.EntropyBatchDecompressor(B).run(frames, foff, ref_out, doff).clone()
        out = torch.full_like(src, 0xA5)
        status = torch.full((B,), -7, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = L.ozl_entropy_decompress_batch_noalloc(frames.data_ptr(), foff.data_ptr(), B, out.data_ptr(),
                                                    doff.data_ptr(), status.data_ptr(), st)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert rc == 0
        assert torch.equal(status, ref_st) and int(status.abs().sum()) == 0
        assert torch.equal(out, ref_out) and torch.equal(out, src)
        assert free1 >= free0 - (2 << 20), (free0, free1)


def test_noalloc_rejects_more_frames_than_its_scratch(env):
    torch, L, batch = env
    mx = int(L.ozl_entropy_decompress_noalloc_max_frames())
    assert mx == 16384
    dummy = torch.zeros(16, dtype=torch.int64, device="cuda")
    rc = L.ozl_entropy_decompress_batch_noalloc(dummy.data_ptr(), dummy.data_ptr(), mx + 1, dummy.data_ptr(),
                                                dummy.data_ptr(), dummy.data_ptr(),
                                                torch.cuda.current_stream().cuda_stream)
    assert rc == 71
