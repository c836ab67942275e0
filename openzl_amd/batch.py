"""Device-resident batch entry points (torch tensors in, torch tensors out).

Thin wrappers over the C ABI of libozl_hip.so; torch only provides device
memory and the current stream.
"""
import torch

from ._lib import lib, check

FSE, HUFFMAN = 0, 1


class EntropyBatchCompressor:
    """Reusable workspace for ozl_entropy_compress_batch on blocks <= max_block."""

    def __init__(self, n_blocks, max_block, device="cuda"):
        L = lib()
        self.n_blocks = n_blocks
        self.max_block = max_block
        self.ws = torch.empty(L.ozl_entropy_compress_workspace(n_blocks, max_block), dtype=torch.uint8, device=device)
        self.cap = n_blocks * L.ozl_entropy_frame_bound(max_block)
        self.dst = torch.empty(self.cap, dtype=torch.uint8, device=device)
        self.frame_off = torch.zeros(n_blocks + 1, dtype=torch.int64, device=device)
        self.status = torch.zeros(n_blocks, dtype=torch.int32, device=device)

    def run(self, mode, src, offsets):
        """src: uint8 device tensor, offsets: int64 device tensor (n_blocks+1).
        Returns (dst, frame_off, status); frames are dst[frame_off[i]:frame_off[i+1]],
        status[i] is 0 or the error code of block i."""
        L = lib()
        B = offsets.numel() - 1
        st = torch.cuda.current_stream().cuda_stream
        check(L.ozl_entropy_compress_batch(mode, src.data_ptr(), offsets.data_ptr(), B, self.max_block,
                                           self.dst.data_ptr(), self.cap, self.frame_off.data_ptr(),
                                           self.status.data_ptr(), self.ws.data_ptr(), self.ws.numel(), st),
              "ozl_entropy_compress_batch")
        return self.dst, self.frame_off, self.status[:B]


class EntropyBatchDecompressor:
    """Reusable workspace for ozl_entropy_decompress_batch."""

    def __init__(self, n_blocks, device="cuda"):
        L = lib()
        self.ws = torch.empty(L.ozl_entropy_decompress_workspace(n_blocks), dtype=torch.uint8, device=device)
        self.status = torch.zeros(n_blocks, dtype=torch.int32, device=device)

    def run(self, frames, frame_off, dst, dst_off):
        """Decodes frames[frame_off[i]:frame_off[i+1]] into dst[dst_off[i]:dst_off[i+1]].
        Returns the per-frame status tensor (0 = ok, else a ZL error code)."""
        L = lib()
        B = frame_off.numel() - 1
        st = torch.cuda.current_stream().cuda_stream
        check(L.ozl_entropy_decompress_batch(frames.data_ptr(), frame_off.data_ptr(), B, dst.data_ptr(),
                                             dst_off.data_ptr(), self.status.data_ptr(), self.ws.data_ptr(),
                                             self.ws.numel(), st),
              "ozl_entropy_decompress_batch")
        return self.status[:B]
