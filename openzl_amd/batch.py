"""Device-resident batch entry points (torch tensors in, torch tensors out).

Thin wrappers over the C ABI of libozl_hip.so; torch only provides device
memory and the current stream.
"""
import ctypes

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


class ColumnBatch:
    """delta_int -> zigzag -> ZL_GRAPH_BITPACK over many numeric columns, one
    frame per column (ozl_columns_compress_batch / ozl_columns_decompress_batch).

    The column layout is fixed at construction: spans (int64, 2K: byte offset
    and byte length of every column in the source buffer, offsets multiples of
    the width) and widths (uint8, K) as device tensors; decoded columns land at
    the same spans of the output buffer. ``tune()`` runs the autotuner on the
    batch; ``graph(direction, ...)`` captures one call in a hipGraph for replay.
    """

    def __init__(self, spans, widths, min_tile=1024, device="cuda"):
        """min_tile: the smallest item size (deltas) the workspace must serve
        (the autotuner's candidates start at 1024)."""
        L = lib()
        self.K = widths.numel()
        self.col_off = spans
        self.widths = widths
        self.total = int(spans.view(-1, 2)[:, 1].sum().item())
        self.ws = torch.empty(L.ozl_columns_workspace(self.K, self.total, min_tile), dtype=torch.uint8,
                              device=device)
        self.cap = self.total + 128 * self.K
        self.dst = torch.empty(self.cap, dtype=torch.uint8, device=device)
        self.frame_off = torch.zeros(self.K + 1, dtype=torch.int64, device=device)
        self.status = torch.zeros(self.K, dtype=torch.int32, device=device)
        self.dstatus = torch.zeros(self.K, dtype=torch.int32, device=device)
        self.tile, self.wg = 4096, 4
        self._graphs = []

    def tune(self, src):
        L = lib()
        t, w = ctypes.c_uint32(0), ctypes.c_uint32(0)
        st = torch.cuda.current_stream().cuda_stream
        check(L.ozl_columns_autotune(src.data_ptr(), self.col_off.data_ptr(), self.widths.data_ptr(), self.K,
                                     self.total, self.dst.data_ptr(), self.cap, self.frame_off.data_ptr(),
                                     self.status.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                     ctypes.byref(t), ctypes.byref(w), st), "ozl_columns_autotune")
        self.tile, self.wg = t.value, w.value
        return self.tile, self.wg

    def compress(self, src):
        """Returns (dst, frame_off, status)."""
        L = lib()
        st = torch.cuda.current_stream().cuda_stream
        check(L.ozl_columns_compress_batch(src.data_ptr(), self.col_off.data_ptr(), self.widths.data_ptr(), self.K,
                                           self.dst.data_ptr(), self.cap, self.frame_off.data_ptr(),
                                           self.status.data_ptr(), self.ws.data_ptr(), self.ws.numel(), self.tile,
                                           self.wg, st), "ozl_columns_compress_batch")
        return self.dst, self.frame_off, self.status

    def decompress(self, frames, frame_off, out):
        """Decodes the K frames into out (column c at col_off[c]); returns status."""
        L = lib()
        st = torch.cuda.current_stream().cuda_stream
        check(L.ozl_columns_decompress_batch(frames.data_ptr(), frame_off.data_ptr(), self.K, out.data_ptr(),
                                             self.col_off.data_ptr(), self.dstatus.data_ptr(), self.ws.data_ptr(),
                                             self.ws.numel(), self.tile, self.wg, st),
              "ozl_columns_decompress_batch")
        return self.dstatus

    def graph(self, direction, a, a_off, b):
        """hipGraph of one call: direction 0 compresses a (source) into the batch's
        frame buffer (b unused); 1 decompresses frames a (offsets a_off) into b."""
        L = lib()
        if direction == 0:
            g = L.ozl_columns_graph_create(0, a.data_ptr(), self.col_off.data_ptr(), self.widths.data_ptr(), self.K,
                                           self.dst.data_ptr(), self.cap, self.frame_off.data_ptr(),
                                           self.status.data_ptr(), self.ws.data_ptr(), self.ws.numel(), self.tile,
                                           self.wg)
        else:
            g = L.ozl_columns_graph_create(1, a.data_ptr(), a_off.data_ptr(), self.widths.data_ptr(), self.K,
                                           b.data_ptr(), 0, self.col_off.data_ptr(), self.dstatus.data_ptr(),
                                           self.ws.data_ptr(), self.ws.numel(), self.tile, self.wg)
        if not g:
            raise RuntimeError("hipGraph capture of the column batch failed")
        self._graphs.append(g)
        return g

    @staticmethod
    def launch(g):
        check(lib().ozl_graph_launch(g, torch.cuda.current_stream().cuda_stream), "ozl_graph_launch")

    def __del__(self):
        try:
            L = lib()
            for g in self._graphs:
                L.ozl_graph_destroy(g)
        except Exception:
            pass
