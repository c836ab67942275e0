"""Learned graph selector over eight candidate graphs, batched on the GPU.

Thin wrappers over the C ABI of libozl_hip.so (ozl_select_* entry points):
exhaustive trial of the candidates, compression of every chunk with a chosen
candidate, the per-chunk features and the MLP that predicts the choice. The
candidates, in trial order (their graphs as the reference builds them):

    0 ZL_GRAPH_STORE
    1 ZL_GRAPH_FSE
    2 ZL_GRAPH_BITPACK
    3 ZL_NODE_DELTA_INT -> ZL_GRAPH_FSE
    4 ZL_NODE_DELTA_INT -> ZL_NODE_ZIGZAG -> ZL_GRAPH_BITPACK
    5 transpose_split -> ZL_GRAPH_FSE
    6 transpose_split -> ZL_GRAPH_HUFFMAN
    7 ZL_NODE_DELTA_INT -> transpose_split -> ZL_GRAPH_FSE

The exhaustive trial is the reference's brute-force selector
(src/openzl/compress/selectors/selector_brute_force.c:9-50) registered over
these eight graphs with ZL_Compressor_registerBruteForceSelectorGraph.
"""
import os

import numpy as np
import torch

from ._lib import lib, check

CANDIDATES = ("store", "fse", "bitpack", "delta_fse", "delta_zigzag_bitpack", "transpose_fse", "transpose_huffman",
              "delta_transpose_fse")
N_FEATURES = 40
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "selector_mlp.bin")


def load_params(path=MODEL_PATH, device="cuda"):
    """The MLP parameter block (bf16 weights, fp32 biases) as a device tensor."""
    L = lib()
    raw = np.fromfile(path, dtype=np.uint8)
    if raw.size != L.ozl_select_mlp_params_bytes():
        raise ValueError(f"{path}: {raw.size} bytes, expected {L.ozl_select_mlp_params_bytes()}")
    return torch.from_numpy(raw).to(device)


class SelectorBatch:
    """K numeric chunks: chunk c = src[spans[2c], spans[2c] + spans[2c+1]) of
    element width widths[c]; spans and widths are device tensors (int64 2K,
    uint8 K). Chunks are at most max_chunk (< 100000) bytes."""

    def __init__(self, spans, widths, max_chunk, total_bytes=None, device="cuda"):
        L = lib()
        self.spans, self.widths = spans, widths
        self.K = widths.numel()
        self.max_chunk = int(max_chunk)
        if total_bytes is None:
            total_bytes = self.K * self.max_chunk
        self.ws = torch.empty(L.ozl_select_workspace(self.K, self.max_chunk), dtype=torch.uint8, device=device)
        self.cap = int(total_bytes + self.K * (L.ozl_select_frame_bound(0) + self.max_chunk // 64 + 64))
        self.dst = torch.empty(self.cap, dtype=torch.uint8, device=device)
        self.frame_off = torch.zeros(self.K + 1, dtype=torch.int64, device=device)
        self.status = torch.zeros(self.K, dtype=torch.int32, device=device)
        self.sizes = torch.zeros((self.K, 8), dtype=torch.int32, device=device)
        self.choice = torch.zeros(self.K, dtype=torch.uint8, device=device)
        self.feats = torch.zeros((self.K, N_FEATURES), dtype=torch.float32, device=device)
        self.logits = torch.zeros((self.K, 8), dtype=torch.float32, device=device)

    def _st(self):
        return torch.cuda.current_stream().cuda_stream

    def trial(self, src):
        """Exhaustive trial: (sizes K x 8 with checksums off, 0 = failed;
        choice K, the brute-force selector's pick)."""
        L = lib()
        check(L.ozl_select_trial(src.data_ptr(), self.spans.data_ptr(), self.widths.data_ptr(), self.K, self.max_chunk,
                                 self.sizes.data_ptr(), self.choice.data_ptr(), self.ws.data_ptr(), self.ws.numel(),
                                 self._st()), "ozl_select_trial")
        return self.sizes, self.choice

    def compress(self, src, choice):
        """Frames of every chunk with candidate choice[c]: (dst, frame_off, status)."""
        L = lib()
        check(L.ozl_select_compress(src.data_ptr(), self.spans.data_ptr(), self.widths.data_ptr(), self.K,
                                    self.max_chunk, choice.data_ptr(), self.dst.data_ptr(), self.cap,
                                    self.frame_off.data_ptr(), self.status.data_ptr(), self.ws.data_ptr(),
                                    self.ws.numel(), self._st()), "ozl_select_compress")
        return self.dst, self.frame_off, self.status

    def features(self, src):
        L = lib()
        check(L.ozl_select_features(src.data_ptr(), self.spans.data_ptr(), self.widths.data_ptr(), self.K,
                                    self.feats.data_ptr(), self._st()), "ozl_select_features")
        return self.feats

    def predict(self, src, params):
        """Features, then the MLP: (choice K, logits K x 8)."""
        L = lib()
        self.features(src)
        check(L.ozl_select_mlp(self.feats.data_ptr(), self.K, params.data_ptr(), self.choice.data_ptr(),
                               self.logits.data_ptr(), self._st()), "ozl_select_mlp")
        return self.choice, self.logits


def pack_chunks(chunks, device="cuda"):
    """Places numpy chunks in one device buffer, each start aligned to 8;
    returns (src, spans, widths, spans as numpy)."""
    spans, p = [], 0
    for c in chunks:
        p += (-p) % 8
        spans += [p, c.nbytes]
        p += c.nbytes
    spans = np.array(spans, np.int64)
    buf = np.zeros(max(p, 1), np.uint8)
    for i, c in enumerate(chunks):
        buf[spans[2 * i]:spans[2 * i] + c.nbytes] = np.ascontiguousarray(c).view(np.uint8)
    widths = np.array([c.dtype.itemsize for c in chunks], np.uint8)
    return (torch.from_numpy(buf).to(device), torch.from_numpy(spans).to(device), torch.from_numpy(widths).to(device),
            spans)
