#!/usr/bin/env python3
"""Entropy-stage benchmark of the OpenZL codec hot path on one MI355X.

Workload (the entropy-only stress configuration): 1 GiB of uint8 symbols drawn
from Zipf(s=1.1) over 256 symbols (ranks permuted by the seed), cut into
16,384 blocks of 64 KiB. Every block is compressed into its own frame twice,
with ZL_GRAPH_FSE and with ZL_GRAPH_HUFFMAN, and every frame is decompressed
again; each coding is one batched call for all blocks (libozl_hip.so C ABI).

One step = FSE encode + Huffman encode + FSE decode + Huffman decode of the
whole 1 GiB batch, inputs resident in HBM. value = (2 x 1 GiB) / step time.

Also reported:
  roofline      the kernel with the largest share of the timed region, its
                algorithmic bytes per launch (input + payload bytes of its
                blocks) over its mean duration (HIP events around every launch
                during the timed steps), against 8 TB/s
  cpu_baseline  the reference OpenZL library (oracle/_ref, compiled from the
                reference sources) compressing and decompressing a bounded
                sample of the same blocks with both graphs, one context per
                thread on all host cores
"""
import argparse
import ctypes
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "encode+decode GB/s on 1 MI355X (% HBM roofline) vs CPU; ratio loss vs exhaustive"
HBM_PEAK_GBS = 8000.0
BLOCK = 65536
N_BLOCKS = 16384
SEED = 0x5EED0002


# ---------------------------------------------------------------- data
from openzl_amd.datagen import zipf_blocks_np as gen_blocks_np, zipf_blocks_torch as gen_blocks_torch  # noqa: E402


# ---------------------------------------------------------------- CPU baseline
def cpu_baseline(sample_blocks, threads, budget_s=8.0):
    """The reference library on the host cores: each thread owns one
    compression and one decompression context and cycles over its share of
    `sample_blocks` blocks of the same workload (FSE and Huffman frame of each
    block, then decompression), until `budget_s` seconds have passed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    if C.refh() is None:
        return None
    data = gen_blocks_np(0, sample_blocks)
    blocks = [data[i * BLOCK:(i + 1) * BLOCK] for i in range(sample_blocks)]
    h = C.refh()
    cap = h.refh_compressBound(BLOCK) + 4096
    t_start = time.perf_counter()

    def worker(tid, res):
        out = np.empty(cap, np.uint8)
        back = np.empty(BLOCK + 64, np.uint8)
        nbytes = 0
        tt = ctypes.c_int(0)
        ww = ctypes.c_int(0)
        i = tid
        while time.perf_counter() - t_start < budget_s:
            b = blocks[i % sample_blocks]
            for kind in (C.FSE, C.HUFFMAN):
                r = h.refh_compress(kind, 0, 1, b.ctypes.data, BLOCK, out.ctypes.data, cap, 1, 1)
                d = h.refh_decompress(out.ctypes.data, r, back.ctypes.data, back.size, ctypes.byref(tt),
                                      ctypes.byref(ww))
                assert r > 0 and d == BLOCK
                nbytes += BLOCK
            i += threads
        res[tid] = nbytes

    res = [0] * threads
    ths = [threading.Thread(target=worker, args=(t, res)) for t in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t_start
    total = sum(res)
    return {
        "value": round(total / wall / 1e9, 4),
        "unit": "GB/s",
        "cores": threads,
        "kind": "reference",
        "sample": f"{sample_blocks} distinct 64 KiB blocks of the same Zipf(1.1) input, cycled; each block "
                  f"compressed with ZL_GRAPH_FSE and with ZL_GRAPH_HUFFMAN and decompressed again by the "
                  f"reference library built from its sources, one context pair per thread; "
                  f"{total / 1e9:.2f} GB (uncompressed, both codings) in {wall:.1f} s wall",
    }


# ---------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--blocks", type=int, default=N_BLOCKS)
    ap.add_argument("--cpu-sample", type=int, default=1024, help="distinct blocks in the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="wall-clock budget of the CPU baseline")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from openzl_amd import batch, _lib

    L = _lib.lib()
    B = args.blocks
    dev = "cuda"
    print(f"[bench] generating {B} x 64 KiB Zipf(1.1) blocks", file=sys.stderr, flush=True)
    src = gen_blocks_torch(torch, B)
    off = torch.arange(0, (B + 1) * BLOCK, BLOCK, dtype=torch.int64, device=dev)
    enc = {m: batch.EntropyBatchCompressor(B, BLOCK) for m in (batch.FSE, batch.HUFFMAN)}
    dec = {m: batch.EntropyBatchDecompressor(B) for m in (batch.FSE, batch.HUFFMAN)}
    out = {m: torch.empty(B * BLOCK, dtype=torch.uint8, device=dev) for m in (batch.FSE, batch.HUFFMAN)}
    torch.cuda.synchronize()

    def step():
        frames = {}
        for m in (batch.FSE, batch.HUFFMAN):
            frames[m] = enc[m].run(m, src, off)
        for m in (batch.FSE, batch.HUFFMAN):
            dst, foff, _ = frames[m]
            dec[m].run(dst, foff, out[m], off)
        return frames

    print("[bench] correctness pass", file=sys.stderr, flush=True)
    frames = step()
    torch.cuda.synchronize()
    csize = {}
    for m in (batch.FSE, batch.HUFFMAN):
        dst, foff, st = frames[m]
        assert int(st.abs().sum()) == 0, "compression reported errors"
        assert int(dec[m].status[:B].abs().sum()) == 0, "decompression reported errors"
        assert torch.equal(out[m], src), "round trip mismatch"
        csize[m] = int(foff[B].item())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    if not args.no_profile:
        L.ozl_profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = {}
    if not args.no_profile:
        buf = ctypes.create_string_buffer(1 << 16)
        L.ozl_profile_end.restype = ctypes.c_size_t
        L.ozl_profile_end.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.ozl_profile_end(buf, len(buf))
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            prof[name] = (int(cnt), float(ms))
    ms_step = (t1 - t0) * 1000 / args.steps
    U = B * BLOCK
    value = 2 * U / (ms_step / 1000) / 1e9

    # payload bytes per coding (frames minus per-frame headers ~ same order; use frame bytes)
    roofline = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][1])
        name, (cnt, tot_ms) = dom
        avg_ms = tot_ms / cnt
        per_launch = {
            "k_fse_encode": U + csize[batch.FSE], "k_fse_decode": U + csize[batch.FSE],
            "k_huf_encode": U + csize[batch.HUFFMAN], "k_huf_decode": U + csize[batch.HUFFMAN],
            "k_plan": U, "k_content_hash": U, "k_dverify": U + (csize[0] + csize[1]) // 2,
            "k_frame_write": (csize[0] + csize[1]) // 2, "k_frame_checksum": (csize[0] + csize[1]) // 2,
            "k_pack_decode": 0, "k_pack_encode": 0, "k_dparse": 0,
        }.get(name, U)
        achieved = per_launch / (avg_ms / 1000) / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": per_launch,
                    "kernel_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in sorted(prof.items())}}

    cpu = None
    if args.cpu_sample > 0:
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = max(1, min(threads, 16))
        print(f"[bench] CPU baseline on {threads} threads", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_sample, threads, args.cpu_seconds)

    rec = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded Zipf(1.1) over 256 symbols, ranks permuted)",
        "config": {
            "workload": "entropy stress: 16384 x 64 KiB uint8 Zipf(1.1) blocks, each framed with ZL_GRAPH_FSE "
                        "and ZL_GRAPH_HUFFMAN, encode+decode, format v23, checksums on",
            "blocks": B, "block_bytes": BLOCK, "input_bytes_per_coding": U,
            "frame_bytes_fse": csize[batch.FSE], "frame_bytes_huffman": csize[batch.HUFFMAN],
            "ratio_fse": round(U / csize[batch.FSE], 4), "ratio_huffman": round(U / csize[batch.HUFFMAN], 4),
            "symbols_per_s": round(2 * U / (ms_step / 1000), 1),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
