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
  host buffers  the same steps with the input, frames and output in pinned
                host memory (copies in and out timed, wall clock)
  workloads     sub-records (config.workloads) of the other workloads of the
                path, each timed with its own steps: the 16 MiB timestamps
                frame and the 1 GiB bf16 tensor through the ZL_CCtx engine,
                4,096 batched columns, learned selection over 8,192 chunks
                against exhaustive trial; each with encode / decode GB/s, its
                roofline fraction and traffic and a reference CPU baseline
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
def host_threads():
    """Every core this process may run on (no cap), and what the host has."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, aff), {"nproc": os.cpu_count(), "affinity": aff}


def cpu_baseline(data, threads, budget_s=8.0):
    """The reference library on the host cores: each thread owns one
    compression and one decompression context and cycles over its share of
    the blocks of `data` (FSE and Huffman frame of each block, then
    decompression), until `budget_s` seconds have passed."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    if C.refh() is None:
        return None
    sample_blocks = data.size // BLOCK
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
        "sample": f"{sample_blocks} distinct 64 KiB blocks of the benchmark's own input, cycled; each block "
                  f"compressed with ZL_GRAPH_FSE and with ZL_GRAPH_HUFFMAN and decompressed again by the "
                  f"reference library built from its sources, one context pair per thread; "
                  f"{total / 1e9:.2f} GB (uncompressed, both codings) in {wall:.1f} s wall",
    }


def load_traffic(name):
    """The committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of a workload
    (profiles/<name>_hbm_traffic.json, copy-calibrated bytes per kernel and
    their total over one encode + decode), or None."""
    path = os.path.join(ROOT, "profiles", f"{name}_hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


# ---------------------------------------------------------------- batched columns
def cpu_baseline_columns(cols, threads, budget_s):
    """The reference library compressing and decompressing the columns (graph
    delta_int -> zigzag -> ZL_GRAPH_BITPACK), one context pair per thread,
    cycling over the columns for `budget_s` seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    if C.refh() is None:
        return None
    h = C.refh()
    t_start = time.perf_counter()

    def worker(tid, res):
        nbytes = 0
        i = tid
        tt = ctypes.c_int(0)
        ww = ctypes.c_int(0)
        while time.perf_counter() - t_start < budget_s:
            c = cols[i % len(cols)]
            cap = h.refh_compressBound(c.nbytes) + 4096
            out = np.empty(cap, np.uint8)
            back = np.empty(c.nbytes + 64, np.uint8)
            r = h.refh_compress(C.DELTA_ZZ_BP, C.NUMERIC, c.dtype.itemsize, c.ctypes.data, c.size, out.ctypes.data,
                                cap, 1, 1)
            d = h.refh_decompress(out.ctypes.data, r, back.ctypes.data, back.size, ctypes.byref(tt),
                                  ctypes.byref(ww))
            assert r > 0 and d == c.nbytes
            nbytes += c.nbytes
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
    return {"value": round(total / wall / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
            "sample": f"columns of the same workload cycled, each compressed (delta_int -> zigzag -> "
                      f"ZL_GRAPH_BITPACK) and decompressed by the reference library, one context pair per "
                      f"thread; {total / 1e9:.2f} GB in {wall:.1f} s wall"}


def bench_columns(args):
    """Batched-columns workload: 4,096 numeric columns (log-uniform lengths in
    [1 Ki, 1 Mi], u16/u32, random walks), one frame per column, encode and
    decode each as one captured hipGraph of persistent work-queue launches with
    the autotuner's item size; against the same kernels launched per column."""
    import torch
    from openzl_amd import batch, _lib
    from openzl_amd import datagen as G
    L = _lib.lib()
    print("[bench] generating the 4,096 columns", file=sys.stderr, flush=True)
    cols = G.random_walk_columns()
    spans, p = [], 0
    for c in cols:
        p += (-p) % c.dtype.itemsize
        spans += [p, c.nbytes]
        p += c.nbytes
    buf = np.zeros(p, np.uint8)
    for i, c in enumerate(cols):
        buf[spans[2 * i]:spans[2 * i] + c.nbytes] = c.view(np.uint8)
    U = sum(c.nbytes for c in cols)
    src = torch.from_numpy(buf).cuda()
    spans_t = torch.tensor(spans, dtype=torch.int64, device="cuda")
    widths = torch.tensor([c.dtype.itemsize for c in cols], dtype=torch.uint8, device="cuda")
    K = len(cols)
    cb = batch.ColumnBatch(spans_t, widths)
    print("[bench] autotune", file=sys.stderr, flush=True)
    tile, wg = cb.tune(src)
    dst, frame_off, status = cb.compress(src)
    out = torch.zeros_like(src)
    frames = dst.clone()
    fo = frame_off.clone()
    dstatus = cb.decompress(frames, fo, out)
    torch.cuda.synchronize()
    assert int(status.abs().sum()) == 0 and int(dstatus.abs().sum()) == 0 and torch.equal(out, src)
    C_bytes = int(fo[-1].item())
    ge = cb.graph(0, src, None, None)
    gd = cb.graph(1, frames, fo, out)
    for _ in range(args.warmup):
        batch.ColumnBatch.launch(ge)
        batch.ColumnBatch.launch(gd)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    enc_ms = dec_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record()
        batch.ColumnBatch.launch(ge)
        ev[1].record()
        batch.ColumnBatch.launch(gd)
        ev[2].record()
        ev[2].synchronize()
        enc_ms += ev[0].elapsed_time(ev[1])
        dec_ms += ev[1].elapsed_time(ev[2])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1000 / args.steps
    enc_ms /= args.steps
    dec_ms /= args.steps
    assert torch.equal(out, src)
    # per-column launches: the same kernels, one call per column, per direction
    print("[bench] per-column launches", file=sys.stderr, flush=True)
    maxb = max(spans[1::2])
    one_dst = torch.empty(maxb + 128, dtype=torch.uint8, device="cuda")
    one_off = torch.zeros(2, dtype=torch.int64, device="cuda")
    one_st = torch.zeros(1, dtype=torch.int32, device="cuda")
    lens = fo[1:] - fo[:-1]
    pairs = torch.stack([torch.zeros_like(lens), lens], 1).contiguous()
    foff_h = fo.cpu().numpy()
    st = torch.cuda.current_stream().cuda_stream
    torch.cuda.synchronize()
    tp0 = time.perf_counter()
    for k in range(K):
        L.ozl_columns_compress_batch(src.data_ptr(), spans_t[2 * k:].data_ptr(), widths[k:].data_ptr(), 1,
                                     one_dst.data_ptr(), one_dst.numel(), one_off.data_ptr(), one_st.data_ptr(),
                                     cb.ws.data_ptr(), cb.ws.numel(), tile, wg, st)
    torch.cuda.synchronize()
    tp1 = time.perf_counter()
    out.zero_()
    torch.cuda.synchronize()
    tp1b = time.perf_counter()
    for k in range(K):
        L.ozl_columns_decompress_batch(frames.data_ptr() + int(foff_h[k]), pairs[k].data_ptr(), 1, out.data_ptr(),
                                       spans_t[2 * k:].data_ptr(), one_st.data_ptr(), cb.ws.data_ptr(),
                                       cb.ws.numel(), tile, wg, st)
    torch.cuda.synchronize()
    tp2 = time.perf_counter()
    assert torch.equal(out, src)
    # per-kernel milliseconds of one eager encode + decode (graph replays are
    # not seen by the launch-time event registry)
    prof = {}
    if not args.no_profile:
        L.ozl_profile_begin()
        L.ozl_profile_tag(b"col_enc/")
        cb.compress(src)
        L.ozl_profile_tag(b"col_dec/")
        cb.decompress(frames, fo, out)
        torch.cuda.synchronize()
        bufp = ctypes.create_string_buffer(1 << 16)
        L.ozl_profile_end.restype = ctypes.c_size_t
        L.ozl_profile_end.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.ozl_profile_end(bufp, len(bufp))
        for line in bufp.value.decode().splitlines():
            name, cnt, ms = line.split()
            prof[name] = round(float(ms) / int(cnt), 4)
        L.ozl_profile_tag(b"")
    enc_gbs = (U + C_bytes) / (enc_ms / 1000) / 1e9
    dec_gbs = (U + C_bytes) / (dec_ms / 1000) / 1e9
    both = 2 * (U + C_bytes) / ((enc_ms + dec_ms) / 1000) / 1e9
    traffic = load_traffic("columns_4096")
    roofline = {"bound": "hbm", "kernel": "whole encode + decode graphs (all launches)", "achieved": round(both, 2),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(both / HBM_PEAK_GBS, 4),
                "traffic": traffic.get("total_bytes") if traffic else None, "algorithmic_bytes": 2 * (U + C_bytes),
                "encode_algorithmic_GBps": round(enc_gbs, 2), "decode_algorithmic_GBps": round(dec_gbs, 2),
                "kernel_ms_per_call": prof}
    cpu = None
    if args.cpu_sample > 0:
        threads, hinfo = host_threads()
        cpu = cpu_baseline_columns(cols, threads, args.cpu_seconds)
        if cpu:
            cpu.update(hinfo)
    value = 2 * U / (ms_step / 1000) / 1e9
    rec = {
        "metric": METRIC, "value": round(value, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u16/u32",
        "data": "synthetic (4,096 seeded random-walk columns, step round(N(0, 2^b)), b in 2..14)",
        "config": {
            "workload": "batched columns: 4096 uint16/uint32 columns, lengths log-uniform in [1Ki, 1Mi], "
                        "delta_int -> zigzag -> ZL_GRAPH_BITPACK, one frame per column, format v23, checksums on",
            "columns": K, "input_bytes": U, "frame_bytes": C_bytes, "ratio": round(U / C_bytes, 4),
            "item_deltas": tile, "workgroups_per_cu": wg,
            "encode_ms": round(enc_ms, 4), "decode_ms": round(dec_ms, 4),
            "encode_GBps": round(U / (enc_ms / 1000) / 1e9, 3), "decode_GBps": round(U / (dec_ms / 1000) / 1e9, 3),
            "per_column_launches_encode_ms": round((tp1 - tp0) * 1000, 2),
            "per_column_launches_decode_ms": round((tp2 - tp1b) * 1000, 2),
            "batched_speedup_vs_per_column": round(((tp1 - tp0) + (tp2 - tp1b)) * 1000 / (enc_ms + dec_ms), 2),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return rec


def cpu_baseline_select(chunks, threads, budget_s):
    """The reference library compressing chunks with its brute-force selector
    over the same eight candidate graphs, one context per thread, cycling over
    the chunks for `budget_s` seconds."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    if C.refh() is None:
        return None
    h = C.refh()
    t_start = time.perf_counter()

    def worker(tid, res):
        nbytes = 0
        i = tid
        while time.perf_counter() - t_start < budget_s:
            c = chunks[i % len(chunks)]
            cap = h.refh_compressBound(c.nbytes) + 4096
            out = np.empty(cap, np.uint8)
            r = h.refh_compress(C.SELECT8, C.NUMERIC, c.dtype.itemsize, c.ctypes.data, c.size, out.ctypes.data,
                                cap, 1, 1)
            assert r > 0
            nbytes += c.nbytes
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
    return {"value": round(total / wall / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
            "sample": f"chunks of the same workload cycled, each compressed by the reference's brute-force "
                      f"selector over the eight candidate graphs, one context per thread; {total / 1e9:.3f} GB "
                      f"in {wall:.1f} s wall"}


def bench_select(args):
    """Learned graph selection: 8,192 chunks of 64 KiB from eight generators;
    the MLP's choice (features + MLP + compression with the chosen candidate)
    against exhaustive trial (the brute-force selector over the eight
    candidates, then compression with its choice)."""
    import torch
    from openzl_amd import select as S, _lib
    from openzl_amd import datagen as G
    from openzl_amd import selector_train as T
    L = _lib.lib()
    print("[bench] generating the 8,192 chunks", file=sys.stderr, flush=True)
    gens, chunks = G.selector_chunks(8192, seed=G.SEEDS[5])
    src, spans, widths, _ = S.pack_chunks(chunks)
    U = sum(c.nbytes for c in chunks)
    K = len(chunks)
    sb = S.SelectorBatch(spans, widths, G.SELECT_CHUNK)
    params = S.load_params()

    def exhaustive():
        _, ch = sb.trial(src)
        return sb.compress(src, ch)

    def learned():
        ch, _ = sb.predict(src, params)
        return sb.compress(src, ch)

    sizes, ex_choice = sb.trial(src)
    sizes = sizes.cpu().numpy().astype(np.int64)
    ex_choice = ex_choice.cpu().numpy().copy()
    dst, fo, st = exhaustive()
    torch.cuda.synchronize()
    assert int(st.abs().sum()) == 0
    C_ex = int(fo[-1].item())
    ml_choice = sb.predict(src, params)[0].cpu().numpy().copy()
    dst, fo, st = learned()
    torch.cuda.synchronize()
    assert int(st.abs().sum()) == 0
    C_ml = int(fo[-1].item())
    raw = np.array([c.nbytes for c in chunks], np.int64)
    loss = T.ratio_loss(sizes, raw, ml_choice)
    agree = float(np.mean(ml_choice == ex_choice))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000 / args.steps

    print("[bench] timing", file=sys.stderr, flush=True)
    ms_ml = timed(learned)
    ms_ex = timed(exhaustive)
    # the selection alone: features + MLP against the exhaustive trial's sizes and choice
    ms_sel_ml = timed(lambda: sb.predict(src, params))
    ms_sel_ex = timed(lambda: sb.trial(src))
    prof = {}
    if not args.no_profile:
        L.ozl_profile_begin()
        L.ozl_profile_tag(b"learned/")
        learned()
        L.ozl_profile_tag(b"exhaustive/")
        exhaustive()
        torch.cuda.synchronize()
        bufp = ctypes.create_string_buffer(1 << 16)
        L.ozl_profile_end.restype = ctypes.c_size_t
        L.ozl_profile_end.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.ozl_profile_end(bufp, len(bufp))
        for line in bufp.value.decode().splitlines():
            name, cnt, ms = line.split()
            prof[name] = round(float(ms), 4)
        L.ozl_profile_tag(b"")
    ml_gbs = (U + C_ml) / (ms_ml / 1000) / 1e9
    traffic = load_traffic("select_8192")
    roofline = {"bound": "hbm", "kernel": "whole learned path (features, MLP, compression with the choice)",
                "achieved": round(ml_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ml_gbs / HBM_PEAK_GBS, 4), "traffic": traffic.get("total_bytes") if traffic else None,
                "algorithmic_bytes": U + C_ml, "kernel_ms_per_call": prof}
    cpu = None
    if args.cpu_sample > 0:
        threads, hinfo = host_threads()
        cpu = cpu_baseline_select(chunks, threads, args.cpu_seconds)
        if cpu:
            cpu.update(hinfo)
    rec = {
        "metric": METRIC, "value": round(U / (ms_ml / 1000) / 1e9, 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_ml, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8/u16/u32/u64",
        "data": "synthetic (8,192 x 64 KiB chunks, eight seeded generators chosen uniformly)",
        "config": {
            "workload": "learned graph selection: 8192 x 64 KiB numeric chunks, MLP 40-64-64-8 (bf16 MFMA) over "
                        "eight candidate graphs, against exhaustive trial (brute-force selector); format v23, "
                        "checksums on",
            "chunks": K, "input_bytes": U, "frame_bytes_learned": C_ml, "frame_bytes_exhaustive": C_ex,
            "ratio_learned": round(U / C_ml, 4), "ratio_exhaustive": round(U / C_ex, 4),
            "ratio_loss_vs_exhaustive": round(loss, 7), "choice_agreement": round(agree, 4),
            "learned_ms": round(ms_ml, 4), "exhaustive_ms": round(ms_ex, 4),
            "learned_GBps": round(U / (ms_ml / 1000) / 1e9, 3), "exhaustive_GBps": round(U / (ms_ex / 1000) / 1e9, 3),
            "learned_speedup_vs_exhaustive": round(ms_ex / ms_ml, 3),
            "selection_ms_learned": round(ms_sel_ml, 4), "selection_ms_exhaustive": round(ms_sel_ex, 4),
            "selection_speedup": round(ms_sel_ex / ms_sel_ml, 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return rec


def bench_engine(args):
    """Workloads through the ZL_CCtx / ZL_DCtx engine, one batch call per
    direction (ozl_CCtx_compressBatch / ozl_DCtx_decompressBatch, which write
    the same frames as ZL_CCtx_compressTypedRef / read them as
    ZL_DCtx_decompressTyped): timestamps: 4,194,304 u32 (16 MiB, one frame),
    DELTA_INT -> transpose_split -> ZL_GRAPH_FSE per plane; bf16: 536,870,912
    bf16 N(0, 0.02) values (1 GiB) in 64 frames of 16 MiB, transpose_split ->
    plane 1 ZL_GRAPH_HUFFMAN, plane 0 ZL_GRAPH_FSE. value: inputs, frames and
    outputs resident in HBM; the same calls on host buffers (PCIe included)
    are reported beside it."""
    import torch
    from openzl_amd import api
    from openzl_amd import datagen as G
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    segment = None
    if args.workload == "timestamps":
        x = G.timestamps_u32()
        parts = [x]
        kind, typ, width = api.DELTA_TS_FSE, api.NUMERIC, 4
        workload = ("timestamps: 4,194,304 u32 (start 1.7e9, Poisson(1000) increments, mod 2^32), one frame, "
                    "DELTA_INT -> transpose_split -> ZL_GRAPH_FSE per plane, format v23, checksums on")
        dtype = "u32"
    elif args.workload == "bf16":
        w = G.bf16_normal()
        seg = 8388608
        parts = [w[a:a + seg] for a in range(0, w.size, seg)]
        kind, typ, width = api.TS_HUF_FSE, api.STRUCT, 2
        workload = ("bf16 weights: 536,870,912 bf16 N(0, 0.02) in 64 frames of 16 MiB, transpose_split -> plane 1 "
                    "ZL_GRAPH_HUFFMAN, plane 0 ZL_GRAPH_FSE, format v23, checksums on")
        dtype = "bf16"
    else:
        # one frame: a segmenter cuts 64 chunks of 16 MiB, each through the same graph
        parts = [G.bf16_normal()]
        kind, typ, width = api.SEGMENTED, api.STRUCT, 2
        segment = (api.TS_HUF_FSE, 8388608)
        workload = ("bf16 weights: 536,870,912 bf16 N(0, 0.02) as one frame of 64 chunks of 16 MiB (fixed-size "
                    "segmenter), transpose_split -> plane 1 ZL_GRAPH_HUFFMAN, plane 0 ZL_GRAPH_FSE, format v23, "
                    "checksums on")
        dtype = "bf16"
    U = sum(p.nbytes for p in parts)
    comp = api.Compressor(kind, segment=segment)
    dec = api.Decompressor()
    # host buffers (PCIe included): the reference frames' check and the side figure
    frames = comp.compress_batch(parts, typ, width)
    Cb = sum(len(f) for f in frames)
    outs = dec.decompress_batch(frames, U + 64)
    assert all(o == p.tobytes() for o, p in zip(outs, parts))
    # device-resident buffers
    din = torch.from_numpy(np.concatenate([p.view(np.uint8) for p in parts])).cuda()
    ptrs = [din.data_ptr() + sum(q.nbytes for q in parts[:i]) for i in range(len(parts))]
    sizes_in = [p.nbytes for p in parts]
    cap = Cb + 4096 * len(parts)
    dfr = torch.empty(cap, dtype=torch.uint8, device="cuda")
    dout = torch.empty(U + 64, dtype=torch.uint8, device="cuda")
    fsz = comp.compress_ptrs(ptrs, sizes_in, typ, width, dfr.data_ptr(), cap)
    assert dfr[:Cb].cpu().numpy().tobytes() == b"".join(frames)
    dec.decompress_ptrs(dfr.data_ptr(), fsz, dout.data_ptr(), U + 64)
    assert torch.equal(dout[:U], din)
    for _ in range(args.warmup):
        comp.compress_ptrs(ptrs, sizes_in, typ, width, dfr.data_ptr(), cap)
        dec.decompress_ptrs(dfr.data_ptr(), fsz, dout.data_ptr(), U + 64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        comp.compress_ptrs(ptrs, sizes_in, typ, width, dfr.data_ptr(), cap)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        dec.decompress_ptrs(dfr.data_ptr(), fsz, dout.data_ptr(), U + 64)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    enc_ms = (t1 - t0) * 1000 / args.steps
    dec_ms = (t2 - t1) * 1000 / args.steps
    # per-kernel milliseconds of one encode + decode (launch-time HIP events)
    prof = {}
    if not args.no_profile:
        from openzl_amd import _lib
        Lk = _lib.lib()
        Lk.ozl_profile_begin()
        Lk.ozl_profile_tag(b"enc/")
        comp.compress_ptrs(ptrs, sizes_in, typ, width, dfr.data_ptr(), cap)
        Lk.ozl_profile_tag(b"dec/")
        dec.decompress_ptrs(dfr.data_ptr(), fsz, dout.data_ptr(), U + 64)
        torch.cuda.synchronize()
        bufp = ctypes.create_string_buffer(1 << 16)
        Lk.ozl_profile_end.restype = ctypes.c_size_t
        Lk.ozl_profile_end.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        Lk.ozl_profile_end(bufp, len(bufp))
        for line in bufp.value.decode().splitlines():
            name, cnt, kms = line.split()
            prof[name] = (int(cnt), float(kms))
        Lk.ozl_profile_tag(b"")
    # host buffers: one step each way
    h0 = time.perf_counter()
    comp.compress_batch(parts, typ, width)
    h1 = time.perf_counter()
    dec.decompress_batch(frames, U + 64)
    h2 = time.perf_counter()
    host_ms = (h2 - h0) * 1000
    cpu = None
    if args.cpu_sample > 0 and C.refh() is not None:
        # the reference library on the same frames, one context per thread:
        # the bf16 frames on all host cores; the single timestamps frame on one
        # thread (the reference codes one frame on one thread, so no number of
        # cores makes one frame faster), with the throughput of all cores each
        # coding its own copy reported beside it
        nproc, hinfo = host_threads()
        ctyp = {api.SERIAL: C.SERIAL, api.STRUCT: C.STRUCT, api.NUMERIC: C.NUMERIC}[typ]
        if segment:
            C.ref_set_segmenter(*segment)
        assert frames[0] == C.ref_compress(kind, parts[0], ctyp, width), "frame differs from the reference's"

        def run_ref(threads, budget):
            res = [0] * threads
            t_start = time.perf_counter()

            def worker(tid):
                i, nbytes = tid, 0
                while time.perf_counter() - t_start < budget:
                    p = parts[i % len(parts)]
                    f = C.ref_compress(kind, p, ctyp, width)
                    C.ref_decompress(f, p.nbytes + 64)
                    nbytes += p.nbytes
                    i += threads
                res[tid] = nbytes
            ths = [threading.Thread(target=worker, args=(t,)) for t in range(threads)]
            for t in ths:
                t.start()
            for t in ths:
                t.join()
            wall = time.perf_counter() - t_start
            return sum(res), wall

        single = len(parts) == 1
        threads = 1 if single else nproc
        nbytes, wall = run_ref(threads, args.cpu_seconds)
        cpu = {"value": round(nbytes / wall / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
               "sample": f"the same frames cycled, compressed and decompressed by the reference library "
                         f"({threads} thread(s)); {nbytes / 1e9:.3f} GB in {wall:.1f} s wall", **hinfo}
        if single and nproc > 1:
            nb_all, wall_all = run_ref(nproc, args.cpu_seconds / 2)
            cpu["all_cores_independent_frames_GBps"] = round(nb_all / wall_all / 1e9, 4)
    ms = enc_ms + dec_ms
    both = 2 * (U + Cb) / (ms / 1000) / 1e9
    traffic = load_traffic({"timestamps": "engine_timestamps_16MiB", "bf16": "engine_bf16_1GiB"}.get(args.workload, ""))
    dom = max(prof.items(), key=lambda kv: kv[1][1]) if prof else None
    rec = {
        "metric": METRIC, "value": round(U / (ms / 1000) / 1e9, 3), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded generator)",
        "config": {"workload": workload, "input_bytes": U, "frame_bytes": Cb, "ratio": round(U / Cb, 4),
                   "frames": len(frames), "encode_ms": round(enc_ms, 3), "decode_ms": round(dec_ms, 3),
                   "encode_GBps": round(U / (enc_ms / 1000) / 1e9, 3),
                   "decode_GBps": round(U / (dec_ms / 1000) / 1e9, 3),
                   "buffers": "inputs, frames and outputs in HBM (one batch call per direction)",
                   "host_buffers_GBps": round(U / (host_ms / 1000) / 1e9, 3),
                   "host_buffers_ms": round(host_ms, 3)},
        "roofline": {"bound": "hbm", "kernel": "whole engine calls (host-orchestrated graph walk, device-resident)",
                     "achieved": round(both, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(both / HBM_PEAK_GBS, 4),
                     "traffic": traffic.get("total_bytes") if traffic else None,
                     "algorithmic_bytes": 2 * (U + Cb),
                     "dominant_kernel": dom[0] if dom else None,
                     "kernel_ms_per_call": {k: round(v[1], 4) for k, v in sorted(prof.items())}},
        "cpu_baseline": cpu,
    }
    return rec


# ---------------------------------------------------------------- main
def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--blocks", type=int, default=N_BLOCKS)
    ap.add_argument("--cpu-sample", type=int, default=N_BLOCKS,
                    help="distinct blocks of the input in the CPU baseline sample (0 = skip)")
    ap.add_argument("--no-configs", action="store_true",
                    help="the entropy stress line only (no sub-records of the other workloads)")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="wall-clock budget of the CPU baseline")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run 3 device-to-device copies of 1 GiB (HBM counter calibration) and exit")
    ap.add_argument("--workload", choices=["entropy", "columns", "select", "timestamps", "bf16", "bf16seg"],
                    default="entropy",
                    help="entropy: the 16,384-block entropy stress (the headline line); columns: the batched "
                         "4,096-column delta/zigzag/bitpack workload; select: learned graph selection over "
                         "8,192 chunks against exhaustive trial; timestamps / bf16 / bf16seg: workloads through the "
                         "ZL_CCtx engine (bf16seg: the 1 GiB tensor as one segmented frame)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (asked for --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "columns":
        return print(json.dumps(bench_columns(args)))
    if args.workload == "select":
        return print(json.dumps(bench_select(args)))
    if args.workload in ("timestamps", "bf16", "bf16seg"):
        return print(json.dumps(bench_engine(args)))
    import torch
    from openzl_amd import batch, _lib

    L = _lib.lib()
    if args.copy_probe:
        a = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print("[bench] copy probe: 3 x 1 GiB device copies", file=sys.stderr, flush=True)
        return
    B = args.blocks
    dev = "cuda"
    print(f"[bench] generating {B} x 64 KiB Zipf(1.1) blocks", file=sys.stderr, flush=True)
    src = gen_blocks_torch(torch, B)
    off = torch.arange(0, (B + 1) * BLOCK, BLOCK, dtype=torch.int64, device=dev)
    enc = {m: batch.EntropyBatchCompressor(B, BLOCK) for m in (batch.FSE, batch.HUFFMAN)}
    dec = {m: batch.EntropyBatchDecompressor(B) for m in (batch.FSE, batch.HUFFMAN)}
    out = {m: torch.empty(B * BLOCK, dtype=torch.uint8, device=dev) for m in (batch.FSE, batch.HUFFMAN)}
    torch.cuda.synchronize()
    tag = {batch.FSE: "fse", batch.HUFFMAN: "huf"}
    L.ozl_profile_tag.argtypes = [ctypes.c_char_p]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    phase_ms = {"encode": 0.0, "decode": 0.0}

    def step(timed=False):
        frames = {}
        if timed:
            ev[0].record()
        for m in (batch.FSE, batch.HUFFMAN):
            L.ozl_profile_tag(f"{tag[m]}_enc/".encode())
            frames[m] = enc[m].run(m, src, off)
        if timed:
            ev[1].record()
        for m in (batch.FSE, batch.HUFFMAN):
            L.ozl_profile_tag(f"{tag[m]}_dec/".encode())
            dst, foff, _ = frames[m]
            dec[m].run(dst, foff, out[m], off)
        if timed:
            ev[2].record()
            ev[2].synchronize()
            phase_ms["encode"] += ev[0].elapsed_time(ev[1])
            phase_ms["decode"] += ev[1].elapsed_time(ev[2])
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

    # encode / decode split, measured in separate steps with events between the phases
    for _ in range(args.steps):
        step(timed=True)
    enc_ms = phase_ms["encode"] / args.steps
    dec_ms = phase_ms["decode"] / args.steps

    # roofline of the dominant kernel: algorithmic bytes per launch = U + C of
    # the launch's coding (every launch processes the whole batch of blocks)
    roofline = None
    if prof:
        dom = max(prof.items(), key=lambda kv: kv[1][1])
        name, (cnt, tot_ms) = dom
        avg_ms = tot_ms / cnt
        coding = batch.FSE if name.startswith("fse") else batch.HUFFMAN
        per_launch = U + csize[coding]
        achieved = per_launch / (avg_ms / 1000) / 1e9
        # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
        # summary of this workload (copy-calibrated), when it covers the kernel
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "entropy_stress_1GiB_hbm_traffic.json")
        if B == N_BLOCKS and os.path.exists(tpath):
            with open(tpath) as f:
                traffic = json.load(f)["kernels"].get(name, {}).get("traffic_bytes")
        roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "avg_launch_ms": round(avg_ms, 4), "bytes_per_launch": per_launch,
                    "kernel_ms_per_step": {k: round(v[1] / args.steps, 4) for k, v in sorted(prof.items())}}

    # host buffers (not `value`): the input in pinned host memory, copied in,
    # encoded, the frames copied out (their size read back first); then the
    # frames copied in, decoded and the output copied out; wall clock
    print("[bench] host-buffer steps (PCIe included)", file=sys.stderr, flush=True)
    pin_in = torch.empty(U, dtype=torch.uint8, pin_memory=True)
    pin_in.copy_(src)
    pin_fr = {m: torch.empty(csize[m] + (1 << 20), dtype=torch.uint8, pin_memory=True) for m in enc}
    pin_out = torch.empty(U, dtype=torch.uint8, pin_memory=True)
    d_in = torch.empty_like(src)
    d_fr = {m: torch.empty(csize[m] + (1 << 20), dtype=torch.uint8, device=dev) for m in enc}
    host_ms = {"encode": 0.0, "decode": 0.0}
    nh = 2
    for it in range(nh + 1):
        torch.cuda.synchronize()
        ta = time.perf_counter()
        d_in.copy_(pin_in, non_blocking=True)
        fr_h = {}
        for m in (batch.FSE, batch.HUFFMAN):
            dst, foff, _ = enc[m].run(m, d_in, off)
            nbytes = int(foff[B].item())
            pin_fr[m][:nbytes].copy_(dst[:nbytes], non_blocking=True)
            fr_h[m] = (nbytes, foff)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        for m in (batch.FSE, batch.HUFFMAN):
            nbytes, foff = fr_h[m]
            d_fr[m][:nbytes].copy_(pin_fr[m][:nbytes], non_blocking=True)
            dec[m].run(d_fr[m], foff, out[m], off)
            pin_out.copy_(out[m], non_blocking=True)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        if it:   # the first pass warms the pinned buffers
            host_ms["encode"] += (tb - ta) * 1000 / nh
            host_ms["decode"] += (tc - tb) * 1000 / nh
    assert torch.equal(out[batch.HUFFMAN], src) and bool(torch.equal(pin_out, pin_in))
    host_rate = 2 * U / ((host_ms["encode"] + host_ms["decode"]) / 1000) / 1e9
    del pin_fr, d_fr, d_in

    cpu = None
    if args.cpu_sample > 0:
        threads, hinfo = host_threads()
        nb = min(args.cpu_sample, B)
        print(f"[bench] CPU baseline on {threads} threads over {nb} blocks", file=sys.stderr, flush=True)
        cpu = cpu_baseline(pin_in[:nb * BLOCK].numpy(), threads, args.cpu_seconds)
        if cpu:
            cpu.update(hinfo)
    del pin_in, pin_out

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
            "symbols_per_s_both_codings": round(2 * U / (ms_step / 1000), 1),
            "encode_GBps": round(2 * U / (enc_ms / 1000) / 1e9, 3),
            "decode_GBps": round(2 * U / (dec_ms / 1000) / 1e9, 3),
            "encode_ms_per_step": round(enc_ms, 3), "decode_ms_per_step": round(dec_ms, 3),
            "host_buffers_GBps": round(host_rate, 3),
            "host_buffers_encode_ms": round(host_ms["encode"], 3),
            "host_buffers_decode_ms": round(host_ms["decode"], 3),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if not args.no_configs:
        del src, out, enc, dec, frames
        torch.cuda.empty_cache()
        rec["config"]["workloads"] = other_workloads(args)
    print(json.dumps(rec))


def summary(sub):
    """The figures of a workload's own record kept in the headline line."""
    c = sub["config"]
    keep = {k: v for k, v in c.items() if k.endswith(("GBps", "_ms", "ratio", "speedup", "agreement"))
            or k.startswith(("ratio", "learned_", "exhaustive_", "selection_", "frame_bytes", "input_bytes"))}
    r = sub.get("roofline") or {}
    cpu = sub.get("cpu_baseline")
    out = {"workload": c["workload"], "value": sub["value"], "unit": sub["unit"], "ms_per_step": sub["ms_per_step"],
           "steps": sub["steps"], "dtype": sub["dtype"], **keep,
           "roofline": {k: r.get(k) for k in ("bound", "kernel", "dominant_kernel", "achieved", "peak", "unit",
                                              "frac", "traffic", "algorithmic_bytes") if k in r},
           "cpu_baseline": cpu,
           "vs_cpu": round(sub["value"] / cpu["value"], 2) if cpu and cpu.get("value") else None}
    return out


def other_workloads(args):
    """The other workloads of the hot path, each timed by its own function
    with fewer steps: the 16 MiB timestamps frame and the 1 GiB bf16 tensor
    through the ZL_CCtx engine, the 4,096 batched columns, and learned
    selection over 8,192 chunks against exhaustive trial."""
    import copy
    import torch
    res = {}
    for name, fn, wl in (("timestamps_16MiB", bench_engine, "timestamps"), ("columns_4096", bench_columns, None),
                         ("bf16_1GiB", bench_engine, "bf16"), ("select_8192", bench_select, None)):
        a = copy.copy(args)
        a.steps = max(1, min(args.steps, 5))
        a.warmup = 1
        if wl:
            a.workload = wl
        print(f"[bench] workload {name}", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        try:
            res[name] = summary(fn(a))
        except Exception as e:   # the headline line is still printed
            res[name] = {"error": f"{type(e).__name__}: {e}"}
        res[name]["wall_s"] = round(time.perf_counter() - t0, 1)
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    return res


if __name__ == "__main__":
    main()
