#!/usr/bin/env python3
"""Writes the golden fixtures of the codec hot path from the reference OpenZL
library compiled from its own sources (oracle/_ref/libref_harness.so, built by
oracle/ref.mk). Run in a checkout that has the reference sources; the GPU tests
only read what this script wrote.

    python tests/golden/make_golden.py [--full]

Outputs (all under tests/golden/):
  small.json            one entry per small case: graph kind, input type and
                        width, the generator call that rebuilds the input, the
                        input's sha256, and the reference frame's size and sha256
  small/<name>.zl       the reference frame of each small case
  full.json             full-size workloads (with --full; --select-full adds
                        only the 8,192-chunk selector batch): per-frame sizes and
                        sha256 prefixes of the reference frames, plus the
                        sha256 of all frames concatenated in order
  full/<name>.npy       per-frame (size, 8-byte sha256 prefix as u64) tables

Graph kinds are those of oracle/ref_harness.c (STORE .. DELTA_BP); types are
0 serial, 1 struct, 2 numeric.
"""
import argparse
import hashlib
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import ozl_checkers as C  # noqa: E402
from openzl_amd import datagen as G  # noqa: E402


def small_cases():
    """(name, kind, type, width, generator name, generator kwargs, view dtype)."""
    out = []
    for n in (1, 5, 100, 1000, 30000, 130000):
        out.append((f"ts_delta_ts_fse_{n}", C.DELTA_TS_FSE, 2, 4, "timestamps_u32", {"n": n}, None))
    for n, w, b in ((16, 2, 2), (3000, 4, 9), (40000, 2, 14), (60000, 4, 5)):
        out.append((f"col_delta_zz_bp_{n}_{w}_{b}", C.DELTA_ZZ_BP, 2, w, "walk_column",
                    {"n": n, "width": w, "b": b, "seed": n}, None))
    for n in (3000, 100000):
        out.append((f"bf16_ts_huf_fse_{n}", C.TS_HUF_FSE, 1, 2, "bf16_normal", {"n": n}, None))
    for n in (10, 17, 64, 500, 1001, 5000, 65536, 99999, 100000, 150000):
        out.append((f"zipf_fse_{n}", C.FSE, 0, 1, "zipf_bytes", {"n": n, "s": 1.1, "seed": n}, None))
        out.append((f"zipf_huf_{n}", C.HUFFMAN, 0, 1, "zipf_bytes", {"n": n, "s": 1.1, "seed": n}, None))
    for n in (64, 1000, 65536):
        for s in (0.0, 2.0, 4.0):
            for kind, kn in ((C.FSE, "fse"), (C.HUFFMAN, "huf")):
                out.append((f"zipf{s}_{kn}_{n}", kind, 0, 1, "zipf_bytes",
                            {"n": n, "s": s, "card": 1 + (n % 255), "seed": 7 * n}, None))
    out.append(("blocks_fse_0", C.FSE, 0, 1, "zipf_blocks_np", {"first_block": 0, "n_blocks": 1}, None))
    out.append(("blocks_huf_0", C.HUFFMAN, 0, 1, "zipf_blocks_np", {"first_block": 0, "n_blocks": 1}, None))
    out.append(("blocks_fse_9999", C.FSE, 0, 1, "zipf_blocks_np", {"first_block": 9999, "n_blocks": 1}, None))
    out.append(("blocks_huf_9999", C.HUFFMAN, 0, 1, "zipf_blocks_np", {"first_block": 9999, "n_blocks": 1}, None))
    out.append(("const_fse_200000", C.FSE, 0, 1, "uniform_ints", {"n": 200000, "hi": 1, "dtype": "uint8"}, None))
    out.append(("rand_fse_120000", C.FSE, 0, 1, "uniform_ints", {"n": 120000, "hi": 256, "dtype": "uint8"}, None))
    out.append(("twosym_huf_70000", C.HUFFMAN, 0, 1, "uniform_ints", {"n": 70000, "hi": 2, "dtype": "uint8"}, None))
    out.append(("bitpack_num_5000", C.BITPACK, 2, 2, "uniform_ints", {"n": 5000, "hi": 1000, "dtype": "uint16"}, None))
    out.append(("bitpack_ser_5000", C.BITPACK, 0, 1, "uniform_ints", {"n": 5000, "hi": 7, "dtype": "uint8"}, None))
    out.append(("store_ser_3000", C.STORE, 0, 1, "uniform_ints", {"n": 3000, "hi": 256, "dtype": "uint8"}, None))
    out.append(("store_num_3000", C.STORE, 2, 4, "uniform_ints", {"n": 3000, "hi": 1 << 30, "dtype": "uint32"}, None))
    out.append(("ts_fse_40000", C.TS_FSE, 1, 4, "uniform_ints", {"n": 40000, "hi": 60, "dtype": "uint32"}, None))
    out.append(("ts_huf_30000", C.TS_HUF, 1, 2, "uniform_ints", {"n": 30000, "hi": 3000, "dtype": "uint16"}, None))
    out.append(("delta_fse_30000", C.DELTA_FSE, 2, 4, "timestamps_u32", {"n": 30000, "seed": 99}, None))
    out.append(("zz_bp_9000", C.ZZ_BP, 2, 2, "walk_column", {"n": 9000, "width": 2, "b": 6, "seed": 3}, None))
    out.append(("delta_bp_9000", C.DELTA_BP, 2, 4, "walk_column", {"n": 9000, "width": 4, "b": 3, "seed": 4}, None))
    return out


def entropy2_cases():
    """ZL_GRAPH_HUFFMAN on 2-byte inputs (struct, and numeric through the
    num->token conversion), one case per branch of the reference's 2-byte
    choice (encode_entropy_binding.c:612-682): constant, bitpack, store,
    tokenize (alphabet bitpacked or, at 16 bits, stored; indices through the
    1-byte Huffman graph), huffman_struct_v2 with 4 streams and with 1,
    height-limited and wide (20,000-symbol) trees, and chunked inputs whose
    chunks take different branches; ZL_GRAPH_FSE on a struct input (converted
    to serial) for contrast."""
    out = []
    H = C.HUFFMAN
    for typ, t in ((1, "s"), (2, "n")):
        out.append((f"e2{t}_const_5000", H, typ, 2, "uniform_ints", {"n": 5000, "hi": 1, "dtype": "uint16"}))
        out.append((f"e2{t}_bp_5000", H, typ, 2, "uniform_ints", {"n": 5000, "hi": 512, "dtype": "uint16", "seed": 5}))
        out.append((f"e2{t}_rand_5000", H, typ, 2, "uniform_ints", {"n": 5000, "hi": 65536, "dtype": "uint16", "seed": 6}))
        out.append((f"e2{t}_tok_5000", H, typ, 2, "u16_spread", {"n": 5000, "card": 100, "hi": 4000, "seed": 1}))
        out.append((f"e2{t}_h16_30000", H, typ, 2, "u16_spread", {"n": 30000, "card": 1000, "hi": 10000, "seed": 3}))
    out.append(("e2s_tok16_40000", H, 1, 2, "u16_spread", {"n": 40000, "card": 100, "hi": 65536, "seed": 2}))
    out.append(("e2s_h16x1_1000", H, 1, 2, "u16_spread", {"n": 1000, "card": 300, "hi": 512, "s": 1.0, "seed": 6}))
    out.append(("e2s_h16deep_32768", H, 1, 2, "u16_spread",
                {"n": 32768, "card": 300, "hi": 300, "s": 2.0, "seed": 8, "perm": False}))
    out.append(("e2s_h16fib", H, 1, 2, "u16_fib", {"scale": 2000}))
    out.append(("e2s_h16wide_99999", H, 1, 2, "u16_spread", {"n": 99999, "card": 20000, "hi": 65536, "s": 1.2, "seed": 9}))
    out.append(("e2s_h16mid_60000", H, 1, 2, "u16_spread", {"n": 60000, "card": 2000, "hi": 2000, "s": 1.6, "seed": 9,
                                                          "perm": False}))
    out.append(("e2s_mixed_200000", H, 1, 2, "u16_mixed", {"seed": 1}))
    out.append(("e2n_mixed_200000", H, 2, 2, "u16_mixed", {"seed": 2}))
    out.append(("e2s_chunks_131073", H, 1, 2, "u16_spread", {"n": 131073, "card": 900, "hi": 30000, "seed": 4}))
    out.append(("e2s_five_5", H, 1, 2, "u16_spread", {"n": 5, "card": 3, "hi": 65536, "seed": 5}))
    out.append(("e2s_two_200", H, 1, 2, "u16_spread", {"n": 200, "card": 2, "hi": 65536, "seed": 7}))
    out.append(("e2s_fse_5000", C.FSE, 1, 2, "u16_spread", {"n": 5000, "card": 1000, "hi": 10000, "seed": 3}))
    return out


def param_cases():
    """Cases compressed with compressor parameters: (name, kind, type, width,
    generator, kwargs, params). minStreamSize 0 is the default, 5 and 64 move
    the automatic store, -1 turns it off (include/openzl/zl_compress.h:117-121);
    the skewed 8-symbol Huffman input has an 8-byte weights stream, stored
    below 10 and FSE-coded at 5 or with the store off."""
    out = []
    for ms in (0, 5, -1, 64):
        tag = "m" if ms < 0 else ""
        p = {"minStreamSize": ms}
        out.append((f"ms{tag}{abs(ms)}_skew8_huf_3000", C.HUFFMAN, 0, 1, "zipf_bytes",
                    {"n": 3000, "s": 2.0, "card": 8, "seed": 11}, p))
        out.append((f"ms{tag}{abs(ms)}_skew8_huf_40", C.HUFFMAN, 0, 1, "zipf_bytes",
                    {"n": 40, "s": 2.0, "card": 8, "seed": 12}, p))
        for n in (1, 5, 17, 500):
            out.append((f"ms{tag}{abs(ms)}_zipf_fse_{n}", C.FSE, 0, 1, "zipf_bytes", {"n": n, "s": 1.1, "seed": n}, p))
            out.append((f"ms{tag}{abs(ms)}_zipf_huf_{n}", C.HUFFMAN, 0, 1, "zipf_bytes", {"n": n, "s": 1.1, "seed": n}, p))
        out.append((f"ms{tag}{abs(ms)}_ts_delta_ts_fse_1000", C.DELTA_TS_FSE, 2, 4, "timestamps_u32", {"n": 1000}, p))
        out.append((f"ms{tag}{abs(ms)}_col_delta_zz_bp_16", C.DELTA_ZZ_BP, 2, 2, "walk_column",
                    {"n": 16, "width": 2, "b": 2, "seed": 16}, p))
        out.append((f"ms{tag}{abs(ms)}_bf16_ts_huf_fse_3000", C.TS_HUF_FSE, 1, 2, "bf16_normal", {"n": 3000}, p))
        out.append((f"ms{tag}{abs(ms)}_zipf_huf_100000", C.HUFFMAN, 0, 1, "zipf_bytes",
                    {"n": 100000, "s": 1.1, "seed": 100000}, p))
        out.append((f"ms{tag}{abs(ms)}_e2s_one_1", C.HUFFMAN, 1, 2, "uniform_ints",
                    {"n": 1, "hi": 65536, "dtype": "uint16"}, p))
        out.append((f"ms{tag}{abs(ms)}_e2s_tok_300", C.HUFFMAN, 1, 2, "u16_spread",
                    {"n": 300, "card": 40, "hi": 4000, "seed": 10}, p))
    return out


def make_input(gen, kw):
    kw = dict(kw)
    if "dtype" in kw:
        kw["dtype"] = np.dtype(kw["dtype"])
    return np.ascontiguousarray(getattr(G, gen)(**kw))


def sha(b):
    return hashlib.sha256(b).hexdigest()


def write_small():
    os.makedirs(os.path.join(HERE, "small"), exist_ok=True)
    entries = []
    total = 0
    cases = [c[:6] + (None,) for c in small_cases()] + [c + (None,) for c in entropy2_cases()] + param_cases()
    for name, kind, typ, width, gen, kw, params in cases:
        x = make_input(gen, kw)
        C.ref_set_min_stream_size((params or {}).get("minStreamSize", 0))
        try:
            fr = C.ref_compress(kind, x, typ, width)
        finally:
            C.ref_set_min_stream_size(0)
        back, t, w = C.ref_decompress(fr, x.nbytes + 64)
        assert back == x.tobytes(), name
        with open(os.path.join(HERE, "small", name + ".zl"), "wb") as f:
            f.write(fr)
        total += len(fr)
        e = {"name": name, "kind": kind, "type": typ, "width": width, "gen": gen, "kwargs": kw,
             "input_bytes": int(x.nbytes), "input_sha256": sha(x.tobytes()),
             "frame_bytes": len(fr), "frame_sha256": sha(fr)}
        if params:
            e["params"] = params
        entries.append(e)
    with open(os.path.join(HERE, "small.json"), "w") as f:
        json.dump(entries, f, indent=1)
    print(f"small: {len(entries)} frames, {total} bytes")


def digest_table(frames):
    tab = np.zeros((len(frames), 2), np.uint64)
    h = hashlib.sha256()
    for i, fr in enumerate(frames):
        tab[i, 0] = len(fr)
        tab[i, 1] = np.frombuffer(hashlib.sha256(fr).digest()[:8], np.uint64)[0]
        h.update(fr)
    return tab, h.hexdigest()


def compress_many(kind, typ, width, inputs, threads=8):
    with ThreadPoolExecutor(threads) as ex:
        return list(ex.map(lambda a: C.ref_compress(kind, a, typ, width), inputs))


def write_full():
    os.makedirs(os.path.join(HERE, "full"), exist_ok=True)
    meta = {}

    def record(name, frames, extra):
        tab, whole = digest_table(frames)
        np.save(os.path.join(HERE, "full", name + ".npy"), tab)
        meta[name] = dict(extra, frames=len(frames), total_frame_bytes=int(tab[:, 0].sum()), sha256_all=whole)
        print(f"{name}: {len(frames)} frames, {int(tab[:, 0].sum())} bytes", flush=True)

    t0 = time.time()
    # timestamps: one frame of 4,194,304 u32 through delta -> transpose_split -> FSE
    x = G.timestamps_u32()
    record("timestamps_delta_ts_fse", [C.ref_compress(C.DELTA_TS_FSE, x, 2, 4)],
           {"gen": "timestamps_u32", "kwargs": {}, "kind": C.DELTA_TS_FSE, "type": 2, "width": 4,
            "input_sha256": sha(x.tobytes())})
    # entropy stress: 16,384 blocks of 64 KiB, FSE and Huffman frames per block
    for kind, kn in ((C.FSE, "fse"), (C.HUFFMAN, "huf")):
        frames = []
        hin = hashlib.sha256()
        for a in range(0, 16384, 1024):
            data = G.zipf_blocks_np(a, 1024)
            hin.update(data.tobytes())
            frames += compress_many(kind, 0, 1, [data[i * 65536:(i + 1) * 65536] for i in range(1024)])
        record(f"zipf_blocks_{kn}", frames, {"gen": "zipf_blocks_np", "kwargs": {"first_block": 0, "n_blocks": 16384},
                                             "kind": kind, "type": 0, "width": 1, "block": 65536,
                                             "input_sha256": hin.hexdigest()})
    # batched columns: 4,096 columns through delta -> zigzag -> BITPACK, one frame each
    cols = G.random_walk_columns()
    hin = hashlib.sha256()
    for c in cols:
        hin.update(c.tobytes())
    frames = []
    for a in range(0, len(cols), 256):
        part = cols[a:a + 256]
        with ThreadPoolExecutor(8) as ex:
            frames += list(ex.map(lambda c: C.ref_compress(C.DELTA_ZZ_BP, c, 2, c.dtype.itemsize), part))
    record("columns_delta_zz_bp", frames, {"gen": "random_walk_columns", "kwargs": {}, "kind": C.DELTA_ZZ_BP,
                                           "type": 2, "width": 0, "input_sha256": hin.hexdigest()})
    # model weights: 536,870,912 bf16 as 64 segments of 8 Mi elements, each one frame
    w = G.bf16_normal()
    seg = 1 << 23
    frames = compress_many(C.TS_HUF_FSE, 1, 2, [w[a:a + seg] for a in range(0, w.size, seg)])
    record("bf16_ts_huf_fse", frames, {"gen": "bf16_normal", "kwargs": {}, "kind": C.TS_HUF_FSE, "type": 1,
                                       "width": 2, "segment_elements": seg, "input_sha256": sha(w.tobytes())})
    with open(os.path.join(HERE, "full.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"full: {time.time() - t0:.0f} s")


def write_select_full():
    """The learned-selection workload's 8,192 chunks (datagen.selector_chunks,
    the benchmark's seed) through the reference's brute-force selector over
    the eight candidates: per-frame digests, added to full.json."""
    path = os.path.join(HERE, "full.json")
    with open(path) as f:
        meta = json.load(f)
    seed = G.SEEDS[5]
    _, chunks = G.selector_chunks(8192, seed=seed)
    hin = hashlib.sha256()
    for c in chunks:
        hin.update(c.tobytes())
    with ThreadPoolExecutor(8) as ex:
        frames = list(ex.map(lambda c: C.ref_compress(C.SELECT8, c, 2, c.dtype.itemsize), chunks))
    tab, whole = digest_table(frames)
    np.save(os.path.join(HERE, "full", "select_8192_brute_force.npy"), tab)
    meta["select_8192_brute_force"] = {"gen": "selector_chunks", "kwargs": {"n_chunks": 8192, "seed": seed},
                                       "kind": C.SELECT8, "type": 2, "width": 0, "input_sha256": hin.hexdigest(),
                                       "frames": len(frames), "total_frame_bytes": int(tab[:, 0].sum()),
                                       "sha256_all": whole}
    with open(path, "w") as f:
        json.dump(meta, f, indent=1)
    print(f"select_8192_brute_force: {len(frames)} frames, {int(tab[:, 0].sum())} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--select-full", action="store_true",
                    help="only add the 8,192-chunk brute-force selector digests to full.json")
    a = ap.parse_args()
    if C.refh() is None:
        sys.exit("oracle/_ref/libref_harness.so is not built (make -C oracle ref)")
    if a.select_full:
        return write_select_full()
    write_small()
    if a.full:
        write_full()
    write_select_full()


if __name__ == "__main__":
    main()
