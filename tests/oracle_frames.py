"""CPU universal decoder of single-input OpenZL frames built from the oracle's
kernels (oracle/liboracle.so). Test infrastructure: it checks the oracle's
decode functions against reference frames, and decodes frames made on the GPU.

Graph reconstruction follows src/openzl/decompress/decode_frameheader.c and
decompress2.c: nodes are listed in reverse creation order; stream ids are
assigned in creation order (input 0, then each node's singleton outputs
followed by its variable outputs); a node's input id is firstOutput - distance;
the streams no node consumes are the stored ones, in decreasing id order.
Per-transform decoders restate src/openzl/codecs/*/decode_*_binding.c.
"""
import ctypes

import numpy as np

from frames import parse_frame
from ozl_checkers import oracle, ptr

SINGLETONS = {1: 1, 3: 1, 4: 0, 6: 1, 7: 1, 8: 1, 10: 1, 27: 1, 28: 1, 29: 2, 36: 2, 40: 0, 44: 1, 45: 1, 47: 0, 48: 0,
              49: 2, 50: 2, 51: 2, 52: 1}


class FrameError(ValueError):
    pass


def _varint(b, p=0):
    v = s = 0
    while True:
        x = b[p]
        p += 1
        v |= (x & 127) << s
        s += 7
        if x < 128:
            return v, p


def _frame_header_len(b):
    _, p = _varint(b, 6)
    return p + (1 if b[4] & 2 else 0)


def _norm_from_int16(raw):
    norm = np.frombuffer(raw, np.int16).copy()
    tot = int(sum(1 if v == -1 else int(v) for v in norm))
    if len(norm) == 0 or tot == 0 or tot & (tot - 1):
        raise FrameError("bad normalized counts")
    return norm, len(norm) - 1, tot.bit_length() - 1


def _decode_node(O, trid, hdr, ins):
    """ins: list of (bytes, type, width) of the node's outputs; returns the
    regenerated input as (bytes, type, width). Types: 's' serial, 't' struct,
    'n' numeric."""
    if trid == 1:  # delta_int (decode_delta_binding.c:10-62)
        d, t, w = ins[0]
        n = len(d) // w + (1 if hdr else 0)
        if hdr and len(hdr) != w:
            raise FrameError("delta header size")
        out = np.empty(n * w, np.uint8)
        dd = np.frombuffer(d, np.uint8) if d else np.zeros(1, np.uint8)
        first = np.frombuffer(hdr, np.uint8) if hdr else np.zeros(8, np.uint8)
        O.orc_delta_decode(ptr(out), ptr(first), ptr(dd), n, w)
        return out.tobytes(), "n", w
    if trid == 3:  # zigzag
        d, t, w = ins[0]
        src = np.frombuffer(d, np.uint8) if d else np.zeros(1, np.uint8)
        out = np.empty(max(len(d), 1), np.uint8)
        O.orc_zigzag_decode(ptr(out), ptr(src), len(d) // w, w)
        return out[:len(d)].tobytes(), "n", w
    if trid == 4:  # transpose_split (decode_transpose_binding.c:37-74)
        w = len(ins)
        n = len(ins[0][0])
        if any(len(p[0]) != n for p in ins):
            raise FrameError("plane sizes differ")
        planes = np.frombuffer(b"".join(p[0] for p in ins), np.uint8) if n else np.zeros(1, np.uint8)
        out = np.empty(max(n * w, 1), np.uint8)
        O.orc_transpose_join(ptr(out), ptr(planes), n, w)
        return out[:n * w].tobytes(), "t", w
    if trid == 6:  # struct -> serial conversion: regenerates a struct stream
        w, _ = _varint(hdr)
        return ins[0][0], "t", w
    if trid == 7:
        return ins[0][0], "t", ins[0][2]
    if trid == 8:  # num -> struct: regenerates numeric
        return ins[0][0], "n", ins[0][2]
    if trid == 10:  # num -> serial: regenerates numeric of width 1 << hdr[0]
        return ins[0][0], "n", 1 << hdr[0]
    if trid in (27, 28):  # bitpack (decode_bitpack_binding.c:10-60)
        src = np.frombuffer(ins[0][0], np.uint8) if ins[0][0] else np.zeros(1, np.uint8)
        h = np.frombuffer(hdr, np.uint8)
        cap = len(ins[0][0]) * 8 + 8
        out = np.empty(cap * 8, np.uint8)
        wo = ctypes.c_int(0)
        n = O.orc_bitpack_decode(ptr(out), cap, ptr(src), len(ins[0][0]), ptr(h), len(hdr), ctypes.byref(wo))
        if n == ctypes.c_size_t(-1).value:
            raise FrameError("bitpack")
        w = wo.value
        return out[:n * w].tobytes(), ("n" if trid == 28 else "s"), w
    if trid == 29:  # flatpack
        alpha = np.frombuffer(ins[0][0], np.uint8)
        packed = np.frombuffer(ins[1][0], np.uint8)
        cap = len(packed) * 8 + 8
        out = np.empty(cap, np.uint8)
        n = O.orc_flatpack_decode(ptr(out), cap, ptr(alpha), len(alpha), ptr(packed), len(packed))
        if n == ctypes.c_size_t(-1).value:
            raise FrameError("flatpack")
        return out[:n].tobytes(), "s", 1
    if trid == 40:  # splitN: concatenation
        return b"".join(p[0] for p in ins), ins[0][1] if ins else "s", ins[0][2] if ins else 1
    if trid in (47, 48):  # splitN of struct / numeric streams: concatenation
        if not ins:
            w, _ = _varint(hdr)
            return b"", "t" if trid == 47 else "n", w
        if any(p[2] != ins[0][2] for p in ins):
            raise FrameError("splitN widths differ")
        return b"".join(p[0] for p in ins), "t" if trid == 47 else "n", ins[0][2]
    if trid == 45:  # constant_fixed (decode_constant_binding.c:10-47)
        n, _ = _varint(hdr)
        v, t, w = ins[0]
        if len(v) != w or n < 1:
            raise FrameError("constant")
        return v * n, "t", w
    if trid == 36:  # tokenize (decode_tokenize_binding.c:12-37)
        alpha, _, w = ins[0]
        idx, _, iw = ins[1]
        a = np.frombuffer(alpha, np.uint8).reshape(-1, w)
        ix = np.frombuffer(idx, {1: np.uint8, 2: np.uint16, 4: np.uint32}[iw]).astype(np.int64)
        if ix.size and ix.max() >= a.shape[0]:
            raise FrameError("tokenize index")
        return a[ix].tobytes(), "t", w
    if trid == 51:  # huffman_struct_v2 (decode_entropy_binding.c:415-472)
        if not 2 <= len(hdr) <= 9:
            raise FrameError("header")
        x4 = hdr[0] & 1
        n = int.from_bytes(hdr[1:], "little")
        wts = np.frombuffer(ins[0][0], np.uint8)
        bits = np.frombuffer(ins[1][0], np.uint8) if ins[1][0] else np.zeros(1, np.uint8)
        out = np.empty(max(n, 1), np.uint16)
        if n < 2 or O.orc_huf16_decode(ptr(out), n, ptr(bits), len(ins[1][0]), ptr(wts), len(wts), x4) != 0:
            raise FrameError("huffman16")
        return out[:n].tobytes(), "t", 2
    if trid == 44:  # constant_serial
        n, _ = _varint(hdr)
        if len(ins[0][0]) != 1:
            raise FrameError("constant")
        return ins[0][0] * n, "s", 1
    if trid == 49:  # fse_v2 (decode_entropy_binding.c:64-116)
        nb_states = hdr[0]
        n = int.from_bytes(hdr[1:], "little")
        norm, max_sym, tl = _norm_from_int16(ins[0][0])
        bits = np.frombuffer(ins[1][0], np.uint8) if ins[1][0] else np.zeros(1, np.uint8)
        out = np.empty(max(n, 1), np.uint8)
        if O.orc_fse_decode(ptr(out), n, ptr(bits), len(ins[1][0]), ptr(norm), max_sym, tl, nb_states) != 0:
            raise FrameError("fse")
        return out[:n].tobytes(), "s", 1
    if trid == 50:  # huffman_v2 (decode_entropy_binding.c:286-347)
        x4 = hdr[0]
        n = int.from_bytes(hdr[1:], "little")
        wts = np.frombuffer(ins[0][0], np.uint8)
        bits = np.frombuffer(ins[1][0], np.uint8) if ins[1][0] else np.zeros(1, np.uint8)
        out = np.empty(max(n, 1), np.uint8)
        if O.orc_huf_decode(ptr(out), n, ptr(bits), len(ins[1][0]), ptr(wts), len(wts), x4) != 0:
            raise FrameError("huffman")
        return out[:n].tobytes(), "s", 1
    if trid == 52:  # fse_ncount (decode_entropy_binding.c:118-144)
        src = np.frombuffer(ins[0][0], np.uint8)
        norm = np.zeros(256, np.int16)
        ms = ctypes.c_uint(0)
        tl = ctypes.c_uint(0)
        if O.orc_fse_read_ncount(ptr(norm), ctypes.byref(ms), ctypes.byref(tl), ptr(src), len(src)) < 0:
            raise FrameError("ncount")
        return norm[:ms.value + 1].tobytes(), "n", 2
    raise FrameError(f"transform {trid} not supported")


def decode_frame(frame):
    """Returns (bytes, type, width); raises FrameError on malformed frames or
    checksum mismatches."""
    O = oracle()
    b = bytes(frame)
    f = parse_frame(b)
    if f["version"] != 23 or len(f["sizes"]) != 1 or len(f["chunks"]) != 1:
        raise FrameError("unsupported frame")
    ch = f["chunks"][0]
    decs = ch["decoders"]
    nodes = list(reversed(decs))  # creation order
    nxt = 1
    info = []
    for nd in nodes:
        s1 = SINGLETONS[nd["trid"]]
        first = nxt
        nout = s1 + nd["nb_vo"]
        nxt += nout
        src = first - nd["dist"][0]
        if src < 0:
            raise FrameError("distance")
        info.append((first, nout, src))
    consumed = {src for _, _, src in info}
    stored_ids = [s for s in range(nxt - 1, -1, -1) if s not in consumed]
    if len(stored_ids) != len(ch["streams"]):
        raise FrameError("stored stream count")
    S = {}
    for sid, data in zip(stored_ids, ch["streams"]):
        S[sid] = (bytes(data), "s", 1)
    for k in range(len(nodes) - 1, -1, -1):
        first, nout, src = info[k]
        ins = [S[first + j] for j in range(nout)]
        S[src] = _decode_node(O, nodes[k]["trid"], bytes(nodes[k]["header"]), ins)
    out, t, w = S[0]
    if len(out) != f["sizes"][0]:
        raise FrameError("size")
    ty = f["types"][0]
    # checksums (cctx.c:1537-1580): XXH3-64 low 32 bits; the frame ends with
    # [content checksum][compressed checksum] and the end-of-frame byte 0
    flags = f["flags"]
    end = f["end"]
    if end != len(b) or b[end - 1] != 0:
        raise FrameError("trailer")
    q = end - 1 - (4 if flags & 2 else 0)  # end of the content checksum
    if flags & 1:
        src = np.frombuffer(out, np.uint8) if out else np.zeros(1, np.uint8)
        if O.orc_xxh3_64(ptr(src), len(out)) & 0xFFFFFFFF != ch["content_checksum"]:
            raise FrameError("content checksum")
    if flags & 2:
        region = np.frombuffer(b[_frame_header_len(b):q], np.uint8)
        if O.orc_xxh3_64(ptr(region), len(region)) & 0xFFFFFFFF != ch["compressed_checksum"]:
            raise FrameError("compressed checksum")
    return out, ty, w
