"""Reader for OpenZL frames (format v21+), used by tests to inspect structure.

Layout restated from src/openzl/common/wire_format.h:16-128 and the writers in
src/openzl/compress/encode_frameheader.c:441-635 (frame header) and :667-832
(chunk header).
"""

TRID_NAMES = {
    1: "delta_int", 3: "zigzag", 4: "transpose_split", 6: "convert_struct_to_serial",
    7: "convert_struct_to_num_le", 8: "convert_num_to_struct_le", 10: "convert_num_to_serial_le",
    27: "bitpack_serial", 28: "bitpack_int", 29: "flatpack", 40: "splitn", 44: "constant_serial",
    36: "tokenize_fixed", 45: "constant_fixed", 47: "splitn_struct", 48: "splitn_num", 49: "fse_v2",
    50: "huffman_v2", 51: "huffman_struct_v2", 52: "fse_ncount",
}


def _varint(b, p):
    v = 0
    s = 0
    while True:
        x = b[p]
        p += 1
        v |= (x & 127) << s
        s += 7
        if x < 128:
            return v, p


def _unpack(b, p, n, nb):
    nbytes = (n * nb + 7) // 8
    v = int.from_bytes(b[p:p + nbytes], "little")
    return [(v >> (i * nb)) & ((1 << nb) - 1) for i in range(n)], p + nbytes


def _next_pow2(x):
    return 0 if x <= 1 else (x - 1).bit_length()


def parse_frame(frame):
    b = bytes(frame)
    magic = int.from_bytes(b[0:4], "little")
    version = magic - 0xD7B1A5C0
    p = 4
    flags = b[p]
    p += 1
    tok = b[p]
    p += 1
    n_in = tok & 15
    types = [(tok >> 4) & 3, (tok >> 6) & 3][:n_in]
    if n_in > 2:
        p += (n_in - 2 + 3) // 4
    sizes = []
    for _ in range(n_in):
        v, p = _varint(b, p)
        sizes.append(v - 1)
    if flags & 2:
        p += 1
    chunks = []
    while True:
        nbdec, p = _varint(b, p)
        if nbdec == 0:
            break
        nbdec -= 1
        nbuf, p = _varint(b, p)
        nbuf += 1
        trt, p = _unpack(b, p, nbdec, 1)
        nstd = sum(1 for t in trt if t == 0)
        ids, p = _unpack(b, p, nstd, 6)
        flags_h, p = _unpack(b, p, nbdec, 1)
        hs = []
        for fl in flags_h:
            if fl:
                v, p = _varint(b, p)
                hs.append(v + 1)
            else:
                hs.append(0)
        flags_v, p = _unpack(b, p, nbdec, 1)
        vos = []
        for fl in flags_v:
            if fl:
                v, p = _varint(b, p)
                vos.append(v + 1)
            else:
                vos.append(0)
        flags_i, p = _unpack(b, p, nbdec, 1)
        nis = []
        for fl in flags_i:
            if fl:
                v, p = _varint(b, p)
                nis.append(v + 2)
            else:
                nis.append(1)
        tot = sum(nis)
        dist, p = _unpack(b, p, tot, _next_pow2(tot + nbuf - 1))
        ssz = []
        for _ in range(nbuf - 1):
            v, p = _varint(b, p)
            ssz.append(v)
        hstream = b[p:p + sum(hs)]
        p += sum(hs)
        decs = []
        di = hp = 0
        for i in range(nbdec):
            decs.append({
                "trid": ids[i] if trt[i] == 0 else None,
                "name": TRID_NAMES.get(ids[i], str(ids[i])),
                "header": hstream[hp:hp + hs[i]],
                "nb_vo": vos[i],
                "nb_in": nis[i],
                "dist": [x + 1 for x in dist[di:di + nis[i]]],
            })
            di += nis[i]
            hp += hs[i]
        streams = []
        for s in ssz:
            streams.append(b[p:p + s])
            p += s
        content = compressed = None
        if flags & 1:
            content = int.from_bytes(b[p:p + 4], "little")
            p += 4
        if flags & 2:
            compressed = int.from_bytes(b[p:p + 4], "little")
            p += 4
        chunks.append({"decoders": decs, "streams": streams, "content_checksum": content,
                       "compressed_checksum": compressed})
    return {"version": version, "flags": flags, "types": types, "sizes": sizes,
            "chunks": chunks, "end": p, "length": len(b)}
