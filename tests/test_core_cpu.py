"""CPU unit build of the device codec core (openzl_amd/csrc/ozl_core.h) against
the oracle: decisions, FSE normalisation/tables/encode/NCount/decode and the
Huffman tree. The same source runs on the GPU inside the plan kernels; this
catches logic errors before any kernel launch. Bit-exact."""
import ctypes
import os

import numpy as np
import pytest

from ozl_checkers import oracle, ptr
from test_oracle import histogram_cases, hist, zipf_bytes

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "cpu_unit", "libozl_core_cpu.so")


@pytest.fixture(scope="module")
def cu():
    if not os.path.exists(SO):
        pytest.skip("CPU unit build missing (run __graft_entry__.build())")
    L = ctypes.CDLL(SO)
    P, SZ, U, I = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint, ctypes.c_int
    L.cu_decide.restype = I
    L.cu_decide.argtypes = [P, SZ, I]
    L.cu_entropy.restype = ctypes.c_double
    L.cu_entropy.argtypes = [P, U, SZ]
    L.cu_fse_normalize.restype = I
    L.cu_fse_normalize.argtypes = [P, U, P, SZ, U]
    L.cu_fse_optimal_tablelog.restype = U
    L.cu_fse_optimal_tablelog.argtypes = [U, SZ, U, U]
    L.cu_fse_build_ctable.restype = I
    L.cu_fse_build_ctable.argtypes = [P, P, P, U, U]
    L.cu_fse_encode.restype = SZ
    L.cu_fse_encode.argtypes = [P, SZ, P, SZ, P, P, U, U]
    L.cu_fse_decode.restype = I
    L.cu_fse_decode.argtypes = [P, SZ, P, SZ, P, U, U, U]
    L.cu_fse_write_ncount.restype = SZ
    L.cu_fse_write_ncount.argtypes = [P, P, U, U]
    L.cu_fse_read_ncount.restype = ctypes.c_long
    L.cu_fse_read_ncount.argtypes = [P, P, P, P, SZ]
    L.cu_huf_build.restype = I
    L.cu_huf_build.argtypes = [P, P, P, U, U]
    L.cu_huf_from_weights.restype = I
    L.cu_huf_from_weights.argtypes = [P, P, P, U]
    L.cu_flatpack.restype = SZ
    L.cu_flatpack.argtypes = [P, P, P, P, SZ]
    return L


def test_decide_entropy(cu):
    O = oracle()
    rng = np.random.default_rng(5)
    xs = histogram_cases() + [zipf_bytes(rng, 256, 2.0, card=12), np.full(50, 3, np.uint8)]
    for x in xs:
        c, ms = hist(x)
        assert cu.cu_entropy(ptr(c), 255, x.size) == O.orc_entropy(ptr(c), 255, x.size)
        for huf in (0, 1):
            assert cu.cu_decide(ptr(x), x.size, huf) == O.orc_entropy_decide(ptr(x), x.size, huf)


def test_fse_tables_encode_decode(cu):
    O = oracle()
    for x in histogram_cases():
        c, ms = hist(x)
        if np.count_nonzero(c) < 2:
            continue
        n = x.size
        tl = cu.cu_fse_optimal_tablelog(11, n, ms, 2)
        assert tl == O.orc_fse_optimal_tablelog(11, n, ms, 2)
        a = np.zeros(256, np.int16)
        b = np.zeros(256, np.int16)
        ra = cu.cu_fse_normalize(ptr(a), tl, ptr(c), n, ms)
        rb = O.orc_fse_normalize(ptr(b), tl, ptr(c), n, ms, 1)
        assert (ra != 0) == (rb != 0)
        if ra:
            continue
        assert np.array_equal(a[: ms + 1], b[: ms + 1])
        st = np.zeros(4096, np.uint16)
        tt = np.zeros(256, np.uint32)
        assert cu.cu_fse_build_ctable(ptr(st), ptr(tt), ptr(a), ms, tl) == 0
        oct_ = ctypes.create_string_buffer(8 + 2 * 4096 + 8 * 256)
        O.orc_fse_build_ctable(oct_, ptr(b), ms, tl)
        ob = np.frombuffer(oct_.raw, np.uint8)
        assert np.array_equal(st[: 1 << tl], ob[8: 8 + 2 * 4096].view(np.uint16)[: 1 << tl])
        for nbStates in (2, 4):
            cap = n + n // 128 + 64
            ea = np.zeros(cap, np.uint8)
            eb = np.zeros(cap, np.uint8)
            sa = cu.cu_fse_encode(ptr(ea), cap, ptr(x), n, ptr(st), ptr(tt), tl, nbStates)
            sb = O.orc_fse_encode(ptr(eb), cap, ptr(x), n, oct_, nbStates)
            assert sa == sb
            if sa:
                assert np.array_equal(ea[:sa], eb[:sb])
                d = np.zeros(n, np.uint8)
                assert cu.cu_fse_decode(ptr(d), n, ptr(ea), sa, ptr(a), ms, tl, nbStates) == 0
                assert np.array_equal(d, x)
        h1 = np.zeros(600, np.uint8)
        h2 = np.zeros(600, np.uint8)
        l1 = cu.cu_fse_write_ncount(ptr(h1), ptr(a), ms, tl)
        l2 = O.orc_fse_write_ncount(ptr(h2), 600, ptr(b), ms, tl)
        assert l1 == l2 and np.array_equal(h1[:l1], h2[:l2])
        back = np.zeros(256, np.int16)
        m2 = ctypes.c_uint(0)
        t2 = ctypes.c_uint(0)
        assert cu.cu_fse_read_ncount(ptr(back), ctypes.byref(m2), ctypes.byref(t2), ptr(h1), l1) == l1
        assert m2.value == ms and t2.value == tl and np.array_equal(back[: ms + 1], a[: ms + 1])


def test_huf_build_and_weights(cu):
    O = oracle()
    for x in histogram_cases():
        c, ms = hist(x)
        if np.count_nonzero(c) < 2:
            continue
        tl = O.orc_huf_optimal_tablelog(11, x.size, ms)
        na = np.zeros(256, np.uint8); va = np.zeros(256, np.uint16)
        nb = np.zeros(256, np.uint8); vb = np.zeros(256, np.uint16)
        ra = cu.cu_huf_build(ptr(na), ptr(va), ptr(c), ms, tl)
        rb = O.orc_huf_build_ctable(ptr(nb), ptr(vb), ptr(c), ms, tl)
        assert ra == rb
        assert np.array_equal(na[: ms + 1], nb[: ms + 1]) and np.array_equal(va[: ms + 1], vb[: ms + 1])
        w = np.where(na[: ms + 1] > 0, ra + 1 - na[: ms + 1].astype(np.int32), 0).astype(np.uint8)
        ln = np.zeros(256, np.uint8); vl = np.zeros(256, np.uint16)
        assert cu.cu_huf_from_weights(ptr(ln), ptr(vl), ptr(w), ms + 1) == ra
        assert np.array_equal(ln[: ms + 1], na[: ms + 1]) and np.array_equal(vl[: ms + 1], va[: ms + 1])


def test_flatpack(cu):
    O = oracle()
    rng = np.random.default_rng(9)
    for n in (1, 7, 100, 3000):
        for card in (2, 3, 9, 100):
            x = rng.choice(rng.permutation(256)[:card], n).astype(np.uint8)
            a1 = np.zeros(256, np.uint8); p1 = np.zeros(n + 8, np.uint8); s1 = ctypes.c_size_t(0)
            a2 = np.zeros(256, np.uint8); p2 = np.zeros(n + 8, np.uint8); s2 = ctypes.c_size_t(0)
            c1 = cu.cu_flatpack(ptr(a1), ptr(p1), ctypes.byref(s1), ptr(x), n)
            c2 = O.orc_flatpack_encode(ptr(a2), ptr(p2), ctypes.byref(s2), ptr(x), n)
            assert c1 == c2 and s1.value == s2.value
            assert np.array_equal(a1[:c1], a2[:c2]) and np.array_equal(p1[: s1.value], p2[: s2.value])


def test_dparse_recognises_golden_frames(cu):
    """The batched decoder's frame recognition (ozl_dparse.h, CPU build) on the
    reference's FSE / Huffman frames below 100000 bytes: every frame is
    accepted, and the normalized counts / code lengths it leaves decode the
    frame's payload back to the input with the sequential decoders."""
    import golden_cases as GC
    P, SZ, U = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint
    cu.cu_dparse.restype = ctypes.c_int
    cu.cu_dparse.argtypes = [P, SZ, P, P, ctypes.c_int]
    cu.cu_dplan_size.restype = SZ
    import struct
    seen = {"fse": 0, "huf": 0}
    for e in GC.small_entries():
        # the batched decoder reads frames written with the default parameters
        if e["kind"] not in (1, 2) or e["type"] != 0 or e["input_bytes"] >= 100000 or e.get("params"):
            continue
        x = GC.make_input(e["gen"], e["kwargs"])
        fr = np.frombuffer(GC.small_frame(e["name"]), np.uint8)
        plan = np.zeros(cu.cu_dplan_size(), np.uint8)
        side = np.zeros(512, np.uint8)
        st = cu.cu_dparse(ptr(fr), fr.size, ptr(plan), ptr(side), 1)
        assert st == 0, e["name"]
        n, bk, tl, nbs, _ = struct.unpack_from("<IBBBB", plan, 0)
        payOff, payLen, _, _, aux = struct.unpack_from("<5I", plan, 8)
        assert n == x.size
        pay = fr[payOff:payOff + payLen].copy()
        if bk == 5:  # FSE
            norm = side.view(np.int16).copy()
            out = np.zeros(max(n, 1), np.uint8)
            assert cu.cu_fse_decode(ptr(out), n, ptr(pay), payLen, ptr(norm), aux, tl, nbs) == 0
            assert out[:n].tobytes() == x.tobytes(), e["name"]
            seen["fse"] += 1
        elif bk == 4:  # Huffman: code lengths are those of the frame's weights
            lens = side[:256]
            assert lens.max() == tl and (lens[np.unique(x)] > 0).all(), e["name"]
            kraft = sum(2.0 ** (tl - l) for l in lens if l)
            assert kraft == 2 ** tl
            seen["huf"] += 1
    assert seen["fse"] >= 5 and seen["huf"] >= 5


def _dparse(cu, fr, words):
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    cu.cu_dparse.restype = ctypes.c_int
    cu.cu_dparse.argtypes = [P, SZ, P, P, ctypes.c_int]
    cu.cu_dplan_size.restype = SZ
    plan = np.zeros(cu.cu_dplan_size(), np.uint8)
    side = np.zeros(512, np.uint8)
    st = cu.cu_dparse(ptr(fr), fr.size, ptr(plan), ptr(side), words)
    return st, plan, side


def test_dparse_word_readers_match_byte_readers(cu):
    """The parser's 32-bit-window readers (NCount, nested weights FSE stream:
    ozl_read_ncount_w / ozl_fse_decode_w, used on the device) against its
    byte-by-byte readers (ozl_fse_read_ncount / ozl_fse_decode_seq, pinned to
    the reference) on the reference's FSE / Huffman frames and on copies with
    single bit flips, truncations and byte swaps in the header, NCount and
    weights area: the same status, plan and tables every time."""
    import golden_cases as GC
    rng = np.random.default_rng(21)
    frames = []
    for e in GC.small_entries():
        if e["kind"] not in (1, 2) or e["type"] != 0 or e["input_bytes"] >= 100000 or e.get("params"):
            continue
        frames.append(np.frombuffer(GC.small_frame(e["name"]), np.uint8).copy())
    assert len(frames) >= 10
    n_ok = n_bad = 0
    for fr in frames:
        variants = [fr]
        head = min(fr.size, 400)
        for _ in range(40):
            v = fr.copy()
            i = int(rng.integers(0, head))
            kind = int(rng.integers(0, 3))
            if kind == 0:
                v[i] ^= np.uint8(1 << int(rng.integers(0, 8)))
            elif kind == 1:
                v = v[: max(1, fr.size - int(rng.integers(1, 16)))].copy()
            else:
                j = int(rng.integers(0, head))
                v[i], v[j] = v[j], v[i]
            variants.append(v)
        for v in variants:
            a = _dparse(cu, v, 1)
            b = _dparse(cu, v, 0)
            assert a[0] == b[0]
            assert np.array_equal(a[1], b[1])
            if a[0] == 0:
                assert np.array_equal(a[2], b[2])
                n_ok += 1
            else:
                n_bad += 1
    assert n_ok >= len(frames) and n_bad >= 50
