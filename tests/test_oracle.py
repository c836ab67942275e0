"""Pins the CPU oracle (oracle/ozl_oracle.c) before it is trusted as a checker.

Pin per oracle function:
  orc_log2_table        reference LOG2_TABLE symbol + closed form (generator)
  orc_entropy           reference ZL_calculateEntropy (bit-equal doubles) + closed
                        forms (uniform over 2^k symbols is exactly k)
  orc_fse_normalize     reference ZS_FSE_normalizeCount, incl. the M2 fallback
  orc_fse_build_ctable  reference ZS_FSE_buildCTable (state table and symbol
                        transforms compared field by field)
  orc_fse_encode        reference ZS_FSE_compress_usingCTable, 2 and 4 states
  orc_fse_write_ncount  reference ZS_FSE_writeNCount
  orc_fse_read_ncount   reference ZS_FSE_readNCount + round trip
  orc_fse_decode        round trip of reference-encoded streams
  orc_huf_build_ctable  reference ZS_HUF_buildCTable (lengths and code values),
                        incl. height-limited trees
  orc_huf_encode_1x/4x  reference ZS_HUF_compress{1X,4X}_usingCTable
  orc_huf_decode        round trip of reference-encoded streams
  orc_delta/zigzag/transpose/bitpack/flatpack  reference ZS_* / ZL_* kernels,
                        plus the known-answer transpose pattern of
                        tests/unittest/transforms/test_transpose_kernel.cpp:10-28
  orc_entropy_decide    backend recorded in reference frames (transform ids)
  orc_xxh3_64           the independent xxhash package (XXH3-64, seed 0) and
                        the low 32 bits stored in reference frames
"""
import ctypes

import numpy as np
import pytest

from ozl_checkers import oracle, ref, ptr
import ozl_checkers as C

O = oracle()
R = ref()
needs_ref = pytest.mark.skipif(R is None, reason="reference library not built here")


def zipf_bytes(rng, n, s, card=256):
    p = 1.0 / np.arange(1, card + 1) ** s
    p /= p.sum()
    perm = rng.permutation(256)[:card]
    return perm[rng.choice(card, n, p=p)].astype(np.uint8)


def histogram_cases():
    rng = np.random.default_rng(1234)
    cases = []
    for n in (2, 3, 5, 17, 100, 999, 1000, 1001, 4096, 32768, 65536):
        for s in (0.0, 0.6, 1.1, 2.0, 4.0):
            cases.append(zipf_bytes(rng, n, s, card=int(rng.integers(2, 257))))
    # very skewed: one dominant symbol plus many singletons (M2 path, deep trees)
    for n in (300, 5000, 70000):
        x = np.zeros(n, np.uint8)
        idx = rng.choice(n, 255, replace=False)
        x[idx] = np.arange(1, 256, dtype=np.uint8)
        cases.append(x)
    # geometric distribution -> Huffman depth > 11 before limiting
    for n in (20000, 65536):
        g = np.minimum(rng.geometric(0.45, n) - 1, 255).astype(np.uint8)
        cases.append(g)
    # two symbols, 3 symbols
    cases.append(np.array([0, 1] * 50, np.uint8))
    cases.append(rng.integers(0, 3, 777).astype(np.uint8))
    return cases


def hist(x):
    c = np.bincount(x, minlength=256).astype(np.uint32)
    ms = int(np.nonzero(c)[0].max())
    return c, ms


def test_log2_table_closed_form_and_reference():
    t = np.ctypeslib.as_array((ctypes.c_uint16 * 256).in_dll(O, "orc_log2_table"))
    import math
    for i in range(256):
        # closed form: floor(log2(1+i/256)*65536); check with a float guard band
        v = math.log2(1 + i / 256) * 65536
        assert abs(t[i] - math.floor(v)) <= (1 if abs(v - round(v)) < 1e-6 else 0)
    if R is not None:
        r = np.ctypeslib.as_array((ctypes.c_uint16 * 256).in_dll(R, "LOG2_TABLE"))
        assert np.array_equal(t, r)


def test_entropy_closed_form():
    for k in range(1, 9):
        c = np.zeros(256, np.uint32)
        c[: 1 << k] = 1000
        h = O.orc_entropy(ptr(c), 255, 1000 << k)
        assert abs(h - k) < 5e-3 * k  # the fixed-point log2 is accurate to ~0.3%


@needs_ref
def test_entropy_matches_reference_bitwise():
    for x in histogram_cases():
        c, _ = hist(x)
        a = O.orc_entropy(ptr(c), 255, x.size)
        b = R.ZL_calculateEntropy(ptr(c), 255, x.size)
        assert a == b


@needs_ref
def test_fse_normalize_and_ctable_match_reference():
    for x in histogram_cases():
        c, ms = hist(x)
        if np.count_nonzero(c) < 2:
            continue
        n = x.size
        for maxTL in (11, 12, 9):
            tl = O.orc_fse_optimal_tablelog(maxTL, n, ms, 2)
            assert tl == R.ZS_FSE_optimalTableLog(maxTL, n, ms)
            na = np.zeros(256, np.int16)
            nb = np.zeros(256, np.int16)
            ra = O.orc_fse_normalize(ptr(na), tl, ptr(c), n, ms, 1)
            rb = R.ZS_FSE_normalizeCount(ptr(nb), tl, ptr(c), n, ms, 1)
            assert (ra != 0) == bool(R.ZS_FSE_isError(rb))
            if ra:
                continue
            assert np.array_equal(na[: ms + 1], nb[: ms + 1])
            ct = ctypes.create_string_buffer(4 * (4 + 4096 // 2 + 8 * 256) + 64)
            assert R.ZS_FSE_buildCTable(ct, ptr(nb), ms, tl) == 0
            raw = np.frombuffer(ct.raw, np.uint8)
            ts = 1 << tl
            st_ref = raw[4: 4 + 2 * ts].view(np.uint16)
            tt_off = 4 + 4 * (ts >> 1)
            tt = raw[tt_off: tt_off + 8 * (ms + 1)].view(np.int32).reshape(-1, 2)
            oc = ctypes.create_string_buffer(8 + 2 * 4096 + 8 * 256)
            assert O.orc_fse_build_ctable(oc, ptr(na), ms, tl) == 0
            ob = np.frombuffer(oc.raw, np.uint8)
            st = ob[8: 8 + 2 * 4096].view(np.uint16)[:ts]
            dfs = ob[8 + 8192: 8 + 8192 + 1024].view(np.int32)[: ms + 1]
            dnb = ob[8 + 8192 + 1024: 8 + 8192 + 2048].view(np.uint32)[: ms + 1]
            assert np.array_equal(st, st_ref)
            used = nb[: ms + 1] != 0
            assert np.array_equal(dfs[used], tt[used, 0])
            assert np.array_equal(dnb.view(np.int32), tt[:, 1])


def _fse_tables(x, tl=None):
    c, ms = hist(x)
    n = x.size
    if tl is None:
        tl = O.orc_fse_optimal_tablelog(11, n, ms, 2)
    norm = np.zeros(256, np.int16)
    assert O.orc_fse_normalize(ptr(norm), tl, ptr(c), n, ms, 1) == 0
    return norm, ms, tl


@needs_ref
def test_fse_encode_ncount_decode_match_reference():
    for x in histogram_cases():
        if np.count_nonzero(np.bincount(x, minlength=256)) < 2:
            continue
        norm, ms, tl = _fse_tables(x)
        rct = ctypes.create_string_buffer(4 * (4 + 4096 // 2 + 8 * 256) + 64)
        assert R.ZS_FSE_buildCTable(rct, ptr(norm), ms, tl) == 0
        oct_ = ctypes.create_string_buffer(8 + 2 * 4096 + 8 * 256)
        O.orc_fse_build_ctable(oct_, ptr(norm), ms, tl)
        for nbStates in (2, 4):
            cap = x.size + x.size // 128 + 64
            a = np.zeros(cap, np.uint8)
            b = np.zeros(cap, np.uint8)
            sa = O.orc_fse_encode(ptr(a), cap, ptr(x), x.size, oct_, nbStates)
            sb = R.ZS_FSE_compress_usingCTable(ptr(b), cap, ptr(x), x.size, rct, nbStates)
            assert sa == sb
            if sa == 0:
                continue
            assert np.array_equal(a[:sa], b[:sb])
            if x.size <= 4096:
                d = np.zeros(x.size, np.uint8)
                assert O.orc_fse_decode(ptr(d), x.size, ptr(b), sb, ptr(norm), ms, tl, nbStates) == 0
                assert np.array_equal(d, x)
        # ncount
        ha = np.zeros(600, np.uint8)
        hb = np.zeros(600, np.uint8)
        la = O.orc_fse_write_ncount(ptr(ha), 600, ptr(norm), ms, tl)
        lb = R.ZS_FSE_writeNCount(ptr(hb), 600, ptr(norm), ms, tl)
        assert la == lb and np.array_equal(ha[:la], hb[:lb])
        back = np.zeros(256, np.int16)
        msb = ctypes.c_uint(0)
        tlb = ctypes.c_uint(0)
        used = O.orc_fse_read_ncount(ptr(back), ctypes.byref(msb), ctypes.byref(tlb), ptr(ha), la)
        assert used == la and msb.value == ms and tlb.value == tl
        assert np.array_equal(back[: ms + 1], norm[: ms + 1])
        rback = np.zeros(256, np.int16)
        rms = ctypes.c_uint(255)
        rtl = ctypes.c_uint(0)
        ru = R.ZS_FSE_readNCount(ptr(rback), ctypes.byref(rms), ctypes.byref(rtl), ptr(ha), la)
        assert ru == la and rms.value == ms and rtl.value == tl


def _ref_huf_ctable(c, ms, maxbits):
    ct = np.zeros(ms + 2, np.uint64)
    r = R.ZS_HUF_buildCTable(ptr(ct), ptr(c), ms, maxbits)
    assert not R.ZS_HUF_isError(r)
    elts = ct[1: ms + 2]
    nb = (elts & 0xFF).astype(np.uint8)
    val = np.where(nb > 0, elts >> (64 - nb.astype(np.uint64)), 0).astype(np.uint16)
    return int(r), nb, val, ct


@needs_ref
def test_huf_ctable_encode_decode_match_reference():
    for x in histogram_cases():
        c, ms = hist(x)
        if np.count_nonzero(c) < 2:
            continue
        n = x.size
        tl = O.orc_huf_optimal_tablelog(11, n, ms)
        assert tl == R.ZS_HUF_optimalTableLog(11, n, ms)
        rtl, rnb, rval, rct = _ref_huf_ctable(c, ms, tl)
        onb = np.zeros(256, np.uint8)
        oval = np.zeros(256, np.uint16)
        otl = O.orc_huf_build_ctable(ptr(onb), ptr(oval), ptr(c), ms, tl)
        assert otl == rtl
        assert np.array_equal(onb[: ms + 1], rnb)
        assert np.array_equal(oval[: ms + 1], rval)
        cap = n + (n >> 8) + 8 + 129
        for x4, rf, of in ((0, R.ZS_HUF_compress1X_usingCTable, O.orc_huf_encode_1x),
                           (1, R.ZS_HUF_compress4X_usingCTable, O.orc_huf_encode_4x)):
            a = np.zeros(cap, np.uint8)
            b = np.zeros(cap, np.uint8)
            sa = of(ptr(a), cap, ptr(x), n, ptr(onb), ptr(oval))
            sb = rf(ptr(b), cap, ptr(x), n, ptr(rct))
            assert sa == sb
            if sa == 0:
                continue
            assert np.array_equal(a[:sa], b[:sb])
            if n <= 5000:
                w = np.where(onb[: ms + 1] > 0, otl + 1 - onb[: ms + 1].astype(np.int32), 0).astype(np.uint8)
                d = np.zeros(n, np.uint8)
                assert O.orc_huf_decode(ptr(d), n, ptr(b), sb, ptr(w), ms + 1, x4) == 0
                assert np.array_equal(d, x)


@needs_ref
@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_numeric_kernels_match_reference(width):
    rng = np.random.default_rng(width)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    for n in (0, 1, 2, 7, 100, 1000, 4099):
        x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
        x[: min(n, 3)] = [np.iinfo(dt).max, 0, 1 << (8 * width - 1)][: min(n, 3)]
        # delta
        fa = np.zeros(1, dt); fb = np.zeros(1, dt)
        da = np.zeros(max(n, 1), dt); db = np.zeros(max(n, 1), dt)
        O.orc_delta_encode(ptr(da), ptr(fa), ptr(x), n, width)
        if n:
            R.ZS_deltaEncode(ptr(fb), ptr(db), ptr(x), n, width)
            assert fa[0] == fb[0] and np.array_equal(da[: n - 1], db[: n - 1])
            back = np.zeros(n, dt)
            O.orc_delta_decode(ptr(back), ptr(fa), ptr(da), n, width)
            assert np.array_equal(back, x)
        # zigzag
        za = np.zeros(max(n, 1), dt); zb = np.zeros(max(n, 1), dt)
        O.orc_zigzag_encode(ptr(za), ptr(x), n, width)
        getattr(R, f"ZL_zigzagEncode{8 * width}")(ptr(zb), ptr(x), n)
        assert np.array_equal(za[:n], zb[:n])
        zback = np.zeros(max(n, 1), dt)
        O.orc_zigzag_decode(ptr(zback), ptr(za), n, width)
        assert np.array_equal(zback[:n], x)
        # transpose
        pa = np.zeros(max(n * width, 1), np.uint8)
        pb = np.zeros(max(n * width, 1), np.uint8)
        O.orc_transpose_split(ptr(pa), ptr(x), n, width)
        ptrs = (ctypes.c_void_p * width)(*[pb.ctypes.data + k * n for k in range(width)])
        R.ZS_splitTransposeEncode(ptrs, ptr(x), n, width)
        assert np.array_equal(pa, pb)
        # bitpack at every legal width
        xb = x.copy()
        for nbBits in sorted({1, 3, 8 * width // 2 + 1, 8 * width}):
            if nbBits > 8 * width:
                continue
            m = (1 << nbBits) - 1
            v = (xb & dt(m)) if nbBits < 64 else xb
            cap = (n * nbBits + 7) // 8 + 16
            ba = np.zeros(cap, np.uint8); bb = np.zeros(cap, np.uint8)
            sa = O.orc_bitpack_encode(ptr(ba), ptr(v), n, width, nbBits)
            sb = R.ZS_bitpackEncode(ptr(bb), cap, ptr(v), n, width, nbBits)
            assert sa == sb and np.array_equal(ba[:sa], bb[:sb])
            hdr = np.zeros(2, np.uint8)
            hs = O.orc_bitpack_header(ptr(hdr), n, width, nbBits, sa)
            out = np.zeros(max(n, 1) + 8, dt)
            wo = ctypes.c_int(0)
            got = O.orc_bitpack_decode(ptr(out), out.size, ptr(ba), sa, ptr(hdr), hs, ctypes.byref(wo))
            if n:
                assert got == n and np.array_equal(out[:n], v)


def test_transpose_known_answer():
    # tests/unittest/transforms/test_transpose_kernel.cpp:10-28: input byte i,
    # output plane j holds bytes j, j+width, j+2*width, ...
    for width in (1, 2, 3, 4, 7, 8):
        for n in (0, 1, 5, 20, 100):
            src = (np.arange(n * width) % 251).astype(np.uint8)
            out = np.zeros(max(n * width, 1), np.uint8)
            O.orc_transpose_split(ptr(out), ptr(src), n, width)
            exp = np.concatenate([src[j::width] for j in range(width)]) if n else np.zeros(1, np.uint8)
            assert np.array_equal(out[: n * width], exp[: n * width])


@needs_ref
def test_flatpack_matches_reference():
    rng = np.random.default_rng(7)
    for n in (1, 2, 9, 100, 1001):
        for card in (1, 2, 3, 5, 17, 200):
            x = rng.choice(rng.permutation(256)[:card], n).astype(np.uint8)
            aa = np.zeros(256, np.uint8); pa = np.zeros(n + 8, np.uint8)
            ab = np.zeros(256, np.uint8); pb = np.zeros(n + 8, np.uint8)
            psz = ctypes.c_size_t(0)
            ca = O.orc_flatpack_encode(ptr(aa), ptr(pa), ctypes.byref(psz), ptr(x), n)
            cb = R.ZS_flatpackEncode(ptr(ab), 256, ptr(pb), n + 8, ptr(x), n)
            assert ca == cb and np.array_equal(aa[:ca], ab[:cb])
            assert np.array_equal(pa[: psz.value], pb[: psz.value])
            back = np.zeros(n, np.uint8)
            assert O.orc_flatpack_decode(ptr(back), n, ptr(aa), ca, ptr(pa), psz.value) == n
            assert np.array_equal(back, x)


def test_xxh3_matches_xxhash_package():
    xxhash = pytest.importorskip("xxhash")
    rng = np.random.default_rng(3)
    lens = list(range(0, 260)) + [1023, 1024, 1025, 2047, 2048, 2049, 4096, 65536, 100003]
    for L in lens:
        b = rng.integers(0, 256, L, dtype=np.uint8)
        got = O.orc_xxh3_64(ptr(b) if L else None, L)
        assert got == xxhash.xxh3_64_intdigest(b.tobytes()), L


@needs_ref
def test_decision_matches_reference_frames():
    """The backend orc_entropy_decide picks is the node the reference put in the frame."""
    from frames import parse_frame
    rng = np.random.default_rng(11)
    bk_for_trid = {44: C_ORC["CONSTANT"], 29: C_ORC["FLATPACK"], 27: C_ORC["BITPACK"],
                   49: C_ORC["FSE"], 50: C_ORC["HUF"]}
    samples = []
    for n in (1, 2, 10, 50, 300, 4000, 65536):
        samples += [zipf_bytes(rng, n, 1.1), zipf_bytes(rng, n, 3.0, card=2),
                    rng.integers(0, 4, n).astype(np.uint8), np.full(n, 7, np.uint8),
                    rng.integers(0, 256, n).astype(np.uint8), zipf_bytes(rng, n, 0.3, card=40)]
    for x in samples:
        for huf, kind in ((0, C.FSE), (1, C.HUFFMAN)):
            fr = parse_frame(C.ref_compress(kind, x, C.SERIAL, 1))
            dec = fr["chunks"][0]["decoders"]
            # the codec that consumed the input is the last decoder listed
            trid = dec[-1]["trid"] if dec else None
            got = O.orc_entropy_decide(ptr(x), x.size, huf)
            if x.size < 10:
                assert trid is None  # stored because of minStreamSize
                continue
            exp = bk_for_trid.get(trid, C_ORC["STORE"])
            assert got == exp, (x.size, huf, trid)


C_ORC = {"STORE": 0, "CONSTANT": 1, "FLATPACK": 2, "BITPACK": 3, "HUF": 4, "FSE": 5}


# ---------------------------------------------------------------- 2-byte inputs
def h16_cases():
    """16-bit sources: Zipf over spread values, deep (height-limited) trees,
    many ties, two symbols, a dominant symbol with singletons."""
    rng = np.random.default_rng(77)
    out = []

    def spread(n, card, hi, s, perm=True):
        vals = (rng.choice(hi, card, replace=False) if perm else np.arange(card)).astype(np.uint16)
        p = 1.0 / np.arange(1, card + 1) ** s
        p /= p.sum()
        x = np.concatenate([vals, vals[rng.choice(card, n - card, p=p)]])
        rng.shuffle(x)
        return x
    for n, card, hi, s in ((1000, 300, 512, 1.0), (600, 260, 300, 0.3), (32768, 300, 300, 2.0),
                           (60000, 2000, 2000, 1.6), (30000, 1000, 10000, 1.1), (32768, 5000, 65536, 0.9),
                           (99999, 20000, 65536, 1.2), (4096, 4000, 4096, 0.0), (50, 2, 65536, 0.0)):
        out.append(spread(n, card, hi, s))
    x = np.zeros(20000, np.uint16)
    x[rng.choice(20000, 400, replace=False)] = np.arange(1, 401, dtype=np.uint16) * 97
    out.append(x)
    # Fibonacci-like counts: the unlimited tree is 30+ deep
    fib = [1, 1]
    while len(fib) < 26:
        fib.append(fib[-1] + fib[-2])
    out.append(np.repeat(np.arange(26, dtype=np.uint16) * 2000, fib))
    return out


@needs_ref
def test_huf16_tables_encode_decode_match_reference():
    h = C.refh()
    for x in h16_cases():
        count = np.zeros(65536, np.uint32)
        ms, cd = ctypes.c_uint(0), ctypes.c_uint(0)
        O.orc_histogram16(ptr(count), ctypes.byref(ms), ctypes.byref(cd), ptr(x), x.size)
        assert ms.value == int(x.max()) and cd.value == len(np.unique(x))
        assert np.array_equal(count[: ms.value + 1], np.bincount(x, minlength=ms.value + 1))
        m = ms.value
        ca = np.zeros(m + 1, np.uint16); na = np.zeros(m + 1, np.uint8)
        cb = np.zeros(m + 1, np.uint16); nb = np.zeros(m + 1, np.uint8)
        ta = O.orc_huf16_build(ptr(ca), ptr(na), ptr(count), m)
        tb = h.refh_huf16_build(ptr(cb), ptr(nb), ptr(count), m)
        assert ta == tb and ta > 0
        assert np.array_equal(na, nb) and np.array_equal(ca, cb)
        assert na.max() == ta and ((na > 0) == (count[: m + 1] > 0)).all()
        for x4 in (0, 1):
            for cap in (2 * x.size + 32, x.size // 3 + 40):
                ea = np.zeros(cap + 64, np.uint8)
                eb = np.zeros(cap + 64, np.uint8)
                sa = O.orc_huf16_encode(ptr(ea), cap, ptr(x), x.size, ptr(ca), ptr(na), x4)
                sb = h.refh_huf16_encode(ptr(eb), cap, ptr(x), x.size, ptr(cb), ptr(nb), m, tb, x4)
                assert sa == sb
                if not sa:
                    continue
                assert np.array_equal(ea[:sa], eb[:sb])
                w = np.where(na > 0, ta + 1 - na.astype(np.int32), 0).astype(np.uint8)
                back = np.zeros(x.size, np.uint16)
                assert O.orc_huf16_decode(ptr(back), x.size, ptr(ea), sa, ptr(w), w.size, x4) == 0
                assert np.array_equal(back, x)
                rback = np.zeros(x.size, np.uint16)
                assert h.refh_huf16_decode(ptr(rback), x.size, ptr(eb), sb, ptr(w), w.size, ta, x4) == x.size
                assert np.array_equal(rback, x)


@needs_ref
def test_entropy16_matches_reference():
    for x in h16_cases():
        count = np.bincount(x, minlength=65536).astype(np.uint32)
        ms = int(x.max())
        assert O.orc_entropy(ptr(count), ms, x.size) == R.ZL_calculateEntropy(ptr(count), ms, x.size)


def test_tokenize2_closed_form():
    rng = np.random.default_rng(3)
    for n, card in ((1, 1), (10, 3), (5000, 100), (40000, 255), (300, 256)):
        vals = rng.choice(65536, card, replace=False).astype(np.uint16)
        x = np.concatenate([vals, vals[rng.integers(0, card, n - card)]]) if n > card else vals[:n]
        rng.shuffle(x)
        a = np.zeros(256, np.uint16)
        idx = np.zeros(x.size, np.uint8)
        k = O.orc_tokenize2(ptr(a), ptr(idx), ptr(x), x.size)
        u, first = np.unique(x, return_index=True)
        order = u[np.argsort(first)]
        assert k == len(order) and np.array_equal(a[:k], order)
        assert np.array_equal(a[idx], x)
    x = np.arange(300, dtype=np.uint16)
    assert O.orc_tokenize2(ptr(np.zeros(300, np.uint16)), ptr(np.zeros(300, np.uint8)), ptr(x), 300) == 0
