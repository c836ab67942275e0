"""Parity of the numeric HIP kernels (delta, zigzag, transpose_split, bitpack)
with the CPU oracle, called through the C ABI of libozl_hip.so.

Sizes cover the reference's kernel tests (empty, one element, wraparound
values around 1<<(w-1): tests/unittest/transforms/test_delta_kernel.cpp:11-64,
test_zigzag_kernel.cpp:10-38; widths 1..8 and nbElts {0..10,20,100,1000}:
test_transpose_kernel.cpp:10-104; all nbBits: tests/unittest/common/
test_bitpack.cpp:13-115) plus large ragged sizes, bit-exact.
"""
import ctypes

import numpy as np
import pytest

from ozl_checkers import oracle, ptr

pytestmark = pytest.mark.gpu

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
SIZES = [0, 1, 2, 3, 10, 17, 20, 100, 1000, 4097, 65536 + 5, 1 << 20]


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from openzl_amd import _lib
    return torch, _lib.lib(), _lib


def dev(torch, a):
    t = torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).copy()).cuda()
    return t


def host(t, dtype, n):
    return t.cpu().numpy().view(dtype)[:n]


def sample(rng, n, w):
    dt = DT[w]
    x = rng.integers(0, np.iinfo(dt).max, n, dtype=dt, endpoint=True)
    specials = [0, 1, (1 << (8 * w - 1)) - 1, 1 << (8 * w - 1), np.iinfo(dt).max]
    x[: min(n, len(specials))] = specials[: min(n, len(specials))]
    return x


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_delta_roundtrip_and_parity(env, w):
    torch, L, lib = env
    O = oracle()
    rng = np.random.default_rng(w)
    st = torch.cuda.current_stream().cuda_stream
    for n in SIZES:
        print(f"delta w={w} n={n}", flush=True)
        x = sample(rng, n, w)
        dx = dev(torch, x) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
        dd = torch.zeros(max(n * w, 16), dtype=torch.uint8, device="cuda")
        df = torch.zeros(16, dtype=torch.uint8, device="cuda")
        lib.check(L.ozl_delta_encode(dx.data_ptr(), n, w, dd.data_ptr(), df.data_ptr(), st), "delta_encode")
        exp_d = np.zeros(max(n, 1), DT[w])
        exp_f = np.zeros(1, DT[w])
        O.orc_delta_encode(ptr(exp_d), ptr(exp_f), ptr(x), n, w)
        torch.cuda.synchronize()
        if n:
            assert host(df, DT[w], 1)[0] == exp_f[0]
            assert np.array_equal(host(dd, DT[w], n - 1), exp_d[: n - 1])
        out = torch.zeros(max(n * w, 16), dtype=torch.uint8, device="cuda")
        scratch = torch.zeros(L.ozl_delta_decode_scratch_size(n, w) + 16, dtype=torch.uint8, device="cuda")
        lib.check(L.ozl_delta_decode(df.data_ptr(), dd.data_ptr(), n, w, out.data_ptr(), scratch.data_ptr(), st),
                  "delta_decode")
        torch.cuda.synchronize()
        assert np.array_equal(host(out, DT[w], n), x)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_zigzag_parity(env, w):
    torch, L, lib = env
    O = oracle()
    rng = np.random.default_rng(10 + w)
    st = torch.cuda.current_stream().cuda_stream
    for n in SIZES:
        x = sample(rng, n, w)
        dx = dev(torch, x) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
        dz = torch.zeros(max(n * w, 16), dtype=torch.uint8, device="cuda")
        lib.check(L.ozl_zigzag_encode(dx.data_ptr(), n, w, dz.data_ptr(), st), "zigzag_encode")
        exp = np.zeros(max(n, 1), DT[w])
        O.orc_zigzag_encode(ptr(exp), ptr(x), n, w)
        torch.cuda.synchronize()
        assert np.array_equal(host(dz, DT[w], n), exp[:n])
        back = torch.zeros_like(dz)
        lib.check(L.ozl_zigzag_decode(dz.data_ptr(), n, w, back.data_ptr(), st), "zigzag_decode")
        torch.cuda.synchronize()
        assert np.array_equal(host(back, DT[w], n), x)


@pytest.mark.parametrize("w", [1, 2, 3, 4, 7, 8])
def test_transpose_parity(env, w):
    torch, L, lib = env
    O = oracle()
    rng = np.random.default_rng(20 + w)
    st = torch.cuda.current_stream().cuda_stream
    for n in [0, 1, 5, 10, 20, 100, 1000, 4099, 1 << 20]:
        src = rng.integers(0, 256, n * w, dtype=np.uint8)
        ds = dev(torch, src) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
        dp = torch.zeros(max(n * w, 16), dtype=torch.uint8, device="cuda")
        lib.check(L.ozl_transpose_split(ds.data_ptr(), n, w, dp.data_ptr(), st), "transpose_split")
        exp = np.zeros(max(n * w, 1), np.uint8)
        O.orc_transpose_split(ptr(exp), ptr(src), n, w)
        torch.cuda.synchronize()
        assert np.array_equal(dp.cpu().numpy()[: n * w], exp[: n * w])
        back = torch.zeros_like(dp)
        lib.check(L.ozl_transpose_join(dp.data_ptr(), n, w, back.data_ptr(), st), "transpose_join")
        torch.cuda.synchronize()
        assert np.array_equal(back.cpu().numpy()[: n * w], src)


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_bitpack_parity_all_widths(env, w):
    torch, L, lib = env
    O = oracle()
    rng = np.random.default_rng(30 + w)
    st = torch.cuda.current_stream().cuda_stream
    for n in [0, 1, 5, 20, 100, 1000, 70001]:
        for nb in range(1, 8 * w + 1):
            if n > 1000 and nb % 7 not in (0, 1):
                continue
            x = sample(rng, n, w)
            if nb < 8 * w:
                x = (x & DT[w]((1 << nb) - 1)).astype(DT[w])
            dx = dev(torch, x) if n else torch.zeros(16, dtype=torch.uint8, device="cuda")
            size = (n * nb + 7) // 8
            dp = torch.zeros(size + 16, dtype=torch.uint8, device="cuda")
            lib.check(L.ozl_bitpack_encode(dx.data_ptr(), n, w, nb, dp.data_ptr(), st), "bitpack_encode")
            exp = np.zeros(size + 1, np.uint8)
            O.orc_bitpack_encode(ptr(exp), ptr(x), n, w, nb)
            torch.cuda.synchronize()
            assert np.array_equal(dp.cpu().numpy()[:size], exp[:size]), (n, nb)
            back = torch.zeros(max(n * w, 16), dtype=torch.uint8, device="cuda")
            lib.check(L.ozl_bitpack_decode(dp.data_ptr(), n, w, nb, back.data_ptr(), st), "bitpack_decode")
            torch.cuda.synchronize()
            assert np.array_equal(host(back, DT[w], n), x), (n, nb)
        if n:
            mx = torch.zeros(1, dtype=torch.int64, device="cuda")
            x = sample(rng, n, w)
            dx = dev(torch, x)
            lib.check(L.ozl_max_async(dx.data_ptr(), n, w, mx.data_ptr(), st), "max")
            torch.cuda.synchronize()
            assert int(mx.cpu().numpy().view(np.uint64)[0]) == int(x.max())
