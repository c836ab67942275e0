"""Learned selector on the host: the CPU restatement of the features and the
MLP (oracle/selector_oracle.py) pinned by closed forms, the chunk generators,
the training export, and the committed model file."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import selector_oracle as SO  # noqa: E402

from openzl_amd import datagen as G  # noqa: E402


def test_bits_closed_form():
    v = np.array([0, 1, 2, 3, 255, 256, (1 << 63), (1 << 64) - 1], np.uint64)
    assert SO._bits(v).tolist() == [0, 1, 2, 2, 8, 9, 64, 64]


def test_zigzag_closed_form():
    w = 2
    d = np.array([0, 1, 0xFFFF, 2, 0xFFFE, 0x7FFF, 0x8000], np.uint64)   # 0, 1, -1, 2, -2, max, min
    assert SO.zigzag(d, w).tolist() == [0, 2, 1, 4, 3, 0xFFFE, 0xFFFF]


def test_features_constant_chunk():
    x = np.full(1000, 0x01020304, np.uint32)
    f = SO.features(x)
    assert f[32] == pytest.approx(2 / 8)           # four equally frequent bytes
    assert f[33] == pytest.approx(2 / 3)
    assert f[34] == 1.0 and f[35] == 0.0 and f[36] == 0.0
    assert f[37] == pytest.approx(25 / 64)         # 0x01020304 has 25 bits
    assert f[38] == 0.0 and f[39] == 0.0
    assert f[:32].sum() == pytest.approx(1.0)
    assert f[0] == pytest.approx(1.0)              # bytes 1..4 all in bin 0


def test_features_linear_chunk():
    x = np.arange(4096, dtype=np.uint16)
    f = SO.features(x)
    assert f[34] == 0.0
    assert f[35] == pytest.approx(2 / 64) and f[36] == pytest.approx(2 / 64)   # zigzag(1) = 2
    assert f[38] == 0.0                            # delta planes: all 1, all 0
    assert f[39] == pytest.approx((8 + 4) / 2 / 8)  # low plane uniform over 256, high over 16


def test_features_uniform_bytes():
    x = np.random.default_rng(1).integers(0, 256, 1 << 20, dtype=np.uint8)
    f = SO.features(x)
    assert f[32] == pytest.approx(1.0, abs=1e-3)
    assert f[39] == pytest.approx(1.0, abs=1e-3)
    assert np.allclose(f[:32], 1 / 32, atol=2e-3)


def test_mlp_oracle_closed_form():
    blob = bytearray((64 * 64 * 2 + 16 * 64) * 2 + (64 + 64 + 16) * 4)
    b3 = np.zeros(16, np.float32)
    b3[:8] = [0, 1, 2, 3, 7, 5, 6, 4]
    blob[-64:] = b3.tobytes()
    lg = SO.mlp(np.random.default_rng(0).random((5, 40)).astype(np.float32), bytes(blob))
    assert np.all(np.argmax(lg, axis=1) == 4)
    # one path: W1[0,0] = 1, W2[0,0] = 2, W3[3,0] = 1 -> logit3 = 2 * bf16(x0)
    W = np.zeros((64 + 64 + 16, 64), np.float32)
    W[0, 0] = 1
    W[64, 0] = 2
    W[128 + 3, 0] = 1
    u = W.view(np.uint32).astype(np.uint64)
    bf = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    blob2 = bf.tobytes() + bytes((64 + 64 + 16) * 4)
    x = np.zeros((1, 40), np.float32)
    x[0, 0] = 0.3
    lg = SO.mlp(x, blob2)
    assert lg[0, 3] == pytest.approx(2 * float(SO.bf16_round(np.float32(0.3))), rel=1e-6)


def test_selector_chunks_are_deterministic():
    g1, c1 = G.selector_chunks(16, seed=9)
    g2, c2 = G.selector_chunks(16, seed=9)
    assert np.array_equal(g1, g2)
    assert all(a.tobytes() == b.tobytes() for a, b in zip(c1, c2))
    assert all(c.nbytes == G.SELECT_CHUNK for c in c1)
    assert [G.SELECT_WIDTHS[g] for g in g1] == [c.dtype.itemsize for c in c1]


def test_export_layout_and_forward():
    from openzl_amd import selector_train as T
    rng = np.random.default_rng(2)
    W1 = rng.normal(0, 0.3, (64, 40)).astype(np.float32)
    W2 = rng.normal(0, 0.3, (64, 64)).astype(np.float32)
    W3 = rng.normal(0, 0.3, (8, 64)).astype(np.float32)
    b1, b2, b3 = (rng.normal(0, 0.1, k).astype(np.float32) for k in (64, 64, 8))
    mean = rng.random(40).astype(np.float32)
    std = (rng.random(40) + 0.5).astype(np.float32)
    blob = T.export(W1, b1, W2, b2, W3, b3, mean, std)
    assert len(blob) == (64 * 64 * 2 + 16 * 64) * 2 + (64 + 64 + 16) * 4
    x = rng.random((32, 40)).astype(np.float32)
    # the folded scaling gives the float32 network on the scaled inputs, up to bf16 rounding
    h = np.maximum(((x - mean) / std) @ W1.T + b1, 0)
    h = np.maximum(h @ W2.T + b2, 0)
    ref = h @ W3.T + b3
    got = SO.mlp(x, blob)
    assert np.abs(got - ref).max() < 0.05 * max(1.0, np.abs(ref).max())
    assert np.allclose(T.forward_bf16(blob, x), got, atol=1e-3 * max(1.0, np.abs(got).max()))


def test_committed_model_file():
    path = os.path.join(ROOT, "openzl_amd", "data", "selector_mlp.bin")
    if not os.path.exists(path):
        pytest.skip("no trained selector model")
    assert os.path.getsize(path) == (64 * 64 * 2 + 16 * 64) * 2 + (64 + 64 + 16) * 4


def test_training_labels_match_reference_sizes():
    """The committed training sizes (GPU exhaustive trial, checksums off) of
    the first training chunks equal the reference library's frame sizes."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ozl_checkers as C
    path = os.path.join(ROOT, "openzl_amd", "data", "selector_train.npz")
    if not os.path.exists(path):
        pytest.skip("no training data")
    if C.refh() is None:
        pytest.skip("reference library not built")
    d = np.load(path)
    n = d["sizes"].shape[0]
    gens, chunks = G.selector_chunks(n, seed=G.SELECT_TRAIN_SEED, limit=48)
    assert np.array_equal(gens, d["gens"][:48])
    kinds = [C.STORE, C.FSE, C.BITPACK, C.DELTA_FSE, C.DELTA_ZZ_BP, C.TS_FSE, C.TS_HUF, C.DELTA_TS_FSE]
    for i, c in enumerate(chunks):
        ref = [len(C.ref_compress(k, c, C.NUMERIC, c.dtype.itemsize, False, False)) for k in kinds]
        assert list(d["sizes"][i]) == ref, i
