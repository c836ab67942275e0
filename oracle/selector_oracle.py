"""CPU restatement (numpy) of the learned selector's features and MLP.

Test infrastructure only: the GPU kernels (ozl_learned.hip) are checked
against it, and it is never part of the product path.

The reference has no such features or network (its selector analogue is the
GBT model of src/openzl/compress/selectors/ml/mlselector.c with the features
of features.c:158-242); these functions restate this project's own
definitions, which the tests pin with closed forms (constant, uniform and
linear inputs) before comparing the GPU against them.
"""
import numpy as np

N_FEATURES = 40


def _entropy(counts):
    counts = counts[counts > 0].astype(np.float64)
    if counts.size == 0:
        return 0.0
    p = counts / counts.sum()
    return float(-(p * np.log2(p)).sum())


def features(x):
    """40 features of one numeric chunk (numpy array of uint8/16/32/64)."""
    x = np.ascontiguousarray(x)
    w = x.dtype.itemsize
    n = x.size
    b = x.view(np.uint8)
    f = np.zeros(N_FEATURES, np.float64)
    h = np.bincount(b, minlength=256)
    f[0:32] = h.reshape(32, 8).sum(axis=1) / b.size
    f[32] = _entropy(h) / 8
    f[33] = {1: 0, 2: 1, 4: 2, 8: 3}[w] / 3
    xv = x.astype(np.uint64)
    planes = b.reshape(n, w)
    f[39] = np.mean([_entropy(np.bincount(planes[:, k], minlength=256)) for k in range(w)]) / 8
    bits_x = _bits(xv)
    f[37] = bits_x.max() / 64
    if n >= 2:
        mask = np.uint64((1 << (8 * w)) - 1) if w < 8 else np.uint64(0xFFFFFFFFFFFFFFFF)
        with np.errstate(over="ignore"):
            d = (xv[1:] - xv[:-1]) & mask
        z = zigzag(d, w)
        bz = _bits(z)
        f[34] = np.count_nonzero(d == 0) / d.size
        f[35] = bz.mean() / 64
        f[36] = bz.max() / 64
        dp = d.astype({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[w]).view(np.uint8).reshape(n - 1, w)
        f[38] = np.mean([_entropy(np.bincount(dp[:, k], minlength=256)) for k in range(w)]) / 8
    return f


def zigzag(d, w):
    """zigzag of w-byte two's complement values held in uint64."""
    d = d.astype(np.uint64)
    sign = (d >> np.uint64(8 * w - 1)) & np.uint64(1)
    mask = np.uint64((1 << (8 * w)) - 1) if w < 8 else np.uint64(0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        return ((d << np.uint64(1)) ^ (np.uint64(0) - sign)) & mask


def _bits(v):
    """bits(0) = 0, else 1 + floor(log2 v), for uint64 arrays (exact)."""
    v = v.astype(np.uint64)
    out = np.zeros(v.shape, np.int64)
    for s in (32, 16, 8, 4, 2, 1):
        big = v >= (np.uint64(1) << np.uint64(s))
        out += np.where(big, s, 0)
        v = np.where(big, v >> np.uint64(s), v)
    return out + (v > 0)


def bf16_round(a):
    """float32 -> nearest-even bfloat16, returned as float32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack_params(blob):
    """Parameter block -> (W1 [64,64], W2 [64,64], W3 [16,64], b1, b2, b3) as float32."""
    blob = np.frombuffer(bytes(blob), np.uint8)
    def bf(a, n):
        return (np.frombuffer(blob[a:a + 2 * n].tobytes(), np.uint16).astype(np.uint32) << 16).view(np.float32)
    o = 0
    W1 = bf(o, 4096).reshape(64, 64); o += 8192
    W2 = bf(o, 4096).reshape(64, 64); o += 8192
    W3 = bf(o, 1024).reshape(16, 64); o += 2048
    b1 = np.frombuffer(blob[o:o + 256].tobytes(), np.float32); o += 256
    b2 = np.frombuffer(blob[o:o + 256].tobytes(), np.float32); o += 256
    b3 = np.frombuffer(blob[o:o + 64].tobytes(), np.float32)
    return W1, W2, W3, b1, b2, b3


def mlp(feats, blob):
    """Logits (K x 8) with the GPU's rounding points: inputs and hidden
    activations rounded to bf16, products accumulated in float64 (the GPU
    accumulates in fp32), biases in fp32."""
    W1, W2, W3, b1, b2, b3 = unpack_params(blob)
    x = np.zeros((feats.shape[0], 64), np.float32)
    x[:, :N_FEATURES] = bf16_round(feats.astype(np.float32))
    h = np.maximum(x.astype(np.float64) @ W1.T.astype(np.float64) + b1, 0)
    h = bf16_round(h.astype(np.float32)).astype(np.float64)
    h = np.maximum(h @ W2.T.astype(np.float64) + b2, 0)
    h = bf16_round(h.astype(np.float32)).astype(np.float64)
    return (h @ W3.T.astype(np.float64) + b3)[:, :8]
