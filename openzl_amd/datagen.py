"""Seeded synthetic inputs of the benchmark configurations (numpy, host).

The same seeds give the same bytes on every machine running this image, so
reference frames computed here and frames computed on the GPU box compare.
"""
import numpy as np

SEEDS = {1: 0x5EED0001, 2: 0x5EED0002, 3: 0x5EED0003, 4: 0x5EED0004, 5: 0x5EED0005}


def timestamps_u32(n=4194304, seed=SEEDS[1]):
    """Sorted uint32 timestamps: start 1,700,000,000, Poisson(1000) increments,
    wrapping modulo 2**32 (the sum passes 2**32 after ~2.6 M elements)."""
    rng = np.random.default_rng(seed)
    inc = rng.poisson(1000, n - 1).astype(np.uint64)
    x = np.empty(n, np.uint64)
    x[0] = 1_700_000_000
    np.cumsum(inc, out=x[1:])
    x[1:] += 1_700_000_000
    return (x & 0xFFFFFFFF).astype(np.uint32)


def random_walk_columns(n_cols=4096, seed=SEEDS[3], min_len=1024, max_len=1 << 20):
    """Columns of the batched-columns configuration: length log-uniform in
    [min_len, max_len], uint16 or uint32 by coin flip, int64 random walk with
    step round(N(0, sigma=2**b)), b uniform in 2..14, truncated to the width."""
    rng = np.random.default_rng(seed)
    cols = []
    for _ in range(n_cols):
        n = int(np.floor(np.exp(rng.uniform(np.log(min_len), np.log(max_len)))))
        width = 2 if rng.integers(0, 2) == 0 else 4
        b = int(rng.integers(2, 15))
        steps = np.rint(rng.normal(0.0, float(2 ** b), n)).astype(np.int64)
        walk = np.cumsum(steps)
        dt = np.uint16 if width == 2 else np.uint32
        cols.append((walk & ((1 << (8 * width)) - 1)).astype(dt))
    return cols


def bf16_normal(n=536870912, seed=SEEDS[4], sigma=0.02):
    """bf16 bit patterns of N(0, sigma) fp32 samples, rounded to nearest even."""
    rng = np.random.default_rng(seed)
    out = np.empty(n, np.uint16)
    step = 1 << 24
    for a in range(0, n, step):
        b = min(n, a + step)
        f = rng.normal(0.0, sigma, b - a).astype(np.float32).view(np.uint32)
        rnd = ((f >> 16) & 1) + 0x7FFF
        out[a:b] = ((f + rnd) >> 16).astype(np.uint16)
    return out
