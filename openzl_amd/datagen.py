"""Seeded synthetic inputs of the benchmark configurations (numpy, host).

The same seeds give the same bytes on every machine running this image, so
reference frames computed here and frames computed on the GPU box compare.
"""
import numpy as np

SEEDS = {1: 0x5EED0001, 2: 0x5EED0002, 3: 0x5EED0003, 4: 0x5EED0004, 5: 0x5EED0005}
BLOCK = 65536


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


def zipf_bytes(n, s=1.1, card=256, seed=1):
    """n u8 symbols, Zipf(s) over `card` ranks mapped to random byte values."""
    rng = np.random.default_rng(seed)
    p = 1.0 / np.arange(1, card + 1) ** s
    p /= p.sum()
    perm = rng.permutation(256)[:card]
    return perm[rng.choice(card, n, p=p)].astype(np.uint8)


def walk_column(n, width, b, seed=1):
    """One column of the batched-columns configuration with fixed length, width and b."""
    rng = np.random.default_rng(seed)
    steps = np.rint(rng.normal(0.0, float(2 ** b), n)).astype(np.int64)
    walk = np.cumsum(steps)
    dt = np.uint16 if width == 2 else np.uint32
    return (walk & ((1 << (8 * width)) - 1)).astype(dt)


def uniform_ints(n, hi, dtype, seed=1):
    rng = np.random.default_rng(seed)
    return rng.integers(0, hi, n).astype(dtype)


def u16_spread(n, card, hi, s=1.1, seed=1, perm=True):
    """n uint16 values over `card` distinct values (each present at least
    once, drawn from [0, hi) when perm, else 0..card-1), the rest Zipf(s)
    over them, shuffled."""
    rng = np.random.default_rng(seed)
    vals = (rng.choice(hi, card, replace=False) if perm else np.arange(card)).astype(np.uint16)
    p = 1.0 / np.arange(1, card + 1) ** s
    p /= p.sum()
    x = np.concatenate([vals[:min(card, n)], vals[rng.choice(card, max(n - card, 0), p=p)]])
    rng.shuffle(x)
    return x


def u16_fib(scale=2000):
    """Values k*scale repeated Fibonacci(k) times (26 values): an unlimited
    Huffman tree of depth ~25."""
    fib = [1, 1]
    while len(fib) < 26:
        fib.append(fib[-1] + fib[-2])
    return np.repeat(np.arange(26, dtype=np.uint16) * scale, fib)


def u16_mixed(seed=1):
    """200,000 uint16 values whose 32,768-element chunks differ: constant,
    small values (bitpack), 100 values (tokenize), 1,000 values (16-bit
    Huffman), uniform noise (store), a height-limited tree, then the rest."""
    rng = np.random.default_rng(seed)
    c = 32768
    parts = [np.full(c, 4242, np.uint16),
             rng.integers(0, 512, c).astype(np.uint16),
             u16_spread(c, 100, 4000, 1.1, seed + 1),
             u16_spread(c, 1000, 10000, 1.1, seed + 2),
             rng.integers(0, 65536, c).astype(np.uint16),
             u16_spread(c, 300, 300, 2.0, seed + 3, perm=False),
             u16_spread(200000 - 6 * c, 700, 65536, 1.3, seed + 4)]
    return np.concatenate(parts)


def zipf_table(seed, s=1.1, card=256):
    """uint32 inverse-CDF thresholds and the rank -> symbol permutation."""
    p = 1.0 / np.arange(1, card + 1, dtype=np.float64) ** s
    cdf = np.cumsum(p / p.sum())
    thr = np.minimum(np.floor(cdf * 4294967296.0), 4294967295).astype(np.uint64)
    thr[-1] = 4294967296
    rng = np.random.default_rng(seed)
    perm = rng.permutation(card).astype(np.uint8)
    return thr, perm


def splitmix_np(idx, seed):
    with np.errstate(over="ignore"):
        z = idx.astype(np.uint64) + np.uint64((seed * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def zipf_blocks_np(first_block, n_blocks, seed=SEEDS[2]):
    """Blocks [first_block, first_block + n_blocks) of the entropy-stress input:
    u8 symbols, Zipf(s=1.1) over 256 ranks permuted by the seed, one splitmix64
    draw per byte (counter-based, so any block range is generated alone)."""
    thr, perm = zipf_table(seed)
    idx = np.arange(first_block * BLOCK, (first_block + n_blocks) * BLOCK, dtype=np.uint64)
    u = splitmix_np(idx, seed) >> np.uint64(32)
    rank = np.searchsorted(thr, u, side="right")
    return perm[rank]


def zipf_blocks_torch(torch, n_blocks, seed=SEEDS[2], device="cuda"):
    """Same values as zipf_blocks_np(0, n_blocks), generated on the GPU in slices."""
    thr, perm = zipf_table(seed)
    thr_t = torch.from_numpy(thr.astype(np.int64)).to(device)
    perm_t = torch.from_numpy(perm).to(device)
    total = n_blocks * BLOCK
    out = torch.empty(total, dtype=torch.uint8, device=device)
    M64 = (1 << 64) - 1

    def s64(v):
        v &= M64
        return v - (1 << 64) if v >= (1 << 63) else v

    add = s64(seed * 0x9E3779B97F4A7C15)
    m1 = s64(0xBF58476D1CE4E5B9)
    m2 = s64(0x94D049BB133111EB)

    def lsr(x, k):  # logical shift right on int64
        return (x >> k) & ((1 << (64 - k)) - 1)

    step = 1 << 26
    for a in range(0, total, step):
        b = min(total, a + step)
        z = torch.arange(a, b, dtype=torch.int64, device=device) + add
        z = (z ^ lsr(z, 30)) * m1
        z = (z ^ lsr(z, 27)) * m2
        z = z ^ lsr(z, 31)
        u = lsr(z, 32)
        rank = torch.searchsorted(thr_t, u, right=True)
        out[a:b] = perm_t[rank]
    return out




# ---------------------------------------------------------------- learned-selector chunks
SELECT_TRAIN_SEED = 0x7EA10001
SELECT_CHUNK = 65536
SELECT_GENERATORS = ("timestamps_u32", "zipf_u16", "bf16_normal", "uniform_u8", "const_runs_u32",
                     "sawtooth_u16", "ids_u64", "walk_i32")
SELECT_WIDTHS = (4, 2, 2, 1, 4, 2, 8, 4)


def _select_chunk(g, rng, nbytes):
    w = SELECT_WIDTHS[g]
    n = nbytes // w
    if g == 0:    # sorted uint32 timestamps, Poisson(1000) increments (mod 2**32)
        x = np.cumsum(rng.poisson(1000, n).astype(np.uint64)) + np.uint64(rng.integers(1_600_000_000, 1_800_000_000))
        return (x & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    if g == 1:    # Zipf(s=1.2) categoricals: 1,000 distinct uint16 codes
        p = 1.0 / np.arange(1, 1001, dtype=np.float64) ** 1.2
        codes = rng.choice(65536, 1000, replace=False).astype(np.uint16)
        return codes[rng.choice(1000, n, p=p / p.sum())]
    if g == 2:    # bf16 N(0, 1), round to nearest even
        f = rng.standard_normal(n).astype(np.float32).view(np.uint32)
        return ((f + ((f >> 16) & 1) + 0x7FFF) >> 16).astype(np.uint16)
    if g == 3:    # uint8 uniform
        return rng.integers(0, 256, n, dtype=np.uint8)
    if g == 4:    # uint32 constant runs, geometric run length of mean 64
        runs = rng.geometric(1.0 / 64, n)
        k = int(np.searchsorted(np.cumsum(runs), n)) + 1
        vals = rng.integers(0, 1 << 32, k, dtype=np.uint64).astype(np.uint32)
        return np.repeat(vals, runs[:k])[:n]
    if g == 5:    # uint16 sawtooth plus noise
        period = int(rng.integers(16, 2049))
        slope = int(rng.integers(1, 65))
        base = int(rng.integers(0, 65536))
        i = np.arange(n, dtype=np.int64)
        v = base + (i % period) * slope + rng.integers(-4, 5, n)
        return (v & 0xFFFF).astype(np.uint16)
    if g == 6:    # uint64 IDs: fixed high 44 bits, 20 random low bits
        hi = np.uint64(int(rng.integers(0, 1 << 44)) << 20)
        return hi | rng.integers(0, 1 << 20, n, dtype=np.uint64)
    # int32 Gaussian random walk, step round(N(0, 2**b)), b in 0..12
    b = int(rng.integers(0, 13))
    walk = np.cumsum(np.rint(rng.normal(0.0, float(2 ** b), n)).astype(np.int64))
    return (walk & 0xFFFFFFFF).astype(np.uint32)


def selector_chunks(n_chunks=8192, seed=SEEDS[5], nbytes=SELECT_CHUNK, limit=None):
    """Chunks of the learned-selector configuration: each from one of the eight
    generators chosen uniformly; returns (generator ids, list of arrays).
    `limit` generates only the first chunks of the same sequence."""
    rng = np.random.default_rng(seed)
    gens = rng.integers(0, 8, n_chunks)
    if limit is not None:
        gens = gens[:limit]
    return gens, [_select_chunk(int(g), rng, nbytes) for g in gens]
