"""Training of the learned selector's MLP from exhaustive-trial sizes.

    python -m openzl_amd.selector_train collect OUT.npz   (GPU: features and
        exhaustive-trial sizes of the training chunks, seed 0x7EA10001)
    python -m openzl_amd.selector_train fit IN.npz [OUT.bin]   (CPU: trains the
        40 -> 64 -> 64 -> 8 network and writes its parameter block)

The labels are the GPU's exhaustive-trial frame sizes of the eight candidate
graphs (openzl_amd.select). The loss is the expected size the choice costs
above the best candidate (the softmax-weighted regret, in units of the chunk
size) plus a small cross-entropy term on the brute-force choice. Training
runs in float32 on the host; the exported weights are rounded to bf16, the
feature scaling folded into the first layer.
"""
import json
import os
import sys

import numpy as np

from . import datagen as G

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "selector_mlp.bin")


def collect(out, n_chunks=8192, seed=G.SELECT_TRAIN_SEED):
    import torch
    from . import select as S
    gens, chunks = G.selector_chunks(n_chunks, seed=seed)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, G.SELECT_CHUNK)
    feats = sb.features(src).cpu().numpy().copy()
    sizes, choice = sb.trial(src)
    torch.cuda.synchronize()
    np.savez_compressed(out, feats=feats, sizes=sizes.cpu().numpy(), choice=choice.cpu().numpy(),
                        gens=gens.astype(np.int8), raw=np.array([c.nbytes for c in chunks], np.int64))
    print(f"[collect] {n_chunks} chunks -> {out}", flush=True)


def _bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((((u + 0x7FFF + ((u >> 16) & 1)) >> 16)) & 0xFFFF).astype(np.uint16)


def _from_bf16(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


def export(W1, b1, W2, b2, W3, b3, mean, std):
    """Parameter block: W1 bf16[64][64], W2 bf16[64][64], W3 bf16[16][64],
    b1 f32[64], b2 f32[64], b3 f32[16]; the scaling (x - mean) / std folded
    into W1 and b1."""
    W1f = W1 / std[None, :]
    b1f = b1 - W1f @ mean
    A = np.zeros((64, 64), np.float32); A[:, :W1.shape[1]] = W1f
    C = np.zeros((16, 64), np.float32); C[:8] = W3
    b3p = np.zeros(16, np.float32); b3p[:8] = b3
    parts = [_bf16(A).tobytes(), _bf16(W2).tobytes(), _bf16(C).tobytes(), b1f.astype(np.float32).tobytes(),
             b2.astype(np.float32).tobytes(), b3p.tobytes()]
    return b"".join(parts)


def forward_bf16(blob, feats):
    """Logits with the GPU's rounding points (bf16 inputs, weights and hidden
    activations, float accumulation)."""
    o = 0
    raw = np.frombuffer(blob, np.uint8)
    def take(nbytes, dt):
        nonlocal o
        a = np.frombuffer(raw[o:o + nbytes].tobytes(), dt)
        o += nbytes
        return a
    W1 = _from_bf16(take(8192, np.uint16)).reshape(64, 64)
    W2 = _from_bf16(take(8192, np.uint16)).reshape(64, 64)
    W3 = _from_bf16(take(2048, np.uint16)).reshape(16, 64)
    b1, b2, b3 = take(256, np.float32), take(256, np.float32), take(64, np.float32)
    x = np.zeros((feats.shape[0], 64), np.float32)
    x[:, :feats.shape[1]] = _from_bf16(_bf16(feats))
    h = _from_bf16(_bf16(np.maximum(x @ W1.T + b1, 0)))
    h = _from_bf16(_bf16(np.maximum(h @ W2.T + b2, 0)))
    return (h @ W3.T + b3)[:, :8]


def brute_cost(sizes, raw):
    """Frame size the brute-force selector ends with (checksums off)."""
    ok = sizes > 0
    best = np.where(ok, sizes, np.iinfo(np.int64).max).astype(np.int64)
    bi = np.argmin(best, axis=1)
    bv = best[np.arange(len(bi)), bi]
    return np.where(bv < raw, bv, sizes[:, 0])


def ratio_loss(sizes, raw, choice):
    chosen = sizes[np.arange(len(choice)), choice].astype(np.float64)
    ex = brute_cost(sizes, raw).astype(np.float64)
    return 1.0 - ex.sum() / chosen.sum()


def fit(npz, out=DEFAULT_OUT, steps=4000, seed=0):
    import torch
    d = np.load(npz)
    X = d["feats"].astype(np.float32)
    S = d["sizes"].astype(np.int64)
    raw = d["raw"].astype(np.int64)
    bad = S <= 0
    Sf = np.where(bad, raw[:, None] * 4, S).astype(np.float64)
    best = Sf.min(axis=1, keepdims=True)
    regret = (Sf - best) / raw[:, None]
    label = np.argmin(Sf, axis=1)
    mean = X.mean(axis=0)
    std = X.std(axis=0) + 1e-3
    torch.manual_seed(seed)
    Xt = torch.from_numpy((X - mean) / std)
    Rt = torch.from_numpy(regret.astype(np.float32))
    Lt = torch.from_numpy(label)
    net = torch.nn.Sequential(torch.nn.Linear(X.shape[1], 64), torch.nn.ReLU(), torch.nn.Linear(64, 64),
                              torch.nn.ReLU(), torch.nn.Linear(64, 8))
    opt = torch.optim.Adam(net.parameters(), lr=3e-3)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, steps)
    for it in range(steps):
        logits = net(Xt)
        p = torch.softmax(logits, dim=1)
        loss = (p * Rt).sum(dim=1).mean() * 50 + 0.05 * torch.nn.functional.cross_entropy(logits, Lt)
        opt.zero_grad()
        loss.backward()
        opt.step()
        sched.step()
    W = [m for m in net if isinstance(m, torch.nn.Linear)]
    g = lambda t: t.detach().numpy().astype(np.float32)
    blob = export(g(W[0].weight), g(W[0].bias), g(W[1].weight), g(W[1].bias), g(W[2].weight), g(W[2].bias),
                  mean.astype(np.float32), std.astype(np.float32))
    choice = np.argmax(forward_bf16(blob, X), axis=1)
    rl = ratio_loss(S, raw, choice)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with open(out, "wb") as f:
        f.write(blob)
    meta = {"train_chunks": int(X.shape[0]), "steps": steps, "train_ratio_loss": float(rl),
            "train_accuracy_vs_brute_force": float(np.mean(choice == np.asarray(d["choice"])))}
    with open(os.path.splitext(out)[0] + ".json", "w") as f:
        json.dump(meta, f, indent=1)
    print(f"[fit] {out}: {meta}", flush=True)
    return blob


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "collect":
        collect(sys.argv[2])
    elif len(sys.argv) >= 3 and sys.argv[1] == "fit":
        fit(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else DEFAULT_OUT)
    else:
        sys.exit(__doc__)
