"""Learned-selector batch through the C ABI (ozl_select_trial,
ozl_select_compress, ozl_select_features, ozl_select_mlp).

* exhaustive trial: every candidate's frame size with checksums off equals
  the reference library's frame size for that candidate graph, on chunks of
  all eight generators (64 KiB and smaller) and on tiny chunks of every width
  around minStreamSize;
* compression with a given candidate: frames byte-identical to the
  reference's frames for that graph, for every candidate;
* trial + compression: frames byte-identical to the reference's brute-force
  selector over the same eight graphs, and they decode back with the
  reference decoder;
* features and MLP against the CPU restatement (oracle/selector_oracle.py);
* the committed model's selection on fresh chunks stays close to exhaustive.
"""
import os
import sys

import numpy as np
import pytest

import ozl_checkers as C

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import selector_oracle as SO  # noqa: E402

pytestmark = pytest.mark.gpu

REF_KINDS = [C.STORE, C.FSE, C.BITPACK, C.DELTA_FSE, C.DELTA_ZZ_BP, C.TS_FSE, C.TS_HUF, C.DELTA_TS_FSE]
DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.fixture(scope="module")
def env():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from openzl_amd import select as S
    from openzl_amd import datagen as G
    return torch, S, G


def chunk_set(G, seed=77):
    _, big = G.selector_chunks(24, seed=seed)
    _, mid = G.selector_chunks(16, seed=seed + 1, nbytes=4000)
    _, small = G.selector_chunks(16, seed=seed + 2, nbytes=200)
    rng = np.random.default_rng(seed)
    tiny = []
    for w in (1, 2, 4, 8):
        for n in (1, 2, 3, 5, 6, 9, 10, 11, 13, 20, 40):
            hi = 1 << min(8 * w, 62)
            tiny.append(rng.integers(0, hi, n, dtype=np.uint64).astype(DT[w]))
            tiny.append(np.cumsum(rng.integers(0, 3, n)).astype(DT[w]))
    return big + mid + small + tiny


def frames_of(dst, frame_off):
    fo = frame_off.cpu().numpy()
    raw = dst[: int(fo[-1])].cpu().numpy()
    return [raw[fo[i]:fo[i + 1]].tobytes() for i in range(len(fo) - 1)]


def test_trial_sizes_match_reference(env):
    torch, S, G = env
    if C.refh() is None:
        pytest.skip("reference library not built")
    chunks = chunk_set(G)
    src, spans, widths, _ = S.pack_chunks(chunks)
    print(f"select trial: K={len(chunks)}", flush=True)
    sb = S.SelectorBatch(spans, widths, 65536)
    sizes, choice = sb.trial(src)
    torch.cuda.synchronize()
    got = sizes.cpu().numpy()
    bad = []
    for i, c in enumerate(chunks):
        for k, kind in enumerate(REF_KINDS):
            ref = len(C.ref_compress(kind, c, C.NUMERIC, c.dtype.itemsize, False, False))
            if got[i, k] != ref:
                bad.append((i, c.dtype, c.size, k, int(got[i, k]), ref))
    assert not bad, bad[:12]
    # brute-force choice
    ch = choice.cpu().numpy()
    for i, c in enumerate(chunks):
        best, bi = c.nbytes, 0
        for k in range(8):
            if 0 < got[i, k] < best:
                best, bi = got[i, k], k
        assert ch[i] == bi, (i, ch[i], bi)


def test_compress_each_candidate_matches_reference(env):
    torch, S, G = env
    if C.refh() is None:
        pytest.skip("reference library not built")
    chunks = chunk_set(G, seed=91)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    for k, kind in enumerate(REF_KINDS):
        print(f"select compress: candidate {k}", flush=True)
        choice = torch.full((len(chunks),), k, dtype=torch.uint8, device="cuda")
        dst, fo, status = sb.compress(src, choice)
        torch.cuda.synchronize()
        assert int(status.abs().sum()) == 0, status.cpu().numpy()
        got = frames_of(dst, fo)
        bad = []
        for i, c in enumerate(chunks):
            ref = C.ref_compress(kind, c, C.NUMERIC, c.dtype.itemsize)
            if got[i] != ref:
                j = next((q for q in range(min(len(ref), len(got[i]))) if ref[q] != got[i][q]), None)
                bad.append((i, c.dtype, c.size, len(got[i]), len(ref), j))
        assert not bad, (k, bad[:8])


def test_selected_frames_match_brute_force(env):
    torch, S, G = env
    if C.refh() is None:
        pytest.skip("reference library not built")
    chunks = chunk_set(G, seed=5)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    _, choice = sb.trial(src)
    choice = choice.clone()
    dst, fo, status = sb.compress(src, choice)
    torch.cuda.synchronize()
    assert int(status.abs().sum()) == 0
    got = frames_of(dst, fo)
    for i, c in enumerate(chunks):
        ref = C.ref_compress(C.SELECT8, c, C.NUMERIC, c.dtype.itemsize)
        assert got[i] == ref, (i, c.dtype, c.size, int(choice[i]))
        out, _, _ = C.ref_decompress(got[i], c.nbytes + 64)
        assert out == c.tobytes(), i


def test_features_match_oracle(env):
    torch, S, G = env
    chunks = chunk_set(G, seed=3)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    f = sb.features(src).cpu().numpy()
    ref = np.stack([SO.features(c) for c in chunks])
    err = np.abs(f - ref)
    assert err.max() < 2e-4, (np.unravel_index(err.argmax(), err.shape), err.max())


def test_mlp_matches_oracle(env):
    torch, S, G = env
    from openzl_amd._lib import lib
    rng = np.random.default_rng(0)
    nbytes = lib().ozl_select_mlp_params_bytes()
    W1 = rng.normal(0, 0.3, (64, 64)).astype(np.float32); W1[:, 40:] = 0
    W2 = rng.normal(0, 0.2, (64, 64)).astype(np.float32)
    W3 = rng.normal(0, 0.2, (16, 64)).astype(np.float32); W3[8:] = 0
    b = rng.normal(0, 0.1, 144).astype(np.float32)
    bf = lambda a: ((a.view(np.uint32).astype(np.uint64) + 0x7FFF + ((a.view(np.uint32) >> 16) & 1)) >> 16).astype(np.uint16)
    blob = bf(W1).tobytes() + bf(W2).tobytes() + bf(W3).tobytes() + b.tobytes()
    assert len(blob) == nbytes
    K = 1000   # not a multiple of 16
    feats = rng.random((K, 40)).astype(np.float32)
    ft = torch.from_numpy(feats).cuda()
    params = torch.from_numpy(np.frombuffer(blob, np.uint8).copy()).cuda()
    widths = torch.ones(K, dtype=torch.uint8, device="cuda")
    spans = torch.zeros(2 * K, dtype=torch.int64, device="cuda")
    sb = S.SelectorBatch(spans, widths, 64)
    sb.feats.copy_(ft)
    from openzl_amd._lib import check
    st = torch.cuda.current_stream().cuda_stream
    check(lib().ozl_select_mlp(sb.feats.data_ptr(), K, params.data_ptr(), sb.choice.data_ptr(), sb.logits.data_ptr(),
                               st), "ozl_select_mlp")
    torch.cuda.synchronize()
    got = sb.logits.cpu().numpy()
    ref = SO.mlp(feats, blob)
    # the oracle rounds at the same points (bf16 inputs, weights and hidden
    # activations); only the accumulation (fp32 MFMA against fp64) differs:
    # the logits stay within one bf16 step of their own magnitude
    err = np.abs(got - ref)
    ulp = np.exp2(np.floor(np.log2(np.maximum(np.abs(ref), 2.0 ** -6))) - 7)
    print(f"mlp: max |err| {err.max():.3g}, max err in logit bf16 steps {(err / ulp).max():.2f}", flush=True)
    assert (err <= ulp).all(), (err / ulp).max()
    ch = sb.choice.cpu().numpy()
    srt = np.sort(ref, axis=1)
    clear = srt[:, -1] - srt[:, -2] > 0.05
    assert np.all(ch[clear] == np.argmax(ref, axis=1)[clear])


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "openzl_amd", "data", "selector_mlp.bin")),
                    reason="no trained selector model")
def test_learned_selection_close_to_exhaustive(env):
    torch, S, G = env
    _, chunks = G.selector_chunks(512, seed=0x5EED0105)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    sizes, ex_choice = sb.trial(src)
    sizes = sizes.cpu().numpy().astype(np.int64)
    ex_choice = ex_choice.cpu().numpy()
    choice, _ = sb.predict(src, S.load_params())
    ch = choice.cpu().numpy()
    r = np.arange(len(chunks))
    loss = 1 - sizes[r, ex_choice].sum() / sizes[r, ch].sum()
    print(f"learned selection ratio loss {loss:.4%}", flush=True)
    assert loss <= 0.01


def test_full_selection_matches_brute_force_digests(env):
    """The learned-selection workload's 8,192 x 64 KiB chunks: exhaustive trial
    then compression on the GPU give the reference brute-force selector's
    frames (per-frame digests in tests/golden/full/)."""
    torch, S, G = env
    import golden_cases as GC
    meta = GC.full_meta()
    if meta is None or "select_8192_brute_force" not in meta:
        pytest.skip("no full-size digests")
    m = meta["select_8192_brute_force"]
    _, chunks = G.selector_chunks(**m["kwargs"])
    print(f"full select: {len(chunks)} chunks", flush=True)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    _, choice = sb.trial(src)
    dst, fo, status = sb.compress(src, choice.clone())
    torch.cuda.synchronize()
    assert int(status.abs().sum()) == 0
    frames = frames_of(dst, fo)
    import hashlib
    got = np.zeros((len(frames), 2), np.uint64)
    for i, fr in enumerate(frames):
        got[i, 0] = len(fr)
        got[i, 1] = np.frombuffer(hashlib.sha256(fr).digest()[:8], np.uint64)[0]
    ref = GC.full_table("select_8192_brute_force")
    bad = np.nonzero((got != ref).any(axis=1))[0]
    assert bad.size == 0, bad[:10]


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "openzl_amd", "data", "selector_mlp.bin")),
                    reason="no trained selector model")
def test_learned_selector_graph_in_the_engine(env):
    """ZL_Compressor_registerSelectorGraph with ozl_learned_selector over the
    eight candidates: ZL_CCtx_compress of each chunk writes the frame of the
    candidate the batch MLP picks for it (ozl_select_mlp + ozl_select_compress),
    and the frame decodes back."""
    torch, S, G = env
    from openzl_amd import api
    _, chunks = G.selector_chunks(48, seed=0x5EED0111)
    src, spans, widths, _ = S.pack_chunks(chunks)
    sb = S.SelectorBatch(spans, widths, 65536)
    choice, _ = sb.predict(src, S.load_params())
    dst, fo, status = sb.compress(src, choice.clone())
    torch.cuda.synchronize()
    assert int(status.abs().sum()) == 0
    want = frames_of(dst, fo)
    comp = api.Compressor(api.LEARNED8)
    dec = api.Decompressor()
    print(f"learned selector graph: {len(chunks)} chunks", flush=True)
    for i, c in enumerate(chunks):
        fr = comp.compress(c, api.NUMERIC, c.dtype.itemsize)
        assert fr == want[i], (i, int(choice[i]))
        out, _, _ = dec.decompress(fr, c.nbytes + 64)
        assert out == c.tobytes(), i
