"""Golden fixtures: reference frames for small cases of every graph of the hot
path (tests/golden/small/), checked on the CPU.

* the recorded inputs regenerate bit-identically from the seeded generators;
* the oracle's universal decoder (tests/oracle_frames.py: oracle FSE / Huffman /
  NCount / delta / zigzag / transpose / bitpack / flatpack / constant / splitN
  decoders, XXH3 checksums) returns every input from its reference frame;
* a mistake planted in a frame (one flipped payload bit) is caught;
* where the reference library is built here, it still writes the same frames.
"""
import numpy as np
import pytest

import golden_cases as GC
import ozl_checkers as C
import oracle_frames as OF

ENTRIES = GC.small_entries()


@pytest.mark.parametrize("e", ENTRIES, ids=[e["name"] for e in ENTRIES])
def test_oracle_decodes_golden_frame(e):
    x = GC.make_input(e["gen"], e["kwargs"])
    assert x.nbytes == e["input_bytes"]
    assert GC.sha256(x.tobytes()) == e["input_sha256"], "generator drifted"
    fr = GC.small_frame(e["name"])
    assert GC.sha256(fr) == e["frame_sha256"]
    out, typ, width = OF.decode_frame(fr)
    assert out == x.tobytes()
    assert typ == e["type"]


def test_planted_bit_flip_is_caught():
    e = next(e for e in ENTRIES if e["name"] == "blocks_huf_0")
    fr = bytearray(GC.small_frame(e["name"]))
    fr[len(fr) // 2] ^= 0x10
    with pytest.raises(Exception):
        out, _, _ = OF.decode_frame(bytes(fr))
        x = GC.make_input(e["gen"], e["kwargs"])
        assert out == x.tobytes()


@pytest.mark.skipif(C.refh() is None, reason="reference library not built here")
def test_reference_still_writes_golden_frames():
    for e in ENTRIES:
        x = GC.make_input(e["gen"], e["kwargs"])
        C.ref_set_min_stream_size(e.get("params", {}).get("minStreamSize", 0))
        try:
            fr = C.ref_compress(e["kind"], x, e["type"], e["width"])
        finally:
            C.ref_set_min_stream_size(0)
        assert GC.sha256(fr) == e["frame_sha256"], e["name"]
