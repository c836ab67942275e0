"""Loader of the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from the reference library)."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def small_entries():
    with open(os.path.join(GOLDEN, "small.json")) as f:
        return json.load(f)


def make_input(gen, kwargs):
    from openzl_amd import datagen as G
    kw = dict(kwargs)
    if "dtype" in kw:
        kw["dtype"] = np.dtype(kw["dtype"])
    return np.ascontiguousarray(getattr(G, gen)(**kw))


def small_frame(name):
    with open(os.path.join(GOLDEN, "small", name + ".zl"), "rb") as f:
        return f.read()


def sha256(b):
    return hashlib.sha256(b).hexdigest()


def full_meta():
    p = os.path.join(GOLDEN, "full.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def full_table(name):
    return np.load(os.path.join(GOLDEN, "full", name + ".npy"))


def frame_digest(fr):
    return int(np.frombuffer(hashlib.sha256(fr).digest()[:8], np.uint64)[0])
