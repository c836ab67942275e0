"""Builds libozl_hip.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Every call recompiles (the object files are cheap); the library is written to
openzl_amd/libozl_hip.so, next to this file, so that it travels with the tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libozl_hip.so")
GRAPHS_LIB = os.path.join(HERE, "libozl_graphs.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def build(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for src, p in procs:
        out = p.communicate()[0].decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"[build] {os.path.basename(src)} failed:\n{out}\n")
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-Wl,-Bsymbolic", "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    build_graphs()
    return LIB


def build_graphs():
    """User graphs in plain C (graphs/ozl_graphs.c) against the public header,
    linked to libozl_hip.so."""
    tmp = GRAPHS_LIB + ".tmp"
    cmd = ["gcc", "-std=gnu11", "-O2", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           os.path.join(HERE, "graphs", "ozl_graphs.c"), "-L", HERE, "-l:libozl_hip.so",
           "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined", "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("graphs build failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, GRAPHS_LIB)
    return GRAPHS_LIB


if __name__ == "__main__":
    print(build(verbose=True))
