"""Builds libsmgpu.so (HIP for gfx950 + host C++) in-tree with hipcc.

Usage: python -m seismicmesh_amd._build   (or seismicmesh_amd._build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libsmgpu.so")
ROOT = os.path.dirname(HERE)

# -ffp-contract=off: the FP64 paths reproduce numpy's rounding step by step
# and the predicate filters rely on uncontracted error bounds.
# -packed-fp32-ops off: packed FP32 instructions are avoided on this target.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950",
    "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-I" + os.path.join(ROOT, "include"),
]


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith(".hip") or f.endswith(".cpp"):
            out.append(os.path.join(CSRC, f))
    return out


def build(verbose=False, jobs=8):
    """Compile every source to an object in parallel, then link."""
    objdir = os.path.join(HERE, "_objs")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        lang = ["-x", "hip"]
        cmd = ["hipcc"] + [f for f in FLAGS if f != "-shared"] + lang + ["-c", src, "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait() if procs[0][1].poll() is None else None
            import time
            time.sleep(0.05)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        msg = "\n".join("%s:\n%s" % (s, t) for s, t in failed)
        raise RuntimeError("libsmgpu build failed:\n" + msg)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "-shared", "-fPIC", "--offload-arch=gfx950"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("libsmgpu link failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
