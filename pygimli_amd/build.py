"""Build libpgert.so (all HIP sources under csrc/) for gfx950, in-tree."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libpgert.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(glob.glob(os.path.join(SRC, "*.hip")))


def build(verbose=True, jobs=8):
    """Compile every .hip file to an object in parallel, then link."""
    os.makedirs(OUT_DIR, exist_ok=True)
    objdir = os.path.join(OUT_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in sources():
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", s, "-o", o]
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = False
    for s, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        # hipcc's host pass warns about the device-only target feature; drop that noise
        lines = [ln for ln in txt.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed = True
            print(f"[pgert build] FAILED {os.path.basename(s)}\n" + "\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines), file=sys.stderr)
    if failed:
        raise RuntimeError("libpgert build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"[pgert build] {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
