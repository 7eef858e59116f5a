"""Builds libradiant_hip.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "radiant_hip", "libradiant_hip.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["kernels.hip", "sweep_bfp.hip", "sweep_awd.hip", "em_source.hip", "sweep_hog.hip", "engine.cpp", "batch.cpp", "abi.cpp", "serve.cpp", "pool.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         # packed FP32 instructions are disabled (they have produced wrong results on this hardware)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def build(verbose=True, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs, procs = [], []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc] + FLAGS + (["-x", "hip"] if src.endswith(".cpp") else []) + ["-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode())
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
        if verbose and out:
            msg = "\n".join(l for l in out.decode().splitlines() if "packed-fp32-ops" not in l and l.strip())
            if msg:
                print(msg)
    tmp = OUT + ".tmp"
    link = [hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs + ["-lpthread"]
    subprocess.run(link, check=True)
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
