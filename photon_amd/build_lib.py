"""Builds the HIP library photon_amd/libphx.so for gfx950 with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libphx.so")
SOURCES = ["adjoint.hip", "capi.cpp", "bridge.hip", "bsdf_batch.hip", "bvh_build.cpp", "display.hip", "integrator.hip", "noalloc.hip", "pbrt.cpp", "proxy.hip", "proxy_bvh.cpp", "scene_data.cpp",
           "render.cpp", "trace.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # IEEE single precision without contraction: the triangle test and the
    # BSDF round like the reference's x86-64 build.
    "-ffp-contract=off", "-fno-fast-math",
    # No packed FP32 instructions (the compiler would form them from float2 code).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, "_build", src + ".o")
        os.makedirs(os.path.dirname(obj), exist_ok=True)
        cmd = [hipcc, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj] + [f for f in FLAGS if f != "-shared"]
        procs.append((src, cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for src, cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"hipcc failed on {src}:\n{' '.join(cmd)}\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("building libphx.so failed")
    tmp = OUT + ".tmp"
    # zlib: the PNG reader of the pbrt scene loader
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("linking libphx.so failed:\n" + r.stdout.decode(errors="replace"))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
