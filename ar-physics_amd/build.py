"""Build step of the package: compiles the gfx950 code object and libarp.so in-tree.

    python ar-physics_amd/build.py

1. hipcc --genco --offload-arch=gfx950 csrc/kernels.hip  -> csrc/_build/kernels.hsaco
2. the code object is embedded (csrc/embed.S, .incbin) into the host runtime
   csrc/runtime.cpp, compiled with g++ against the HIP runtime API, and linked with the
   host-only scene-file reader (csrc/toml.cpp, csrc/scene_file.cpp) and the per-GPU scratch
   array (csrc/scratch.hip, hipcc)                                        -> arp/libarp.so
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
OUT = os.path.join(HERE, "arp", "libarp.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIPCC_FLAGS = [
    "--genco",
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    # packed FP32 instructions stay off (hand-checked scalar FP32 only)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    co = os.path.join(BUILD, "kernels.hsaco")
    run([hipcc] + HIPCC_FLAGS + ["-o", co, os.path.join(CSRC, "kernels.hip")])
    emb = os.path.join(BUILD, "embed.o")
    run(["gcc", "-c", "-DARP_CO_PATH=\"%s\"" % co, "-o", emb, os.path.join(CSRC, "embed.S")])
    rt = os.path.join(BUILD, "runtime.o")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
         "-I" + os.path.join(ROCM, "include"), "-c", "-o", rt, os.path.join(CSRC, "runtime.cpp")])
    # host-only scene files: no contraction, so the f32 arithmetic is the reference's
    host_objs = []
    for name in ("toml", "scene_file"):
        o = os.path.join(BUILD, name + ".o")
        run(["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-ffp-contract=off", "-fno-fast-math",
             "-I" + os.path.join(ROCM, "include"), "-c", "-o", o, os.path.join(CSRC, name + ".cpp")])
        host_objs.append(o)
    # the fixed per-GPU scratch array of the allocation-free entry points: a HIP translation
    # unit of its own, so the HIP runtime registers the array and places it on each device
    scr = os.path.join(BUILD, "scratch.o")
    run([hipcc, "-c", "-fPIC", "--offload-arch=gfx950", "-O2", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-o", scr, os.path.join(CSRC, "scratch.hip")])
    tmp = OUT + ".tmp"
    run(["g++", "-shared", "-o", tmp, rt, scr] + host_objs + [emb, "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
         "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-Wl,--no-undefined"])
    shutil.move(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
