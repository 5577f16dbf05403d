"""Builds libcortex_rt.so in-tree for gfx950 with hipcc (no JIT cache: the .so travels with the tree)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcortex_rt.so")
SOURCES = ["kernels.hip", "bvh_gpu.hip", "noalloc.hip", "scene_build.cpp", "api.cpp", "serve.cpp", "tune.cpp",
           "obj_loader.cpp", "scene_builder.cpp", "scenefile.cpp", "image_io.cpp"]
CLI = os.path.join(HERE, "crt_render")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# packed FP32 (v_pk_*_f32) is disabled for every device compile
DEVICE_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# no FMA contraction anywhere: every kernel variant (uniform / per-lane scene, one / two levels,
# wavefront / megakernel) then rounds the same expressions the same way, so a request gives the same
# bits whichever variant traced it, and the arithmetic matches the oracle (g++ -ffp-contract=off)
FP_FLAGS = ["-ffp-contract=off"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose=False, jobs=4):
    """Compiles every source to an object (in parallel) and links the shared library."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + DEVICE_FLAGS + FP_FLAGS + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append(out)
        elif verbose and out.strip():
            print(out)
    if failed:
        sys.stderr.write("\n".join(failed) + "\n")
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC] + DEVICE_FLAGS + ["-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread", "-lz"])
    os.replace(tmp, LIB)
    # command-line renderer with the reference executable's flags, linked against the library
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "crt_cli.cpp"), "-o", tmp,
          "-L" + HERE, "-lcortex_rt", "-Wl,-rpath,$ORIGIN", "-lz"])
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
