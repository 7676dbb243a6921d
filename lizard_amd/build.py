"""Builds the product library lizard_amd/liblizard_amd.so with hipcc for gfx950 (in-tree, so it travels with the
repository to a GPU machine). Usage: python -m lizard_amd.build [-j N]."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
OUT = os.path.join(HERE, "liblizard_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
SOURCES = ["format.cpp", "api.cpp", "batch.cpp", "register.cpp", "word.cpp", "find.cpp", "shake.cpp", "replay.cpp",
           "ldpc.hip", "spectrum.hip", "refine.hip", "channel.hip", "finder.hip", "xfer.hip", "anchor.hip", "border.hip"]
# -packed-fp32-ops: no packed FP32 instructions are generated (v_pk_*_f32).
# -ffp-contract=off: every float multiply and add rounds on its own, as in every reference build
# (liblizard/build.sh:15-17); the warp, detrend, pilot turn and soft scale reproduce the reference's float order.
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-ffp-contract=off", "-Wall", "-Wno-unused-result", "-fvisibility=hidden"]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(jobs=4):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"), "-Wl,-z,defs"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 4
    print(build(j))
