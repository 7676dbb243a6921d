"""Builds the product library lizard_amd/liblizard_amd.so with hipcc for gfx950 (in-tree, so it travels with the
repository to a GPU machine). Usage: python -m lizard_amd.build [-j N] [--sanitize].

--sanitize builds a second copy, lizard_amd/_san/liblizard_amd.so, whose host code (not the GPU code) carries
AddressSanitizer and UndefinedBehaviorSanitizer, for the tests that run the host entries under them (load it with the
sanitizer runtime preloaded and LIZARD_AMD_LIB pointing at it)."""
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
SOURCES = ["format.cpp", "api.cpp", "batch.cpp", "register.cpp", "word.cpp", "find.cpp", "shake.cpp", "replay.cpp", "xfer_format.cpp",
           "ldpc.hip", "spectrum.hip", "refine.hip", "channel.hip", "finder.hip", "xfer.hip", "anchor.hip", "border.hip"]
# -packed-fp32-ops: no packed FP32 instructions are generated (v_pk_*_f32).
# -ffp-contract=off: every float multiply and add rounds on its own, as in every reference build
# (liblizard/build.sh:15-17); the warp, detrend, pilot turn and soft scale reproduce the reference's float order.
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-ffp-contract=off", "-Wall", "-Wno-unused-result", "-fvisibility=hidden"]


# host code only: every sanitizer flag after -Xarch_host, so no GPU code is built with it
SAN_FLAGS = ("-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=all "
             "-Xarch_host -fno-omit-frame-pointer").split()
SAN_OBJ = os.path.join(HERE, "_obj_san")
SAN_OUT = os.path.join(HERE, "_san", "liblizard_amd.so")


def _compile(src, san=False):
    obj = os.path.join(SAN_OBJ if san else OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + (SAN_FLAGS if san else []) + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr))
    return obj


def build(jobs=4, sanitize=False):
    obj_dir, out = (SAN_OBJ, SAN_OUT) if sanitize else (OBJ, OUT)
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(os.path.dirname(out), exist_ok=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, sanitize), SOURCES))
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + [
        "-Wl,--version-script=" + os.path.join(CSRC, "exports.map")]
    # the sanitizers' runtime comes from the process that loads the library; the product links with no undefined
    # symbol left
    cmd += (SAN_FLAGS + ["-shared-libasan"]) if sanitize else ["-Wl,-z,defs"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 4
    print(build(j, sanitize="--sanitize" in sys.argv))
