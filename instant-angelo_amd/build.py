"""Builds libangelo_amd.so (HIP, gfx950) in-tree.

    python instant-angelo_amd/build.py [-j N]

Every csrc/*.hip is compiled to an object with hipcc and the objects are
linked into angelo_amd/lib/libangelo_amd.so.  Always rebuilds.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "angelo_amd", "lib")
LIB = os.path.join(LIBDIR, "libangelo_amd.so")
ARCH = os.environ.get("ANGELO_AMD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -ffp-contract=off: every a*b+c stays two rounded ops unless the source asks
# for fmaf, which keeps indices / positions bit-exact with the oracle.
# -packed-fp32-ops: no v_pk_*_f32 generation.
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-parameter",
]


# MFMA accumulators in VGPRs rather than AGPRs (on the unified register file
# the AGPRs add to a wave's registers): the radiance forward goes from one to
# two waves a SIMD (0.36 -> 0.23 ms a step).  Not for geometry.hip (the SDF
# forward at three waves a SIMD measured slower, 3.54 -> 3.90 ms) nor
# mlp.hip (the option crashes the compiler there).
FILE_FLAGS = {"radiance.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CFLAGS + FILE_FLAGS.get(os.path.basename(src), []) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    args = ap.parse_args()
    try:
        build(args.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
