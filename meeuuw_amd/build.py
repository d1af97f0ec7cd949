"""Build the HIP library libmeeuuw_hip.so (gfx950) in-tree.

Usage: python -m meeuuw_amd.build   (or meeuuw_amd.build.build()).
Every translation unit under csrc/ is compiled with hipcc, in parallel, and
linked into meeuuw_amd/libmeeuuw_hip.so next to this file.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmeeuuw_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: FE/PIC kernels round exactly as written (no implicit FMA).
# packed-fp32-ops disabled: packed FP32 instructions are avoided on this target.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
PER_FILE = {
    # solver kernels may contract a*b+c into FMA: they are tolerance-checked
    "mw_solver.hip": ["-ffp-contract=fast"],
    "mw_cnn.hip": ["-ffp-contract=fast"],
    "mw_energy.hip": ["-ffp-contract=fast"],
}


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    flags = list(COMMON)
    extra = PER_FILE.get(os.path.basename(src), [])
    if "-ffp-contract=fast" in extra:
        flags = [f for f in flags if f != "-ffp-contract=off"]
    cmd = [HIPCC, *flags, *extra, "-I", CSRC, "-I", INCLUDE, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
