"""Build libphysgpu.so (HIP, gfx950) in-tree: phys-engine_amd/physgpu/libphysgpu.so.

Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked into
one shared library exporting the C ABI of include/physgpu.h.

Flags that matter for results:
  -ffp-contract=off          no a*b+c fusion: the reference's FP32 operation
                             order is replayed exactly (IEEE division and
                             sqrt stay hipcc's correctly rounded defaults)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                             no packed FP32 instructions
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "physgpu", "libphysgpu.so")
# The kernels use gfx950-only instructions (v_permlane16/32_swap, MI355X's
# LDS size): the build is for gfx950 alone.
ARCH = "gfx950"
if os.environ.get("PHYSGPU_ARCH", ARCH) != ARCH:
    raise RuntimeError(f"PHYSGPU_ARCH={os.environ['PHYSGPU_ARCH']}: this library builds for "
                       f"{ARCH} (MI355X) only")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument", "-Wno-unused-value",
    "-Wno-unused-result",
]


def hipcc():
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src)[:-4] + ".o")
    cmd = [hipcc(), *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    warnings = "\n".join(l for l in r.stderr.splitlines()
                         if "packed-fp32-ops" not in l and l.strip())
    return obj, warnings


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, w in results:
            if w:
                print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
