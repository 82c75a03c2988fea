"""Build libvcl_amd.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m vcl_amd.build [--force]
Each .hip source compiles to an object in vcl_amd/_build/ (parallel), then links into
vcl_amd/libvcl_amd.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvcl_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VCL_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-munsafe-fp-atomics",
    # no packed-FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-result",
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    deps = [src] + glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*"))
    if os.path.exists(obj) and all(os.path.getmtime(obj) >= os.path.getmtime(d) for d in deps):
        return obj
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(force: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    if force:
        for o in glob.glob(os.path.join(OUT_DIR, "*.o")):
            os.remove(o)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    if force or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
        os.replace(tmp, LIB)
    return LIB


TOOLS = {
    # tests/cpp/preset_fit: host program on the drop-in layer header (include/vcl_layers.hpp)
    os.path.join(ROOT, "tests", "cpp", "preset_fit"): os.path.join(ROOT, "tests", "cpp", "preset_fit.cpp"),
    # tests/cpp/preset_pose: the reference's pose preset (OBJ, SSIM, Rotation / Camera grads)
    os.path.join(ROOT, "tests", "cpp", "preset_pose"): os.path.join(ROOT, "tests", "cpp", "preset_pose.cpp"),
    # tests/cpp/device_rates: measured atomic and VALU peak rates (the scatter / shading kernels' ceilings)
    os.path.join(ROOT, "tests", "cpp", "device_rates"): os.path.join(ROOT, "tests", "cpp", "device_rates.hip"),
}


def build_tools(force: bool = False) -> list:
    """Host programs linked against libvcl_amd.so (rpath to vcl_amd/)."""
    out = []
    for exe, src in TOOLS.items():
        deps = [src, LIB] + glob.glob(os.path.join(ROOT, "include", "*"))
        if force or not os.path.exists(exe) or any(os.path.getmtime(d) > os.path.getmtime(exe) for d in deps):
            rel = os.path.relpath(HERE, os.path.dirname(exe))
            cmd = [HIPCC, "-O2", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                   "-I", os.path.join(ROOT, "include"), src, "-o", exe, "-L", HERE, "-lvcl_amd", f"-Wl,-rpath,$ORIGIN/{rel}"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"tool build failed for {src}:\n{r.stderr[-6000:]}")
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
    print(build_tools(force="--force" in sys.argv))
