"""Build liblofloat_hip.so in-tree with hipcc for gfx950.

Every ``csrc/*.hip`` file is compiled to an object under ``csrc/_objs`` and linked into
``lofloat_amd/liblofloat_hip.so``.  The library links against the HIP runtime that torch
ships (same soname as ROCm's), so a process that imported torch first uses one runtime.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(CSRC, "_objs")
LIB = os.path.join(HERE, "liblofloat_hip.so")
ARCH = os.environ.get("LOFLOAT_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-fvisibility=hidden",
    "-fno-gpu-flush-denormals-to-zero",
    # keep the compiler from forming packed FP32 ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"),
    "-I", CSRC,
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: cannot build the lofloat_amd HIP library")


def _torch_lib_dir() -> str | None:
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
        if spec and spec.origin:
            d = os.path.join(os.path.dirname(spec.origin), "lib")
            if os.path.exists(os.path.join(d, "libamdhip64.so")):
                return d
    except Exception:
        pass
    return None


def _compile(src: str, obj: str) -> None:
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        msg = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                        if "packed-fp32-ops" not in l)
        raise RuntimeError(f"hipcc failed for {os.path.basename(src)}:\n{msg}")


def build(verbose: bool = True, jobs: int | None = None) -> str:
    """Compile all HIP sources and link liblofloat_hip.so.  Returns the library path."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJDIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = jobs or min(8, os.cpu_count() or 4, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, o): s for s, o in zip(srcs, objs)}
        for f in cf.as_completed(futs):
            f.result()
            if verbose:
                print(f"[lofloat_amd] compiled {os.path.basename(futs[f])}", file=sys.stderr)
    link = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", LIB + ".tmp", *objs]
    tl = _torch_lib_dir()
    if tl:
        link += [f"-L{tl}", f"-Wl,-rpath,{tl}"]
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}{res.stderr}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"[lofloat_amd] linked {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
