"""Build libd2fem_hip.so in-tree: every csrc/*.hip compiled by hipcc for gfx950 (and the
host-only csrc/*.cpp by the host compiler) and linked into one shared library next to this file.

    python -m diff_2d_fem_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
OUT = HERE / "libd2fem_hip.so"
ARCH = os.environ.get("D2FEM_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed FP32 (v_pk_*_f32) instructions: some of their forms have given wrong results
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",               # hardware float64 atomic add
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (Path(cand).exists() or cand == "hipcc"):
            return cand
    raise RuntimeError("hipcc not found")


def gen_headers():
    """Regenerate csrc/scene_dev.h (device mirror of d2f_scene) from include/d2fem.h."""
    sys.path.insert(0, str(CSRC))
    try:
        import gen_scene_dev
    finally:
        sys.path.pop(0)
    return gen_scene_dev.generate()


HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall"]
# sources compiled as several objects, part k with -DD2F_SCENE_PART=k (the scene kernel: its
# float and float64 instantiations)
SPLIT = {"step_scene.hip": 2}


def build(verbose=False):
    """Compile and link; returns the path of the library. Raises on a compiler error. csrc/*.hip
    go through hipcc for gfx950, csrc/*.cpp (host-only code) through the host compiler; the
    sources compile in parallel (at most 8 at a time)."""
    from concurrent.futures import ThreadPoolExecutor
    gen_headers()
    jobs = []
    for src in sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")):
        parts = SPLIT.get(src.name, 1)
        jobs += [(src, k, parts) for k in range(parts)]

    def compile_one(job):
        src, k, parts = job
        obj = CSRC / (src.stem + (f".{k}" if parts > 1 else "") + ".o")
        if src.suffix == ".cpp":
            cmd = [os.environ.get("CXX", "g++"), *HOST_FLAGS, "-c", str(src), "-o", str(obj)]
        else:
            part = [f"-DD2F_SCENE_PART={k}"] if parts > 1 else []
            cmd = [hipcc(), *FLAGS, *part, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compiler failed on {src.name}:\n{r.stderr}")
        return obj
    with ThreadPoolExecutor(max_workers=min(8, max(1, len(jobs)))) as ex:
        objs = list(ex.map(compile_one, jobs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    for o in objs:
        o.unlink()
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
