"""Build ``libasuka_jk.so`` in-tree with hipcc for gfx950.

Each ``csrc/*.hip`` translation unit is compiled to an object in parallel,
then linked into ``asuka_amd/libasuka_jk.so``. Flags:

* ``--offload-arch=gfx950``: CDNA4 only;
* ``-munsafe-fp-atomics``: hardware FP64 atomic add for the J/K/F outputs
  (plain device memory, where it is exact IEEE addition);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 VALU ops.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
# ASUKA_BUILD_DIR puts the objects and the library of a variant build
# elsewhere (with extra hipcc flags from ASUKA_BUILD_FLAGS); default in-tree
_OUT = os.environ.get("ASUKA_BUILD_DIR")
LIB = os.path.join(_OUT, "libasuka_jk.so") if _OUT else os.path.join(HERE, "libasuka_jk.so")
OBJDIR = os.path.join(_OUT, "_objs") if _OUT else os.path.join(HERE, "_objs")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[str]:
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    extra = os.environ.get("ASUKA_BUILD_FLAGS", "").split()
    cmd = [hipcc(), *HIPCC_FLAGS, *extra, "-I", CSRC, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def _local_includes(path: str, seen: set) -> None:
    """Adds path and the files its #include "..." lines name, recursively."""
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return
    seen.add(path)
    with open(path) as f:
        for line in f:
            t = line.strip()
            if t.startswith("#include") and '"' in t:
                _local_includes(os.path.join(os.path.dirname(path), t.split('"')[1]), seen)


def _stale(src: str) -> bool:
    """An object is stale when it is older than its source or any header it
    includes (directly or not) from this tree."""
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    if not os.path.exists(obj):
        return True
    deps: set = set()
    _local_includes(src, deps)
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in deps)


def build(jobs: int | None = None, incremental: bool = False) -> str:
    """Compile every HIP source (with incremental=True only those whose object
    is stale) and link the shared library; returns its path."""
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if _stale(s)] if incremental else srcs
    jobs = jobs or min(max(1, len(todo)), os.cpu_count() or 1, 16)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_compile, todo))
    objs = [os.path.join(OBJDIR, os.path.basename(s)[:-4] + ".o") for s in srcs]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    import sys

    print(build(incremental="-i" in sys.argv[1:]))
