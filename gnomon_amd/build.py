"""In-tree build of libgnomon_hip.so for gfx950 (hipcc, ahead of time).

Each source is compiled to its own object in parallel, then the objects are
linked with --no-undefined, so a source missing from SOURCES fails the build
instead of leaving unresolved symbols in the library.
"""
from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCE_NAMES = ("gnm_kernels.hip", "gnm_dense.hip", "gnm_i8.hip", "gnm_bf16.hip", "gnm_sparse.hip", "gnm_plan.hip",
                "gnm_api.hip")
SOURCES = [os.path.join(HERE, "csrc", f) for f in SOURCE_NAMES]
OUTPUT = os.path.join(HERE, "libgnomon_hip.so")
OBJDIR = os.path.join(HERE, "csrc", "_build")

# -packed-fp32-ops off: hipcc -O3 may otherwise form v_pk_{add,mul,fma}_f32 from
# scalar float code; the host pass warns that it does not know the feature.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC", "-Wl,--no-undefined"]

_NOISE = "is not a recognized feature for this target"


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [ln for ln in (res.stdout + res.stderr).splitlines() if _NOISE not in ln]
    return res.returncode, msgs


def _headers():
    hs = [os.path.join(HERE, "csrc", f) for f in os.listdir(os.path.join(HERE, "csrc")) if f.endswith(".hpp")]
    return hs + [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]


def build_library(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source (incremental=True: only sources newer than their
    object or than any header) and link the library."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    inc = "-I" + os.path.join(ROOT, "include")
    objs = [os.path.join(OBJDIR, os.path.basename(s) + ".o") for s in SOURCES]
    newest_header = max(os.path.getmtime(h) for h in _headers())

    def compile_one(so):
        src, obj = so
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header):
            return 0, []
        return _run([hipcc, *HIPCC_FLAGS, inc, "-c", src, "-o", obj])

    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, zip(SOURCES, objs)))
    msgs = [m for _, ms in results for m in ms]
    if any(rc != 0 for rc, _ in results):
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    tmp = OUTPUT + ".tmp"
    rc, lmsgs = _run([hipcc, *LINK_FLAGS, *objs, "-o", tmp])
    msgs += lmsgs
    if rc != 0:
        raise RuntimeError("link failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    import sys
    print(build_library(verbose=True, incremental="--incremental" in sys.argv))
