"""Builds the in-tree gfx950 library ed_code_amd/libed_amd.so with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libed_amd.so")
SOURCES = ["ed_amd.hip", "ed_hv_fast.hip", "ed_hv64.hip", "ed_block.hip", "ed_lanczos.hip", "ed_obs.hip",
           "ed_calib.hip", "ed_refapi.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{err}")


OBJDIR = os.path.join(HERE, "_build")


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compiles every source to an object in parallel, then links.  With
    incremental=True the objects are kept in ed_code_amd/_build and a source
    is recompiled only when it or a header is newer than its object."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    cflags = [f for f in FLAGS if f != "-shared"]
    hdrs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    hdrs += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    newest_hdr = max(os.path.getmtime(h) for h in hdrs)

    def compile_link(objdir):
        objs = [os.path.join(objdir, os.path.basename(s) + ".o") for s in srcs]
        jobs = [[HIPCC] + cflags + inc + ["-c", s, "-o", o] for s, o in zip(srcs, objs)
                if not incremental or not os.path.exists(o)
                or os.path.getmtime(o) < max(os.path.getmtime(s), newest_hdr)]
        if jobs:
            with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
                for f in [ex.submit(_run, j, verbose) for j in jobs]:
                    f.result()
        _run([HIPCC] + FLAGS + objs + ["-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib",
                                       "-o", LIB + ".tmp"], verbose)

    if incremental:
        os.makedirs(OBJDIR, exist_ok=True)
        compile_link(OBJDIR)
    else:
        with tempfile.TemporaryDirectory(prefix="edamd_") as tmp:
            compile_link(tmp)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
