"""Builds libmoao_hip.so (gfx950) in-tree: moao_amd/_lib/libmoao_hip.so.

    python -m moao_amd.build            # compile every source, link the shared library and
                                        # the command-line driver moao_amd/_lib/moao

hipcc compiles each source of moao_amd/csrc/ to an object (in parallel), then links them.
Every HIP compile carries -Xclang -target-feature -Xclang -packed-fp32-ops so the compiler
emits no packed-FP32 VALU instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(OUT_DIR, "obj")
LIB = os.path.join(OUT_DIR, "libmoao_hip.so")
ARCH = "gfx950"


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), "-c", "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}", "-o", obj]
    if src.endswith(".hip"):
        cmd += ["-x", "hip", src]
    else:
        cmd += [src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the command-line driver (C++ over the C-ABI only), next to the library
    cli = os.path.join(OUT_DIR, "moao")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(HERE, "cli", "moao_main.cpp"), "-o", cli + ".tmp", f"-L{OUT_DIR}",
           "-lmoao_hip", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"cli build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(cli + ".tmp", cli)
    # the same driver under the reference check program's name runs its check pass
    chk = os.path.join(OUT_DIR, "check_moao")
    shutil.copyfile(cli, chk + ".tmp")
    os.chmod(chk + ".tmp", 0o755)
    os.replace(chk + ".tmp", chk)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources and {cli}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
