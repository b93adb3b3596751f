"""Build recipe of the native library (libstmtree.so) for gfx950.

Every HIP source is compiled with hipcc for gfx950 only, then linked into one shared
library in the package directory, so the .so travels with the source tree.

    python -m eirene_amd.build            # build (incremental: skips up-to-date objects)
    python -m eirene_amd.build --force    # rebuild everything
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = ROOT / "_build"
LIB = PKG / "libstmtree.so"
ARCH = os.environ.get("STM_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = [
    "stm_sort.hip",
    "stm_fold.hip",
    "stm_tree.hip",
    "stm_learned.hip",
    "stm_scan.hip",
    "stm_api.hip",
    "stm_server.hip",
]
CPP_SOURCES = [
    "stm_gen.cpp",
]
HEADERS = sorted(p.name for p in CSRC.glob("*.hpp")) + ["../../include/stmtree.h"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-munsafe-fp-atomics",
    # Packed-FP32 instructions stay disabled in every HIP object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT / 'include'}",
]
CPP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-pthread", f"-I{ROOT / 'include'}"]


def _newest_header_mtime() -> float:
    return max((CSRC / h).stat().st_mtime for h in HEADERS if (CSRC / h).exists())


def _compile(src: str, force: bool) -> Path:
    obj = BUILD / (src + ".o")
    src_path = CSRC / src
    if (not force and obj.exists() and obj.stat().st_mtime >= src_path.stat().st_mtime
            and obj.stat().st_mtime >= _newest_header_mtime()):
        return obj
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(src_path), "-o", str(obj)]
    else:
        cmd = ["g++", *CPP_FLAGS, "-c", str(src_path), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    # The host pass of hipcc warns about the packed-fp32 feature; keep other warnings visible.
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l
                     and "is not a recognized feature" not in l)
    if warn.strip():
        sys.stderr.write(warn + "\n")
    return obj


def build(force: bool = False, verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = HIP_SOURCES + CPP_SOURCES
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", str(tmp),
           *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    # stm_test: the reference-compatible command line, linked against the library.
    cli = PKG / "stm_test"
    cmd = ["g++", *CPP_FLAGS, "-o", str(cli), str(CSRC / "stm_test.cpp"), f"-L{PKG}", "-lstmtree",
           "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {LIB} and {cli}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    args = ap.parse_args()
    build(force=args.force)
