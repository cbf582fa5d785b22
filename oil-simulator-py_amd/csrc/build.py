"""Build recipe of libimpes_hip.so (gfx950 only).

Compiles every ``*.hip`` translation unit of this directory with hipcc for
``--offload-arch=gfx950`` and links them into
``oil-simulator-py_amd/impes_hip/libimpes_hip.so``.

Usage: ``python oil-simulator-py_amd/csrc/build.py [--incremental] [-j N]``
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
BUILD_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(PKG, "impes_hip", "libimpes_hip.so")
HEADERS = [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".h")] + [
    os.path.join(PKG, "..", "include", "impes_hip.h")
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CFLAGS = [
    "-O3",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-std=c++17",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[str]:
    return sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".hip"))


def _obj(src: str) -> str:
    return os.path.join(BUILD_DIR, os.path.basename(src) + ".o")


def _stale(src: str) -> bool:
    obj = _obj(src)
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + HEADERS
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _compile(src: str) -> tuple[str, int, str]:
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", _obj(src)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    return src, p.returncode, p.stdout + p.stderr


def build(incremental: bool = False, jobs: int | None = None, verbose: bool = False) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = sources()
    todo = [s for s in srcs if (not incremental) or _stale(s)]
    jobs = jobs or min(len(todo) or 1, os.cpu_count() or 4, 16)
    errors = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for src, rc, log in ex.map(_compile, todo):
            if rc != 0:
                errors.append(f"--- {os.path.basename(src)}\n{log}")
            elif verbose and log.strip():
                print(f"--- {os.path.basename(src)}\n{log}")
    if errors:
        raise RuntimeError("HIP compilation failed:\n" + "\n".join(errors))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *[_obj(s) for s in srcs]]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, OUT)
    return OUT


def clean() -> None:
    shutil.rmtree(BUILD_DIR, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    try:
        print(build(a.incremental, a.j, a.v))
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
