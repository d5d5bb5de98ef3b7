"""Build recipe of the CPU baseline library oracle/libvpm_cpu.so (test infrastructure).

gcc -O3 -ffast-math -fopenmp, targeting x86-64-v3 (AVX2/FMA) so the library
runs on any current x86 host, including the one that holds the GPU.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "vpm_cpu.c"
LIB = HERE / "libvpm_cpu.so"
FLAGS = ["-O3", "-ffast-math", "-fopenmp", "-march=x86-64-v3", "-mtune=generic", "-fPIC",
         "-shared", "-std=c11", "-Wall"]


def build(force: bool = False) -> Path:
    if force or not LIB.exists() or SRC.stat().st_mtime > LIB.stat().st_mtime:
        tmp = LIB.with_suffix(".so.tmp")
        cc = os.environ.get("CC", "gcc")
        r = subprocess.run([cc, *FLAGS, str(SRC), "-o", str(tmp), "-lm"], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"building {LIB.name} failed:\n{r.stderr}")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force=True))
