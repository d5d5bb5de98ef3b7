"""Build recipe of the CPU baseline libraries (test infrastructure).

gcc -O3 -ffast-math -fopenmp, twice:
  oracle/libvpm_cpu.so     -march=x86-64-v3 (AVX2/FMA): runs on any current x86 host;
  oracle/libvpm_cpu_v4.so  -march=x86-64-v4 (AVX-512 F/BW/CD/DQ/VL): loaded instead
                           when the host advertises those (cpu_baseline.load).
Both are built here and travel with the tree, so the GPU machine's host (an
EPYC with AVX-512) runs the wider build without compiling anything.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "vpm_cpu.c"
LIB = HERE / "libvpm_cpu.so"
LIB_V4 = HERE / "libvpm_cpu_v4.so"
COMMON = ["-O3", "-ffast-math", "-fopenmp", "-mtune=generic", "-fPIC", "-shared", "-std=c11",
          "-Wall"]
ISAS = {LIB: "x86-64-v3", LIB_V4: "x86-64-v4"}
AVX512 = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")


def host_has_avx512() -> bool:
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags.update(line.split(":", 1)[1].split())
                    break
        return all(x in flags for x in AVX512)
    except OSError:
        return False


def build(force: bool = False) -> Path:
    for lib, isa in ISAS.items():
        if force or not lib.exists() or SRC.stat().st_mtime > lib.stat().st_mtime:
            tmp = lib.with_suffix(".so.tmp")
            cc = os.environ.get("CC", "gcc")
            r = subprocess.run([cc, *COMMON, f"-march={isa}", str(SRC), "-o", str(tmp), "-lm"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"building {lib.name} failed:\n{r.stderr}")
            os.replace(tmp, lib)
    return LIB


def runtime_lib() -> tuple[Path, str]:
    """The library this host runs and its ISA level."""
    if host_has_avx512() and LIB_V4.exists():
        return LIB_V4, ISAS[LIB_V4]
    return LIB, ISAS[LIB]


if __name__ == "__main__":
    print(build(force=True))
