"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU.

1. The C/OpenMP FP64 port (oracle/vpm_cpu.c) with a small driver
   (tests/sanitize/cpu_port_checks.c): every function, all and strided
   targets, a coincident pair.
2. The host side of libvpm_hip's C ABI (argument validation, workspace
   layout, status codes): every csrc/*.hip file compiled with the
   sanitizers on the host code only (-Xarch_host; the device code is built
   at -O0 and never runs), linked with tests/sanitize/abi_checks.cpp, which
   calls every entry point with malformed arguments on a machine without a
   GPU and checks the documented status.

Builds go to tests/sanitize/_build and are reused while the sources are
unchanged (the HIP build takes about a minute).
"""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"
BUILD = SAN / "_build"
CSRC = ROOT / "openonda_amd" / "csrc"
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")


def _stale(target: Path, sources) -> bool:
    return not target.exists() or any(p.stat().st_mtime > target.stat().st_mtime for p in sources)


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    return r.returncode, r.stdout + r.stderr


def _clean(out: str):
    assert "ERROR: AddressSanitizer" not in out, out[-4000:]
    assert "runtime error:" not in out, out[-4000:]


def test_cpu_port_under_sanitizers():
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    BUILD.mkdir(exist_ok=True)
    exe = BUILD / "cpu_port_checks"
    srcs = [ROOT / "oracle" / "vpm_cpu.c", SAN / "cpu_port_checks.c"]
    if _stale(exe, srcs):
        rc, out = _run([gcc, "-O1", "-g", "-fopenmp", "-fno-omit-frame-pointer",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        *map(str, srcs), "-o", str(exe), "-lm"])
        assert rc == 0, out
    rc, out = _run([str(exe)], env=ENV)
    _clean(out)
    assert rc == 0 and "ok: 0 non-finite" in out, out[-2000:]


def test_abi_host_code_under_sanitizers():
    hipcc = Path("/opt/rocm/bin/hipcc")
    if not hipcc.exists():
        pytest.skip("hipcc not available")
    BUILD.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.h")) + [ROOT / "include" / "vpm_hip.h"]
    flags = ["--offload-arch=gfx950", "-O0", "-g0", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             f"-I{ROOT / 'include'}", f"-I{CSRC}"]
    objs, procs = [], []
    for src in sorted(CSRC.glob("*.hip")):
        obj = BUILD / (src.stem + ".san.o")
        objs.append(obj)
        if _stale(obj, [src, *headers]):
            procs.append((src, subprocess.Popen([str(hipcc), *flags, "-c", str(src), "-o", str(obj)],
                                                stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                                                text=True)))
    for src, p in procs:
        out = p.communicate()[0]
        assert p.returncode == 0, f"{src.name}: {out[-3000:]}"
    drv = BUILD / "abi_checks.o"
    exe = BUILD / "abi_checks"
    clangxx = "/opt/rocm/llvm/bin/clang++"
    if _stale(drv, [SAN / "abi_checks.cpp", ROOT / "include" / "vpm_hip.h"]):
        rc, out = _run([clangxx, "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
                        f"-I{ROOT / 'include'}", "-c", str(SAN / "abi_checks.cpp"), "-o", str(drv)])
        assert rc == 0, out
    if _stale(exe, [drv, *objs]):
        rc, out = _run([str(hipcc), "--offload-arch=gfx950", "-Xarch_host", "-fsanitize=address",
                        "-Xarch_host", "-fsanitize=undefined", str(drv), *map(str, objs),
                        "-o", str(exe)])
        assert rc == 0, out
    rc, out = _run([str(exe)], env=ENV)
    _clean(out)
    assert rc == 0, out[-4000:]
    assert "\nok " in out and "0 failed" in out, out[-2000:]
