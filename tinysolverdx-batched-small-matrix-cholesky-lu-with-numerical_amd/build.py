"""Builds the native parts of tinysolverdx-hip in-tree.

    python build.py [--jobs N] [--incremental]

Outputs (all next to the Python package, so they travel with the source tree):
    tinysolverdx/libtsdx_hip.so   HIP kernels for gfx950 + the C ABI of include/tsdx_hip.h
    tinysolverdx/_core*.so        pybind11 module: the C++ host API over that ABI

Every HIP unit is compiled with -ffp-contract=off (the kernels write each fused
multiply-add explicitly; nothing else may be contracted), IEEE division and
square root (hipcc's default correctly rounded fp32 divide/sqrt), no denormal
flushing, and packed-fp32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys
import sysconfig

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
CSRC = HERE / "csrc"
PKG = HERE / "tinysolverdx"
BUILD = HERE / "_build"
ARCH = os.environ.get("TSDX_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd, log):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(map(str, cmd))}\n{out}")
    if log and out.strip():
        print(out)
    return out


def _includes(path, seen=None):
    """The unit and every local header it includes, recursively."""
    import re
    seen = set() if seen is None else seen
    if path in seen or not path.exists():
        return seen
    seen.add(path)
    for inc in re.findall(r'#include\s+"([^"]+)"', path.read_text()):
        _includes((path.parent / inc).resolve(), seen)
    return seen


def _stale(obj, deps):
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def hip_units():
    """Every HIP translation unit: kernels (*.hip) and the C ABI (capi*.cpp)."""
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("capi*.cpp"))


def build(jobs=None, incremental=False, verbose=False):
    """Compiles every HIP unit for gfx950, links libtsdx_hip.so and the _core module."""
    jobs = jobs or max(1, min(os.cpu_count() or 1, 16))
    BUILD.mkdir(exist_ok=True)
    units = hip_units()
    objs = []
    todo = []
    for u in units:
        o = BUILD / (u.name + ".o")
        objs.append(o)
        if not incremental or _stale(o, sorted(_includes(u.resolve()))):
            extra = ["-x", "hip"] if u.suffix == ".cpp" else []
            todo.append([HIPCC] + HIP_FLAGS + extra + ["-I", str(ROOT / "include"), "-c", str(u), "-o", str(o)])
    if todo:
        print(f"[build] compiling {len(todo)} HIP units with {jobs} jobs", flush=True)
        with cf.ThreadPoolExecutor(jobs) as ex:
            futs = {ex.submit(_run, c, verbose): c for c in todo}
            for f in cf.as_completed(futs):
                f.result()
    lib = PKG / "libtsdx_hip.so"
    tmp = PKG / "libtsdx_hip.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs], verbose)
    os.replace(tmp, lib)
    build_core(verbose)
    print(f"[build] {lib.relative_to(ROOT)} and _core ready", flush=True)
    return lib


def build_core(verbose=False):
    """pybind11 `_core` (host C++ only, compiled with the system C++ compiler)."""
    import pybind11

    host = CSRC / "host"
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = PKG / f"_core{ext}"
    tmp = PKG / f"_core{ext}.tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-ffp-contract=off", "-pthread",
           f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}", f"-I{host}",
           f"-I{ROOT / 'include'}",
           str(host / "runtime.cpp"), str(host / "bindings.cpp"), str(host / "device_arrays.cpp"),
           f"-L{PKG}", "-ltsdx_hip", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    for p in PKG.glob("*.so"):
        p.unlink()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
        sys.exit(0)
    build(a.jobs, a.incremental, a.verbose)
