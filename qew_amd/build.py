"""Builds libqew_amd.so (HIP kernels + C ABI, gfx950) and the command-line
tools in-tree with hipcc. Usage: python -m qew_amd.build [--jobs N]."""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
APPS = os.path.join(PKG, "apps")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libqew_amd.so")
BIN = os.path.join(PKG, "bin")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("QEW_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-adds unless written explicitly, so the
# FP64 kernels keep the reference's rounding sequence. Packed FP32 math is
# disabled for all device code.
DEVICE_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(jobs=8, apps=True, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    objs = [os.path.join(BUILD, os.path.basename(s)[:-4] + ".o") for s in srcs]

    def one(pair):
        src, obj = pair
        cmd = [HIPCC, *DEVICE_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(one, zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if apps:
        build_apps(jobs)
    return LIB


def hdf5_flags():
    """Compiler flags that give the tools HDF5 output: the HDF5 headers (searched
    after the system's, so nothing else in that prefix shadows them) and the
    library the tools dlopen by default. HDF5 is never linked: the tools load it
    at run time and fall back to an error naming it when it is absent. The
    location is $QEW_HDF5_ROOT, else the first prefix that holds include/hdf5.h
    and lib/libhdf5.so."""
    roots = [os.environ.get("QEW_HDF5_ROOT", ""), "/usr/local", "/opt/conda"]
    for r in roots:
        if r and os.path.exists(os.path.join(r, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(r, "lib", "libhdf5.so")):
            lib = os.path.join(r, "lib", "libhdf5.so")
            return ["-idirafter", os.path.join(r, "include"), f'-DQEW_HDF5_LIBRARY="{lib}"']
    return []


def build_apps(jobs=8):
    if not os.path.isdir(APPS):
        return []
    os.makedirs(BIN, exist_ok=True)
    mains = sorted(f for f in os.listdir(APPS) if f.endswith(".cpp"))
    inc = os.path.join(ROOT, "include")

    def one(f):
        out = os.path.join(BIN, f[:-4])
        _run(["g++", "-O2", "-std=c++17", "-I", inc, "-I", "/opt/rocm/include",
              "-D__HIP_PLATFORM_AMD__", *hdf5_flags(), os.path.join(APPS, f), "-o", out,
              "-L", PKG, "-lqew_amd", "-ldl", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/.."])
        return out

    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        return list(ex.map(one, mains))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-apps", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, not a.no_apps, a.verbose))


if __name__ == "__main__":
    main()
