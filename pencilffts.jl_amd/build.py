"""Build the HIP library of the package in-tree (gfx950 only).

Produces
  lib/libpencilffts_hip.so  -- the product: kernels + plan + C ABI (include/pfft.h)
  lib/libpencilffts_emu.so  -- host emulator of the generic pass (tests only)
  lib/libpfft_bandwidth.so  -- streaming-copy ceiling (benchmark only)

Usage: python pencilffts.jl_amd/build.py [-v]
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions stay off: plain float code only.
DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include")]

PRODUCT_SOURCES = ["kernels.hip", "fast.hip", "fast_nt.hip", "plan.cpp", "capi.cpp"]
EMU_SOURCES = ["emu.cpp", "plan.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(shlex.quote(c) for c in cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    if verbose and (r.stdout or r.stderr):
        # hipcc's host pass warns that it does not know the packed-fp32 feature
        lines = [l for l in (r.stdout + r.stderr).splitlines()
                 if "packed-fp32-ops" not in l and l.strip()]
        if lines:
            print("\n".join(lines))


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    jobs, objs, eobjs = [], [], []
    for src in PRODUCT_SOURCES:
        obj = os.path.join(LIB, src.replace(".", "_") + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        dev = DEVICE_FLAGS if src.endswith(".hip") else []
        jobs.append([HIPCC] + COMMON + dev + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # host emulator (tests only): host compilation of the same phase functions
    for src in EMU_SOURCES:
        obj = os.path.join(LIB, "emu_" + src.replace(".", "_") + ".o")
        jobs.append([HIPCC] + COMMON + ["-x", "hip", "--offload-host-only", "-O2",
                                        "-c", os.path.join(CSRC, src), "-o", obj])
        eobjs.append(obj)
    # streaming-copy ceiling of the benchmark (not part of the library)
    bw_out = os.path.join(LIB, "libpfft_bandwidth.so")
    jobs.append([HIPCC] + COMMON + DEVICE_FLAGS + ["-x", "hip", "-shared",
                                                   os.path.join(CSRC, "bandwidth.hip"),
                                                   "-o", bw_out + ".tmp"])
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        for f in [ex.submit(_run, cmd, verbose) for cmd in jobs]:
            f.result()
    os.replace(bw_out + ".tmp", bw_out)
    out = os.path.join(LIB, "libpencilffts_hip.so")
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC", "-o", out + ".tmp"] + objs, verbose)
    os.replace(out + ".tmp", out)
    eout = os.path.join(LIB, "libpencilffts_emu.so")
    _run([HIPCC, "-shared", "-fPIC", "-o", eout + ".tmp"] + eobjs, verbose)
    os.replace(eout + ".tmp", eout)
    for o in objs + eobjs:
        os.remove(o)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
