"""Builds libamt.so (gfx950 HIP + host C++) in-tree with hipcc.

    python -m am_ml_suite_amd.build                  # every source
    python -m am_ml_suite_amd.build --incremental    # only sources with a changed dependency
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libamt.so")

SOURCES = ["amt.hip", "unet.hip", "batch.hip", "simp_py.hip", "ke0.cpp", "topology_optimizer.cpp", "problem.cpp"]
CLI = os.path.join(PKG, "am_topopt")  # the reference's standalone program on this library
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    if verbose and out.strip():
        print(out)


def _up_to_date(obj, dep):
    """True when obj is newer than every file its dependency list (from -MD) names."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    text = open(dep).read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    t = os.path.getmtime(obj)
    return all(os.path.exists(f) and os.path.getmtime(f) <= t for f in files)


def _compile(src, bdir, verbose, force):
    path = os.path.join(CSRC, src)
    obj = os.path.join(bdir, src + ".o")
    dep = obj + ".d"
    if not force and _up_to_date(obj, dep):
        return obj
    cmd = [HIPCC] + COMMON
    if src.endswith(".hip"):
        cmd += ["-x", "hip", "--offload-arch=" + ARCH] + DEVICE_FLAGS
    else:
        cmd += ["-x", "c++"]
    cmd += ["-MD", "-MF", dep, "-c", path, "-o", obj]
    _run(cmd, verbose)
    return obj


def build(verbose=False, force=True):
    """Compiles every source (force) or only those whose object is older than a dependency."""
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, bdir, verbose, force), SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "am_topopt.cpp"), "-o", tmp, "-L", PKG, "-lamt",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, force="--incremental" not in sys.argv))
