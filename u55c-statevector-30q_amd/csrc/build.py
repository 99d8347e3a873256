"""Build libsvx.so (gfx950) in-tree with hipcc.  Usage: python csrc/build.py [--jobs N]"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
OUT = os.path.join(PKG, "lib", "libsvx.so")
SOURCES = ["svx_kernels.hip", "svx_mux.hip", "svx_api.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mcode-object-version=5", "-Wno-unused-result",
         "-I", os.path.join(ROOT, "include")]


def _obj(src):
    return os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")


def _compile(src):
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(HERE, src), "-o", _obj(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return _obj(src)


def build(verbose=True):
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    # every kernel's host stub must be defined (an undefined one only shows up at dlopen)
    nm = subprocess.run(["nm", "-C", "--undefined-only", tmp], capture_output=True, text=True)
    missing = [l for l in nm.stdout.splitlines() if "svx::" in l]
    if missing:
        raise RuntimeError("undefined libsvx symbols:\n" + "\n".join(missing))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
