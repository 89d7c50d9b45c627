"""Build libedge_hip.so (gfx950) in-tree with hipcc.

Usage: python edge-flow-maps_amd/csrc/build.py [--jobs N] [--out PATH] [-D KNOB=VALUE ...]
Each .hip translation unit is compiled to an object in parallel, then linked
into edge-flow-maps_amd/edge_hip/libedge_hip.so (or PATH).  Always recompiles.
Tuning knobs (see edge_device.h): EDGE_WPE_MARCH, EDGE_WPE_FACES;
a library built with other knobs is loaded by setting EDGE_HIP_LIBRARY.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "edge_hip", "libedge_hip.so")
OBJDIR = os.path.join(PKG, "_build")

ARCH = os.environ.get("EDGE_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-munsafe-fp-atomics",
    # packed FP32 (v_pk_*_f32) off: ordinary float code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def sources():
    return sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".hip"))


def compile_one(src, objdir=OBJDIR, defines=()):
    obj = os.path.join(objdir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + [f"-D{d}" for d in defines] + ["-I", HERE, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(jobs=None, verbose=True, out=OUT, defines=()):
    """Compile every .hip source and link `out`; `defines` are NAME=VALUE
    tuning knobs passed as -D to every translation unit."""
    objdir = OBJDIR if out == OUT else out + ".objs"
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda f: compile_one(f, objdir, defines), srcs))
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out} from {len(srcs)} sources", file=sys.stderr)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--out", default=OUT, help="output library path")
    ap.add_argument("-D", dest="defines", action="append", default=[], help="tuning knob NAME=VALUE")
    a = ap.parse_args()
    build(a.jobs, out=os.path.abspath(a.out), defines=a.defines)
