"""Build libsteam_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: python steam_amd/csrc/build.py [-j N] [--resource-usage]
"""

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_build")
TARGET = os.path.join(PKG, "libsteam_hip.so")
SOURCES = ["util.hip", "noise.hip", "envelope.hip", "regrid.hip", "conv.hip",
           "batch.hip", "scalar.hip"]
# The scalar kernels restate float32 numpy arithmetic operation by operation:
# no a*b+c may be fused into one fma there (the __f*_rn intrinsics are header
# functions, outside the reach of a pragma in the source file).
SOURCE_FLAGS = {"scalar.hip": ["-ffp-contract=off"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src, extra, obj_dir=None):
    obj = os.path.join(obj_dir or OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *SOURCE_FLAGS.get(src, []), *extra, "-c", os.path.join(HERE, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, res


def build(jobs=4, resource_usage=False, verbose=False, defines=(), target=None):
    """Compile every source and link the library.  ``defines`` (NAME=VALUE
    strings) and ``target`` build a measurement variant elsewhere (its
    objects beside it), leaving the in-tree library alone."""
    obj_dir = OBJ if target is None else os.path.join(os.path.dirname(os.path.abspath(target)),
                                                      "_obj")
    os.makedirs(obj_dir, exist_ok=True)
    extra = ["-Rpass-analysis=kernel-resource-usage"] if resource_usage else []
    extra += [f"-D{d}" for d in defines]
    sources = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    objs = []
    failed = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as pool:
        for src, obj, res in pool.map(lambda s: _compile(s, extra, obj_dir), sources):
            if verbose or resource_usage or res.returncode:
                sys.stderr.write(res.stderr)
            if res.returncode:
                failed.append(src)
            objs.append(obj)
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    out = TARGET if target is None else os.path.abspath(target)
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", out + ".tmp"]
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=4)
    ap.add_argument("--resource-usage", action="store_true")
    ap.add_argument("-v", action="store_true")
    ap.add_argument("-D", dest="defines", action="append", default=[],
                    help="NAME=VALUE for a measurement variant (with --out)")
    ap.add_argument("--out", default=None, help="library path of a variant build")
    a = ap.parse_args()
    if a.defines and a.out is None:
        ap.error("-D builds a variant: give its --out path")
    print(build(a.j, a.resource_usage, a.v, a.defines, a.out))
