"""Builds the HIP step library (libmeteolab.so, gfx950) in-tree with hipcc.

Run as ``python -m meteolab_amd.build`` or through ``__graft_entry__.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "meteolab.hip")]
LIB = os.path.join(HERE, "libmeteolab.so")
ARCH = os.environ.get("METEOLAB_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # No packed FP32 instructions: plain scalar float code only.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-value", "-Wno-unused-function",
    "-I", os.path.join(ROOT, "include"),
]


def build(verbose=False):
    """Compiles libmeteolab.so; raises CalledProcessError with the compiler's output on failure."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = LIB + ".tmp"
    cmd = [hipcc, *FLAGS, "-o", tmp, *SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(line for line in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in line)
    if res.returncode != 0:
        print(out, file=sys.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, output=out)
    if verbose and out.strip():
        print(out, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


NODE_SRC = os.path.join(ROOT, "node", "meteolab_node.cc")
NODE_ADDON = os.path.join(ROOT, "node", "meteolab.node")
NODE_INCLUDE = os.environ.get("NODE_INCLUDE", "/usr/include/node")


def build_node():
    """Compiles the N-API addon of the C ABI (node/meteolab.node) against libmeteolab.so with g++, when Node's
    headers are installed; returns its path, or None without them."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        return None
    cmd = ["g++", "-std=c++17", "-O2", "-shared", "-fPIC", "-Wall", "-DNODE_GYP_MODULE_NAME=meteolab", "-I", NODE_INCLUDE,
           "-o", NODE_ADDON + ".tmp", NODE_SRC, "-L", HERE, "-lmeteolab", "-Wl,-rpath,$ORIGIN/../meteolab_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        print(res.stdout + res.stderr, file=sys.stderr)
        raise subprocess.CalledProcessError(res.returncode, cmd, output=res.stdout + res.stderr)
    os.replace(NODE_ADDON + ".tmp", NODE_ADDON)
    return NODE_ADDON


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_node())
