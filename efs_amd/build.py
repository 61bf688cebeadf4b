"""Build steps for the native parts of efs_amd.

``build_engine()`` compiles the HIP engine (efs_amd/csrc/efs_engine.hip, the
single-grid engine, and efs_amd/csrc/efs_batch.hip, the batched form) for
gfx950 into efs_amd/libefs_engine.so; ``build_stream_ref()`` compiles the
benchmark's streaming-rate reference into efs_amd/libefs_stream_ref.so;
``build_napi()`` compiles the Node
N-API addon (efs_amd/napi/efs_napi.c) into efs_amd/js/efs_engine.node.
Both are plain compiler invocations so they run on a machine without a GPU.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ENGINE_SRCS = [os.path.join(HERE, "csrc", f) for f in ("efs_engine.hip", "efs_batch.hip")]
ENGINE_LIB = os.path.join(HERE, "libefs_engine.so")
STREAM_REF_SRC = os.path.join(HERE, "csrc", "stream_ref.hip")
STREAM_REF_LIB = os.path.join(HERE, "libefs_stream_ref.so")
NAPI_SRC = os.path.join(HERE, "napi", "efs_napi.c")
NAPI_OUT = os.path.join(HERE, "js", "efs_engine.node")
NODE_INCLUDE = "/usr/include/node"

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
    "-Wno-unused-result", "-Wno-unused-value",
    # packed FP32 instructions stay off (hipcc's host pass warns that it does
    # not know the feature; that is expected)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_engine(verbose=False):
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", ENGINE_LIB + ".tmp", *ENGINE_SRCS]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("engine build failed:\n" + res.stdout + res.stderr)
    os.replace(ENGINE_LIB + ".tmp", ENGINE_LIB)
    if verbose:
        print("built", ENGINE_LIB)
    return ENGINE_LIB


def build_stream_ref(verbose=False):
    """The benchmark's streaming-rate reference (efs_amd/csrc/stream_ref.hip)
    as its own library; the engine library does not contain it."""
    cmd = [hipcc(), *HIPCC_FLAGS, "-o", STREAM_REF_LIB + ".tmp", STREAM_REF_SRC]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("stream reference build failed:\n" + res.stdout + res.stderr)
    os.replace(STREAM_REF_LIB + ".tmp", STREAM_REF_LIB)
    if verbose:
        print("built", STREAM_REF_LIB)
    return STREAM_REF_LIB


def build_napi(verbose=False):
    """Node addon: C against node_api.h, resolving napi_* from the node
    binary at load time and efs_* from libefs_engine.so (rpath $ORIGIN/..)."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("node headers not found; N-API addon skipped")
        return None
    cc = shutil.which("gcc") or "cc"
    cmd = [cc, "-O2", "-fPIC", "-shared", "-Wall", "-I", NODE_INCLUDE, "-I", os.path.join(ROOT, "include"),
           "-o", NAPI_OUT, NAPI_SRC, "-L", HERE, "-lefs_engine", "-Wl,-rpath,$ORIGIN/..",
           "-Wl,--unresolved-symbols=ignore-in-object-files"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("N-API addon build failed:\n" + res.stdout + res.stderr)
    if verbose:
        print("built", NAPI_OUT)
    return NAPI_OUT


def build_all(verbose=False):
    build_engine(verbose)
    build_stream_ref(verbose)
    if os.path.exists(NAPI_SRC):
        build_napi(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
