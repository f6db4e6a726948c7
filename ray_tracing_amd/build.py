"""Builds the gfx950 path-tracing library (librt_hip.so) in-tree with hipcc.

Every HIP translation unit is compiled for gfx950 only, with FMA contraction
off and correctly rounded FP32 division/sqrt (the CPU oracle runs the same
operation sequences), and without packed-FP32 instructions.  Host-only C++
is compiled by the ROCm clang with the same floating-point flags.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librt_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANGXX = os.path.join(ROCM, "llvm", "bin", "clang++")

HIP_SOURCES = ["rt_bvh.hip", "rt_trace.hip", "rt_wavefront.hip", "rt_post.hip", "rt_nn.hip",
               "rt_nn_api.hip", "rt_cache.hip", "rt_cache_api.hip", "rt_api.hip"]
CXX_SOURCES = ["rt_scenes.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++20", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function"]
HIP_FLAGS = COMMON + ["--offload-arch=gfx950", "-fhip-fp32-correctly-rounded-divide-sqrt",
                      "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang",
                      "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src: str, build_dir: str = BUILD, defines=()) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(build_dir, src + ".o")
    dflags = [f"-D{d}" for d in defines]
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + dflags + ["-c", path, "-o", obj]
    else:
        cmd = [CLANGXX] + COMMON + dflags + ["-I", os.path.join(ROCM, "include"), "-c", path,
                                             "-o", obj]
    out = _run(cmd)
    # the host pass does not know the packed-fp32 feature name: not an error
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")
    return obj


def build(jobs: int = 4, verbose: bool = True, out: str = LIB, defines=()) -> str:
    """Compiles and links the library; `defines` (NAME=VALUE) and `out` build a
    variant (e.g. another LDS stack size) into another file."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    build_dir = BUILD if out == LIB else out + ".objs"
    os.makedirs(build_dir, exist_ok=True)
    sources = HIP_SOURCES + CXX_SOURCES
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _compile(src, build_dir, defines), sources))
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


HOST = os.path.join(HERE, "host")
TESTS_HOST = os.path.join(os.path.dirname(HERE), "tests", "host")
# (output path, sources): the CLI beside the library, the report test binary
# beside its source under tests/
HOST_TARGETS = {"render_realtime": (os.path.join(HERE, "render_realtime"),
                                    [os.path.join(HOST, "render_realtime.cpp"),
                                     os.path.join(HOST, "rt_host.cpp")]),
                "test_report": (os.path.join(TESTS_HOST, "test_report"),
                                [os.path.join(TESTS_HOST, "test_report.cpp"),
                                 os.path.join(HOST, "rt_host.cpp")]),
                "test_host_api": (os.path.join(TESTS_HOST, "test_host_api"),
                                  [os.path.join(TESTS_HOST, "test_host_api.cpp"),
                                   os.path.join(HOST, "rt_host.cpp")]),
                "test_pool_gpu": (os.path.join(TESTS_HOST, "test_pool_gpu"),
                                  [os.path.join(TESTS_HOST, "test_pool_gpu.cpp"),
                                   os.path.join(HOST, "rt_host.cpp")])}
# targets that call the HIP runtime themselves (pointer attributes)
HIP_HOST_TARGETS = {"test_pool_gpu"}


def source_provenance():
    """(revision, SHA-256 of the source state) for the run reports: the git
    commit when the tree is a checkout, and a digest over the library's,
    host layer's and public header's source files (path and contents, in
    sorted order), which also covers uncommitted changes."""
    import hashlib
    root = os.path.dirname(HERE)
    r = subprocess.run(["git", "-C", root, "rev-parse", "HEAD"], stdout=subprocess.PIPE,
                       stderr=subprocess.DEVNULL, text=True)
    revision = r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "unversioned"
    h = hashlib.sha256()
    for d in (CSRC, HOST, os.path.join(root, "include", "rt")):
        for name in sorted(os.listdir(d)):
            path = os.path.join(d, name)
            if os.path.isfile(path) and os.path.splitext(name)[1] in (".h", ".hpp", ".hip", ".cpp"):
                h.update(os.path.relpath(path, root).encode())
                with open(path, "rb") as f:
                    h.update(f.read())
    return revision, h.hexdigest()


def build_host(verbose: bool = True):
    """C++ host API, the render_realtime CLI and the report test, linked
    against librt_hip.so (found next to the executables' parent directory)."""
    outs = []
    revision, state = source_provenance()
    for name, (out, srcs) in HOST_TARGETS.items():
        rpath = os.path.relpath(HERE, os.path.dirname(out))
        cmd = ["g++", "-std=c++20", "-O2", "-Wall", f'-DRT_SOURCE_REVISION="{revision}"',
               f'-DRT_SOURCE_STATE_SHA256="{state}"', "-I", HOST, "-o", out] + srcs
        cmd += ["-L", HERE, "-lrt_hip", "-Wl,-rpath,$ORIGIN/" + rpath]
        if name in HIP_HOST_TARGETS:
            cmd += ["-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
                    "-L", os.path.join(ROCM, "lib"), "-lamdhip64"]
        _run(cmd)
        outs.append(out)
        if verbose:
            print(f"built {out}")
    return outs


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-j", type=int, default=4)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    ap.add_argument("-o", dest="out", default=LIB)
    a = ap.parse_args()
    if a.clean:
        clean()
    build(a.j, out=os.path.abspath(a.out), defines=a.defines)
    if os.path.abspath(a.out) == LIB:
        build_host()
