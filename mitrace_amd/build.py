"""Builds the gfx950 Tracer library ``mitrace_amd/lib/libmitrace_gpu.so`` with hipcc, and the
command-line renderer ``mitrace_amd/bin/mitrace_gpu_cli`` over it with g++.

Every translation unit is compiled with ``-ffp-contract=off`` (the traversal and shading
arithmetic follow a fixed operation order, see mt_device.hpp) and with packed-FP32 code
generation disabled.  ``python -m mitrace_amd.build`` or ``build()``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libmitrace_gpu.so")
SOURCES = ["mt_scene.cpp", "mt_render.hip", "mt_capi.cpp", "mt_batch.cpp", "mt_mlp.hip", "mt_bvh_build.hip"]
ARCH = os.environ.get("MITRACE_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return ["-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
            "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-result"]


def build(verbose=False, jobs=4, extra_flags=(), lib_path=None):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmd = [cc] + flags() + list(extra_flags) + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    errors = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            errors.append(f"--- {src}\n{out.decode(errors='replace')}")
        elif verbose and out:
            print(out.decode(errors="replace"))
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    out_lib = lib_path or LIB
    tmp = out_lib + ".tmp"
    cmd = [cc, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out_lib)
    if lib_path is None:
        build_cli(verbose)
    return out_lib


BINDIR = os.path.join(HERE, "bin")
CLI = os.path.join(BINDIR, "mitrace_gpu_cli")


def build_cli(verbose=False):
    """The command-line renderer (cli/mitrace_gpu_cli.cpp): g++ over the C++ host headers,
    linked against the library (found next to it at run time through its rpath)."""
    os.makedirs(BINDIR, exist_ok=True)
    include = os.path.join(os.path.dirname(HERE), "include")
    tmp = CLI + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-Werror",
           "-I", include, os.path.join(HERE, "cli", "mitrace_gpu_cli.cpp"), "-o", tmp,
           "-L", LIBDIR, "-lmitrace_gpu", "-Wl,-rpath,$ORIGIN/../lib", "-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
