"""Build the native pieces of bamrex_amd in-tree.

* libbamrex_hip.so  -- the HIP/CDNA4 hot path behind include/bamrex_hip.h
                       (hipcc, --offload-arch=gfx950)
* bamrex_run        -- the C++ host driver (AmrLevelAdv-shaped time loop that
                       reads the reference's input files), linked to the library

Device code is compiled with -ffp-contract=off so that every FP64 operation
is rounded exactly as written (bit-for-bit agreement with the CPU oracle) and
with the packed-FP32 feature disabled.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")

HIP_SOURCES = ["bx_explicit.hip", "bx_sweep2d.hip", "bx_sweep3d.hip", "bx_api.hip",
               "bx_helmholtz.hip", "bx_cnn.hip", "bx_imex.hip", "bx_amr.hip", "bx_plotfile.hip",
               "bx_checkpoint.hip"]
HOST_SOURCES = ["ParmParse.cpp", "AmrLevelAdv.cpp", "ImexLevel.cpp", "Prob.cpp", "main.cpp"]
HOST_BIN = os.path.join(HERE, "bamrex_run")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BAMREX_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

LIB = os.path.join(HERE, "libbamrex_hip.so")


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def _headers():
    return [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + [
        os.path.join(ROOT, "include", "bamrex_hip.h")]


def _included_sources(src):
    """Sources a translation unit pulls in with #include "x.hip" (the
    per-dimension sweep units include the explicit-scheme source)."""
    out = []
    with open(src) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#include") and '.hip"' in line:
                out.append(os.path.join(os.path.dirname(src), line.split('"')[1]))
    return out


def build(force=True, jobs=6, verbose=False, defines=(), lib_path=None, obj_dir=None,
          host=True):
    """Compile every HIP source for gfx950 and link libbamrex_hip.so.

    force=True recompiles everything (as __graft_entry__.build() asks); force=False only
    rebuilds stale objects."""
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    obj = obj_dir or OBJ
    lib_out = lib_path or LIB
    os.makedirs(obj, exist_ok=True)
    hdrs = _headers()
    flags = HIP_FLAGS + ["-D" + d for d in defines]
    objs = []
    jobs_list = []
    for src in HIP_SOURCES:
        s = os.path.join(CSRC, src)
        o = os.path.join(obj, src.replace(".hip", ".o"))
        objs.append(o)
        if force or _stale(o, [s] + _included_sources(s) + hdrs):
            jobs_list.append([HIPCC] + flags + ["-c", s, "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for out in ex.map(_run, jobs_list):
            if verbose and out.strip():
                print(out)
    if force or _stale(lib_out, objs):
        tmp = lib_out + ".tmp"
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
        os.replace(tmp, lib_out)
    if not host:
        return lib_out
    host = [os.path.join(CSRC, "host", f) for f in HOST_SOURCES]
    if force or _stale(HOST_BIN, host + hdrs + [LIB]):
        tmp = HOST_BIN + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", tmp] + host +
             ["-L" + HERE, "-lbamrex_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, HOST_BIN)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print("built", LIB)
