"""Build libiwgpu.so (HIP, gfx950) in-tree.

``python -m iwrepo_amd.build`` compiles every HIP source for gfx950 with hipcc
(cross-compiles without a GPU).  The Fortran drop-in shims
(iwrepo_amd/fortran/iwshim.F) are compiled inside the caller's own Fortran
build, because they include the caller's dim.h and COMMON headers.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libiwgpu.so")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wno-unused-value",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-result", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode())
        raise RuntimeError("build failed: " + " ".join(cmd[:3]))
    return r.stdout.decode()


def build_hip(verbose=False):
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(CSRC, os.path.basename(s) + ".o")
        objs.append(o)
        procs.append(subprocess.Popen([HIPCC] + HIPFLAGS + ["-c", s, "-o", o],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode())
            raise RuntimeError("hipcc failed")
        if verbose and out:
            sys.stderr.write(out.decode())
    tmp = LIB + ".tmp"
    _run([HIPCC] + HIPFLAGS + ["-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_all(verbose=False):
    return build_hip(verbose)


if __name__ == "__main__":
    print(build_all(verbose=True))
