"""Build libiwgpu.so (HIP, gfx950) in-tree, plus the Fortran drop-in shims.

``python -m iwrepo_amd.build`` compiles every HIP source for gfx950 with hipcc
(cross-compiles without a GPU).  The Fortran shims (iwrepo_amd/fortran) are
compiled with ROCm flang when it is present.
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
FLANG = "/opt/rocm/lib/llvm/bin/flang"

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


def build_fortran():
    """Fortran drop-in shims (same names/arguments as the reference routines)."""
    fdir = os.path.join(PKG, "fortran")
    if not os.path.exists(FLANG) or not os.path.isdir(fdir):
        return None
    out = os.path.join(PKG, "libiwfort.so")
    srcs = [os.path.join(fdir, f) for f in sorted(os.listdir(fdir)) if f.endswith(".f90")]
    if not srcs:
        return None
    _run([FLANG, "-O2", "-fPIC", "-shared", "-J", fdir, "-o", out] + srcs +
         ["-L" + PKG, "-liwgpu", "-Wl,-rpath,$ORIGIN"])
    return out


def build_all(verbose=False):
    lib = build_hip(verbose)
    try:
        build_fortran()
    except RuntimeError as e:  # the shims are optional for the product path
        sys.stderr.write("fortran shims not built: %s\n" % e)
    return lib


if __name__ == "__main__":
    print(build_all(verbose=True))
