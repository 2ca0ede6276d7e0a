#!/usr/bin/env python3
"""Recipe that compiles the reference Fortran into ``oracle/_ref/<name>/iwref``.

TEST INFRASTRUCTURE ONLY.  The reference (jmsegret/IWrepo, ``mode1_bbl``) is
compiled from its own sources where they lie under ``/root/reference``: the
build folder holds symlinks to them, never copies, plus a generated ``dim.h``
(the reference sizes every array at compile time in ``dim.h:12-28`` and ships
several variants of it, e.g. ``dim_65.h``, ``dim_multi.h``).  Flags follow
``makefile:24-31``: ``-autodouble`` becomes ``-fdefault-real-8
-fdefault-double-8`` and ifort's static local arrays become ``-fno-automatic``.

``various.f`` is not compiled (flang rejects its REAL subscript at
``various.f:503``); ``ref_harness.f`` supplies its ``zero``/``copy`` helpers and
replaces ``main.f`` (which flang rejects at ``main.f:571``).

Outputs go only under ``oracle/_ref/`` (git-ignored).  Nothing here runs when
``/root/reference`` is absent.
"""
import argparse
import os
import subprocess
import sys

REF = "/root/reference/mode1_bbl"
HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "iwrepo_amd")
OUT = os.path.join(HERE, "_ref")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
FFLAGS = ["-fdefault-real-8", "-fdefault-double-8", "-fno-automatic",
          "-O2", "-ffixed-line-length=132", "-mcmodel=medium"]
# hot-path sources of the reference (SURVEY-listed files), minus main.f/various.f
SOURCES = ["solver.f", "penalty.f", "band_lu.f", "bandludcmp.f", "ludcmp.f",
           "derv.f", "horfft.f", "fft.f", "filtering.f", "setup.f",
           "forcing.f", "multistep.f", "erfc.f", "sgscalc.f", "output.f", "restart.f"]

DIM_TEMPLATE = """\
      parameter   ( nx = {nx}, ny={ny}, nz= {nz}, nzloc = {nzloc}, nsubd = {nsubd})
      parameter   (nz2={nz2}, nzaux= {nzaux})
      parameter   ( nsdtype = 1)
      parameter   ( ndims = {ndims})
      parameter   ( nxpp = nx + 2,     nxh=nx/2, nxhp = nxh+1,
     &              nxf = 3*nx/2 + 1,  nyf = 2*ny,  nyh= ny/2,
     &              nxppy = nxpp*ny,   ntot = nxppy*nz,
     &              ntot2 = nxppy*nz2, nzm = nzloc - 1)
      parameter(nxhpy=nxhp*ny)
      parameter(nzfl=96,nzh = nzfl/2)
"""


def build(name, nx, ny, nzloc, nsubd, noslip=False, ref_dim=False, jobs=8, shim=False):
    if not os.path.isdir(REF):
        print("reference not mounted; skipping", name)
        return None
    bdir = os.path.join(OUT, name)
    src = os.path.join(bdir, "src")
    os.makedirs(src, exist_ok=True)
    for f in os.listdir(REF):
        if not (f.endswith(".f") or f.endswith(".h")):
            continue
        if f == "dim.h" and not ref_dim:
            continue
        dst = os.path.join(src, f)
        if os.path.lexists(dst):
            os.remove(dst)
        os.symlink(os.path.join(REF, f), dst)
    if not ref_dim:
        nz = nzloc * nsubd
        txt = DIM_TEMPLATE.format(nx=nx, ny=ny, nz=nz, nzloc=nzloc, nsubd=nsubd,
                                  nz2=nsubd * (nzloc - 1) + 1, nzaux=max(nz, 68),
                                  ndims=3 if ny > 1 else 2)
        dpath = os.path.join(src, "dim.h")
        if os.path.lexists(dpath):
            os.remove(dpath)
        with open(dpath, "w") as fh:
            fh.write(txt)
    srcs = SOURCES + ["calcuvwt_noslip.f" if noslip else "calcuvwt.f"]
    objs = []
    procs = []
    for s in srcs:
        o = os.path.join(bdir, s.replace(".f", ".o"))
        objs.append(o)
        procs.append(subprocess.Popen([FLANG] + FFLAGS + ["-c", os.path.join(src, s), "-o", o],
                                      stdout=subprocess.DEVNULL, stderr=subprocess.PIPE))
        if len(procs) >= jobs:
            _wait(procs)
    _wait(procs)
    ho = os.path.join(bdir, "ref_harness.o")
    subprocess.run([FLANG] + FFLAGS + ["-I", src, "-c", os.path.join(HERE, "ref_harness.f"),
                                       "-o", ho], check=True)
    exe = os.path.join(bdir, "iwref")
    subprocess.run([FLANG] + FFLAGS + [ho] + objs + ["-o", exe], check=True)
    print("built", exe)
    if shim:
        build_shim(bdir, src, ho, objs, noslip)
    return exe


def build_shim(bdir, src, ho, objs, noslip):
    """The same harness with the product's Fortran shims (iwshim.F) linked in
    front of the reference objects: every call of horfft, dpdz, convec,
    calcuvwt, interface_avg3d, filter_3dlegenfourier and the multistep
    functions then reaches libiwgpu.so, while setup and the step loop stay
    the reference's own.  The shim object comes first on the link line and
    ``--allow-multiple-definition`` keeps the first definition."""
    lib = os.path.join(PKG, "libiwgpu.so")
    if not os.path.exists(lib):
        print("libiwgpu.so not built; skipping shim harness")
        return None
    so = os.path.join(bdir, "iwshim.o")
    cmd = [FLANG] + FFLAGS + ["-cpp", "-I", src, "-J", bdir]
    if noslip:
        cmd.append("-DIW_NOSLIP")
    subprocess.run(cmd + ["-c", os.path.join(PKG, "fortran", "iwshim.F"), "-o", so], check=True)
    exe = os.path.join(bdir, "iwshim")
    rel = os.path.relpath(PKG, bdir)
    subprocess.run([FLANG] + FFLAGS + [ho, so] + objs +
                   ["-o", exe, "-Wl,--allow-multiple-definition", "-L" + PKG, "-liwgpu",
                    "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath,/opt/rocm/lib"], check=True)
    print("built", exe)
    return exe


def _wait(procs):
    while procs:
        p = procs.pop(0)
        _, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(err.decode())
            raise SystemExit("flang failed")


# Named reference builds used by the tests (sizes of the BASELINE-derived
# parity cases plus the reference's own shipped configuration).
BUILDS = {
    "c1": dict(nx=64, ny=16, nzloc=17, nsubd=4, shim=True),
    "c1ns": dict(nx=64, ny=16, nzloc=17, nsubd=4, noslip=True, shim=True),
    "small": dict(nx=16, ny=8, nzloc=9, nsubd=3, shim=True),
    "smallns": dict(nx=16, ny=8, nzloc=9, nsubd=3, noslip=True),
    "g3": dict(nx=256, ny=1, nzloc=25, nsubd=12, ref_dim=True),
}

if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("names", nargs="*", default=list(BUILDS))
    a = ap.parse_args()
    for n in a.names:
        build(n, **BUILDS[n])
