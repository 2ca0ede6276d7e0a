#!/usr/bin/env python3
"""Recipe that compiles the reference Fortran into ``oracle/_ref/<name>/iwref``.

TEST INFRASTRUCTURE ONLY.  The reference (jmsegret/IWrepo, ``mode1_bbl``) is
compiled from its own sources where they lie under ``/root/reference``: the
build folder holds symlinks to them, never copies, plus a generated ``dim.h``
(the reference sizes every array at compile time in ``dim.h:12-28`` and ships
several variants of it, e.g. ``dim_65.h``, ``dim_multi.h``).  Flags follow
``makefile:24-31``: ``-autodouble`` becomes ``-fdefault-real-8
-fdefault-double-8`` and ifort's static local arrays become ``-fno-automatic``.

``ref_harness.f`` replaces ``main.f`` (which flang rejects at ``main.f:571``).
``various.f`` (``maxcalc``, ``various.f:665-987``) is compiled with the
preprocessor renaming the three REAL loop counters of ``testdump``
(``various.f:493-523``, which flang rejects as subscripts) to INTEGER names;
that routine is not called.  flang also rejects constant subscripts outside the
arrays (``various.f:194-204`` for ny = 1, ``various.f:646, 745`` for nxpp < 32 or
nz < 58), so builds where it does not compile leave it out and the harness
supplies its ``zero``/``copy`` helpers instead.  ``postp.f`` (the TKE budget
of ``postp``, ``postp.f:9-76, 415-732``) is compiled the same way, and the
harness calls ``postp`` where ``main.f:635, 935`` do when it is in the build.

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

# testdump's REAL loop counters used as subscripts (various.f:493-523)
VARIOUS_FLAGS = ["-cpp", "-Dcounter=icounter", "-Dcounter2=icounter2", "-Dcounter3=icounter3"]

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
    vo = os.path.join(bdir, "various.o")
    r = subprocess.run([FLANG] + FFLAGS + VARIOUS_FLAGS + ["-c", os.path.join(src, "various.f"), "-o", vo],
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with_maxcalc = r.returncode == 0
    marker = os.path.join(bdir, "has_maxcalc")
    if with_maxcalc:
        open(marker, "w").close()
    else:
        if os.path.exists(marker):
            os.remove(marker)
        print(name, ": various.f does not compile at these sizes; maxcalc left out")
    objs = [vo] if with_maxcalc else []
    po = os.path.join(bdir, "postp.o")
    r = subprocess.run([FLANG] + FFLAGS + ["-c", os.path.join(src, "postp.f"), "-o", po],
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    with_postp = r.returncode == 0
    pmark = os.path.join(bdir, "has_postp")
    if with_postp:
        open(pmark, "w").close()
        objs.append(po)
    else:
        if os.path.exists(pmark):
            os.remove(pmark)
        print(name, ": postp.f does not compile at these sizes; postp left out")
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
    hdef = (["-DIW_MAXCALC"] if with_maxcalc else []) + (["-DIW_POSTP"] if with_postp else [])
    subprocess.run([FLANG] + FFLAGS + ["-cpp"] + hdef + ["-I", src, "-c", os.path.join(HERE, "ref_harness.f"),
                                                         "-o", ho], check=True)
    exe = os.path.join(bdir, "iwref")
    link([FLANG] + FFLAGS + [ho] + objs + ["-o", exe])
    print("built", exe)
    if shim:
        build_shim(bdir, src, ho, objs, noslip)
        build_shimdrv(name, src, noslip)
    return exe


# Symbols that may stay unresolved: referenced only from routines the harness
# never calls.  time: the ifort library routine behind timecall
# (various.f:990-1010); curvefit: wakelengthcalc (postp.f:736-981, its call
# at postp.f:39-41 commented out; curvefit.f is rejected by flang);
# umeanprof: umeanreset (postp.f:364-412, relaxation only).
UNRESOLVED_OK = ("time_", "curvefit_", "umeanprof_")


def link(cmd):
    """Link, leaving only the symbols of UNRESOLVED_OK unresolved (reported
    as warnings); any other unresolved symbol fails the build."""
    cmd = cmd + ["-Wl,--warn-unresolved-symbols"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    out = r.stdout.decode()
    bad = [l for l in out.splitlines()
           if "undefined symbol" in l and not any(("symbol: " + u) in l for u in UNRESOLVED_OK)]
    if r.returncode != 0 or bad:
        sys.stderr.write(out)
        raise SystemExit("link failed")


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
    link([FLANG] + FFLAGS + [ho, so] + objs +
         ["-o", exe, "-Wl,--allow-multiple-definition", "-L" + PKG, "-liwgpu",
          "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath,/opt/rocm/lib"])
    print("built", exe)
    return exe


SHIMOUT = os.path.join(HERE, "_shim")


def build_shimdrv(name, src, noslip):
    """The harness linked with the product's Fortran shims and
    ``shimdrv_support.f`` only: no object of the reference is in the link
    (its headers give the COMMON layout the shims are compiled against, as
    in any caller's build).  Output ``oracle/_shim/<name>/shimdrv``, which
    the GPU tests run against the oracle."""
    lib = os.path.join(PKG, "libiwgpu.so")
    if not os.path.exists(lib):
        print("libiwgpu.so not built; skipping shim driver")
        return None
    odir = os.path.join(SHIMOUT, name)
    os.makedirs(odir, exist_ok=True)
    cmd = [FLANG] + FFLAGS + ["-cpp", "-DIW_SHIMDRV", "-DIW_MAXCALC", "-DIW_POSTP", "-I", src, "-J", odir]
    if noslip:
        cmd.append("-DIW_NOSLIP")
    objs = []
    for f in (os.path.join(PKG, "fortran", "iwshim.F"), os.path.join(HERE, "shimdrv_support.f"),
              os.path.join(HERE, "ref_harness.f")):
        o = os.path.join(odir, os.path.basename(f).split(".")[0] + ".o")
        subprocess.run(cmd + ["-c", f, "-o", o], check=True)
        objs.append(o)
    exe = os.path.join(odir, "shimdrv")
    rel = os.path.relpath(PKG, odir)
    r = subprocess.run([FLANG] + FFLAGS + objs + ["-o", exe, "-L" + PKG, "-liwgpu",
                                                  "-Wl,-rpath,$ORIGIN/" + rel, "-Wl,-rpath,/opt/rocm/lib"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode())
        raise SystemExit("shim driver link failed")
    for o in objs:
        os.remove(o)
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
