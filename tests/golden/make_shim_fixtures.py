#!/usr/bin/env python3
"""Write tests/golden/shim_outputs.json: the printed outputs of the compiled
reference (oracle/_ref/<cfg>/iwref, built by oracle/build_ref.py from the
sources under /root/reference) for the cases the Fortran-shim GPU tests run.

Run in a container where /root/reference is mounted and the reference builds
exist.  The GPU box has neither: there the shim driver (oracle/_shim, the
product's shims with no reference object linked) is checked against the C
oracle for the state and against these recorded outputs for what the
reference prints.  Data only, each entry an input description plus the
reference's outputs:

* ``maxcalc_c1_50``: 50 steps of the c1 case (64x16, 4x17) with the
  reference maxcalc (various.f:665-987, adaptive dt with Tso kept): the
  per-step dt log (istep, t, dt, cdt, ivardt, idtstep), the 'At max CFLZ'
  values (dz, cflzmax), the temperature extremum positions and the fort.50
  lines (E12.6).
* ``meanwave_c1_20``: the same for 20 steps with the shear current
  (fcshear 0.5) and the nine /sol2dvar/ fields of tests/helpers.meanwave_fields.
* ``diag_small_6``: 6 steps of the small case (16x8, 3x9) with the estimation
  filter and postp every 3 steps, plotf and speclegencalc every 2: the
  /energetics/ rows (istep, t, sixteen values; postp.f:53-58 order) and the
  speclegen*.dat spectra (E12.6).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import oracle as O  # noqa: E402
from helpers import case, meanwave_fields  # noqa: E402


def _mc(m):
    return {k: m[k] for k in ("dtlog", "cflz", "tloc", "fort50")}


def main():
    out = {"provenance": "oracle/_ref/<cfg>/iwref: the reference Fortran (jmsegret/IWrepo mode1_bbl) built "
                         "with ROCm flang -fdefault-real-8 -fdefault-double-8 -fno-automatic by "
                         "oracle/build_ref.py, driven by oracle/ref_harness.f"}
    P, flds = case("c1")
    _, _, _, m = O.run_ref("c1", P, flds, nsteps=50, maxcalc=True)
    out["maxcalc_c1_50"] = _mc(m)
    P, flds = case("c1")
    P.fcshear = 0.5
    mw = meanwave_fields(P, O.Oracle(P).table(7, P.nzloc * P.nsubd))
    _, _, _, m = O.run_ref("c1", P, flds, nsteps=20, maxcalc=True, meanwave=mw)
    out["meanwave_c1_20"] = _mc(m)
    P, flds = case("small")
    P.ifilter, P.nstepfilter, P.npostp, P.n3doutput, P.ncalclegspec = 1, 3, 3, 2, 2
    _, _, _, er, lr = O.run_ref("small", P, flds, nsteps=6, energy=True, speclegen=True)
    out["diag_small_6"] = {"energy": er.tolist(),
                           "speclegen": {"%s|%s" % k: v.tolist() for k, v in sorted(lr.items())}}
    with open(os.path.join(HERE, "shim_outputs.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "shim_outputs.json"))


if __name__ == "__main__":
    main()
