"""Error of the mixed-precision evolve loop against FP64, per quantity.

Workload: the 1000 m x 1000 m dam break over a sloping bed with a Gaussian
bump (rectangular_cross, ``--cells`` x ``--cells`` cells, 4 triangles each;
2 m / 0.5 m stage either side of x = 500 m, Manning n 0.03, reflective
walls).  The FP64 loop runs ``--steps`` Euler steps and reaches time T; the
mixed loop (FP32 flux and reconstruction arithmetic and FP32 edge state;
FP64 accumulation, timestep and depth tests) runs to the same T.  Reported per
quantity (stage, xmomentum, ymomentum centroid values): max |error| and the
relative L2 error ||mixed - fp64|| / ||fp64||, the wet/dry mask differences,
the step counts and both loops' triangle-steps/s.

The FP64 loop is the reference here: it reproduces the FP64 C oracle to
within 1e-12 per step (the GPU parity tests check this, and check the mixed
loop against the oracle directly on a small mesh).

    python -m anuga_cuda_amd.precision_error --cells 1000 --steps 500 [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import scenarios

QUANTITIES = ("stage", "xmomentum", "ymomentum")


def _errors(got, ref, bed, eps):
    out = {}
    for q in QUANTITIES:
        a, b = got[q], ref[q]
        nrm = float(np.linalg.norm(b))
        out[q] = {"max_abs": float(np.max(np.abs(a - b))),
                  "rel_l2": float(np.linalg.norm(a - b) / nrm) if nrm > 0 else 0.0}
    wa = (got["stage"] - bed) > eps
    wb = (ref["stage"] - bed) > eps
    out["wet_dry_mismatch"] = int(np.count_nonzero(wa != wb))
    out["wet_cells"] = int(np.count_nonzero(wb))
    return out


def _run_gpu(cells, precision, steps=None, finaltime=None):
    d = scenarios.dam_break_channel(cells, cells, precision=precision)
    h = d._dev()
    if finaltime is None:
        h.call("sw_evolve_prepare", 1e9, 0.0, 0)
    else:
        h.call("sw_evolve_prepare", 1e9, finaltime, 1)
    h.call("sw_synchronize")
    t0 = time.perf_counter()
    done, ev = h.evolve_steps(steps if steps is not None else 1 << 40)
    h.call("sw_synchronize")
    el = time.perf_counter() - t0
    s = h.get_state()
    vals = {q: d.quantities[q].centroid_values.copy() for q in QUANTITIES}
    bed = d.quantities["elevation"].centroid_values.copy()
    res = {"steps": int(done), "time": s.time, "seconds": el,
           "triangle_steps_per_s": d.number_of_elements * done / el}
    eps = d.minimum_allowed_height
    n = d.number_of_elements
    d.close()
    return res, vals, bed, eps, n


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cells", type=int, default=1000)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--json")
    a = ap.parse_args(argv)

    r64, v64, bed, eps, n = _run_gpu(a.cells, "fp64", steps=a.steps)
    T = r64["time"]
    print("fp64 : %d triangles, %d steps to t = %.6f s, %.3e triangle-steps/s" % (
        n, r64["steps"], T, r64["triangle_steps_per_s"]), flush=True)
    rmx, vmx, _, _, _ = _run_gpu(a.cells, "mixed", finaltime=T)
    print("mixed: %d steps to t = %.6f s, %.3e triangle-steps/s" % (
        rmx["steps"], rmx["time"], rmx["triangle_steps_per_s"]), flush=True)
    report = {"workload": "dam break, rectangular_cross %dx%d (%d triangles), to t = %.6f s" % (
        a.cells, a.cells, n, T), "fp64": r64, "mixed": rmx,
        "mixed_vs_fp64": _errors(vmx, v64, bed, eps)}
    for key in ("mixed_vs_fp64",):
        if key in report:
            e = report[key]
            print("%-15s " % key + "  ".join("%s max|d| %.3e relL2 %.3e" % (q, e[q]["max_abs"], e[q]["rel_l2"])
                                             for q in QUANTITIES)
                  + "  wet/dry mismatches %d of %d wet" % (e["wet_dry_mismatch"], e["wet_cells"]), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(report, fh, indent=1)
    return report


if __name__ == "__main__":
    sys.exit(0 if main() else 1)
