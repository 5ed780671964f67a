#!/usr/bin/env python3
"""Write c2_4096.json: the 4096^2 reference deck solved by the CPU oracle.

The three-state reference deck (rho=100, e=1e-4 background; two rho=0.1
rectangles at x in [0,1] and [1,6], y in [1,2]) on [0,10]^2 at 4096^2,
dt 0.004, eps 1e-15, each solver with its own stop rule: CG, Chebyshev with 8
CG pre-steps and PPCG with 10 inner Chebyshev steps.  For every timestep the
file holds the field summary (vol, mass, ie, temp) and the solver statistics
(iterations, CG iterations before the switch, Chebyshev estimate, eigenvalue
bounds, final error).  The GPU tests compare the library with these numbers
without running the oracle.

A 4096^2 step costs the oracle minutes on 8 cores (CG about 2 min, Chebyshev
about 1.5 min, PPCG longer), so this script is separate from make_golden.py
and writes the file after every step; a run that is stopped keeps the steps
it finished.

    python tests/golden/make_golden_4096.py [--cg N] [--cheby N] [--ppcg N] [--threads T]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tealeaf.jl_amd")]

import oracle as O  # noqa: E402
from tealeaf_hip import decks  # noqa: E402

OUT = os.path.join(HERE, "c2_4096.json")
KEEP = ("iterations", "cg_iters", "cheby_iters", "est_iters", "error", "eigmin", "eigmax", "rel_resid")


def deck(solver: str, steps: int) -> str:
    return decks.reference_deck(4096, steps, solver=solver, presteps=8, ppcg_inner_steps=10)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--cg", type=int, default=20)
    ap.add_argument("--cheby", type=int, default=20)
    ap.add_argument("--ppcg", type=int, default=2)
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    out = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            out = json.load(f)
    for solver, n in (("ppcg", a.ppcg), ("cheby", a.cheby), ("cg", a.cg)):
        if n <= 0:
            continue
        text = deck(solver, 20)
        o = O.Oracle(O.parse_deck(text), threads=a.threads)
        o.initialise()
        rec = {"deck": text, "summary": [], "stats": []}
        for t in range(n):
            t0 = time.time()
            st = o.step()
            rec["summary"].append(o.summary().tolist())
            rec["stats"].append({k: st[k] for k in KEEP})
            out[solver] = rec
            with open(OUT + ".tmp", "w") as f:
                json.dump(out, f, indent=1)
                f.write("\n")
            os.replace(OUT + ".tmp", OUT)
            print(f"{solver} step {t + 1}/{n}: {st['iterations']} iterations, {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
