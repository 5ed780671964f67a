#!/usr/bin/env python3
"""Write the golden fixtures of this directory from the CPU oracle.

The reference (TeaLeaf.jl) ships no test vectors, known-answer tests or
fixtures, and it cannot run here (no Julia), so these files are written by
the oracle, whose functions are pinned separately by closed forms and
textbook routines (tests/test_oracle.py).  They freeze the oracle's answers
as data so that the GPU tests can compare against them without running the
oracle, and so that any change of the oracle shows up as a diff.

    python tests/golden/make_golden.py            # rewrite every fixture

Fixtures (JSON; floats written as shortest round-trip decimals):
  c1_512_cg.json       the 512^2 reference deck, 10 steps of CG: per-step
                       field summary (vol, mass, ie, temp), iterations,
                       final error; intended semantics and, separately,
                       the reference as written (literal mode, defects D1-D4)
  solvers_64.json      the same deck at 64^2, 3 steps, for CG, Chebyshev,
                       PPCG and Jacobi: summaries, iteration counts, eigen
                       bounds
  rect_32_u.json       a 32^2 random-rectangle deck (seed 9): the interior
                       of u after one CG step
  rect_256_fp64.json   a 256^2 random-rectangle deck (seed 3), 3 steps of
                       FP64 CG to a relative residual of 1e-12: the
                       reference answer of the mixed-precision (FP32-refined)
                       solver, which must agree within 1e-9
  tea.problems         written to tealeaf.jl_amd/decks/: checking values
                       (final temp) for the reference decks above, in the
                       reference's "x y endstep value" format
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tealeaf.jl_amd")]


import oracle as O  # noqa: E402
from tealeaf_hip import decks  # noqa: E402

LITERAL_ALL = O.LIT_D1 | O.LIT_D2 | O.LIT_D3 | O.LIT_D4


def _stats(st):
    keep = ("iterations", "cg_iters", "cheby_iters", "est_iters", "error", "eigmin", "eigmax", "rel_resid")
    return [{k: s[k] for k in keep} for s in st]


def c1():
    text = decks.reference_deck(512, 10)
    _, (s, st) = O.run_deck(text, threads=8)
    _, (sl, stl) = O.run_deck(text, literal=LITERAL_ALL, threads=8)
    return {"deck": text, "steps": 10,
            "intended": {"summary": s.tolist(), "stats": _stats(st)},
            "literal": {"summary": sl.tolist(), "stats": _stats(stl)}}


def solvers():
    out = {}
    for solver in ("cg", "cheby", "ppcg", "jacobi"):
        text = decks.reference_deck(64, 3, solver=solver, presteps=30, max_iters=20000)
        _, (s, st) = O.run_deck(text, threads=8)
        out[solver] = {"deck": text, "summary": s.tolist(), "stats": _stats(st)}
    return out


def rect_u():
    text = decks.random_rect_deck(9, n=32, steps=1)
    d = O.parse_deck(text)
    o = O.Oracle(d, threads=8)
    o.initialise()
    s, st = o.run(1)
    return {"deck": text, "summary": s.tolist(), "stats": _stats(st), "u_interior": o.interior("u").tolist()}


def rect_fp64():
    text = decks.random_rect_deck(3, n=256, steps=3, precision="fp32_ir")
    d = O.parse_deck(text)
    d.stop_rule, d.rel_tol, d.maxiters = O.STOP_RELATIVE, 1e-12, 100000
    o = O.Oracle(d, threads=8)
    o.initialise()
    s, st = o.run(3)
    return {"deck": text, "rel_tol": 1e-12, "summary": s.tolist(), "stats": _stats(st)}


def main():
    c = c1()
    sv = solvers()
    r = rect_u()
    rf = rect_fp64()
    for name, obj in (("c1_512_cg.json", c), ("solvers_64.json", sv), ("rect_32_u.json", r),
                      ("rect_256_fp64.json", rf)):
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(obj, f, indent=1)
            f.write("\n")
    lines = [f"512 512 10 {c['intended']['summary'][-1][3]!r}",
             f"64 64 3 {sv['cg']['summary'][-1][3]!r}"]
    with open(os.path.join(ROOT, "tealeaf.jl_amd", "decks", "tea.problems"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", ", ".join(["c1_512_cg.json", "solvers_64.json", "rect_32_u.json", "rect_256_fp64.json",
                              "tea.problems"]))


if __name__ == "__main__":
    main()
