"""Golden fixtures (tests/golden/, written by make_golden.py from the oracle)
and the checking-value / QA path of the library's host code.

The oracle must reproduce its own fixtures bit for bit (its reductions are
deterministic across thread counts), so a change of the oracle shows up
here; the GPU side compares the library with the same files in
test_gpu_golden.py."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle as O
import tealeaf_hip as T
from tealeaf_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
PROBLEMS = os.path.join(ROOT, "tealeaf.jl_amd", "decks", "tea.problems")


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def test_oracle_reproduces_c1_fixture():
    g = _load("c1_512_cg.json")
    _, (s, st) = O.run_deck(g["deck"], threads=4)
    assert np.array_equal(s, np.array(g["intended"]["summary"]))
    assert [x["iterations"] for x in st] == [x["iterations"] for x in g["intended"]["stats"]]
    lit = O.LIT_D1 | O.LIT_D2 | O.LIT_D3 | O.LIT_D4
    _, (sl, _) = O.run_deck(g["deck"], literal=lit, threads=4)
    assert np.array_equal(sl, np.array(g["literal"]["summary"]))


def test_literal_fixture_follows_diagonal_closed_form():
    # as written, smvp keeps only its diagonal term (kernels.jl:181), so each
    # step divides u by D = 1 + kx[k+1] + kx[k] + ky[j+1] + ky[j] cell by cell:
    # ie = sum vol*u must fall every step while vol and mass stay fixed
    g = np.array(_load("c1_512_cg.json")["literal"]["summary"])
    assert np.all(g[:, 0] == g[0, 0]) and np.all(g[:, 1] == g[0, 1])
    assert np.all(np.diff(g[:, 2]) < 0)


@pytest.mark.parametrize("solver", ["cg", "cheby", "ppcg", "jacobi"])
def test_oracle_reproduces_solver_fixtures(solver):
    g = _load("solvers_64.json")[solver]
    _, (s, st) = O.run_deck(g["deck"], threads=4)
    assert np.array_equal(s, np.array(g["summary"]))
    for a, b in zip(st, g["stats"]):
        assert (a["iterations"], a["cg_iters"]) == (b["iterations"], b["cg_iters"])


def test_intended_fixtures_conserve_ie_and_agree_across_solvers():
    gs = _load("solvers_64.json")
    ref = np.array(gs["cg"]["summary"])
    for solver in ("cheby", "ppcg", "jacobi"):
        np.testing.assert_allclose(np.array(gs[solver]["summary"]), ref, rtol=1e-9)
    c1 = np.array(_load("c1_512_cg.json")["intended"]["summary"])
    np.testing.assert_allclose(c1[:, 2], c1[0, 2], rtol=1e-11)  # 1^T A = 1^T


def test_checking_value_lookup_and_qa_rule(tmp_path):
    L = T.load()
    g = _load("c1_512_cg.json")
    temp = g["intended"]["summary"][-1][3]
    v, found = C.c_double(), C.c_int()
    assert L.tl_checking_value(PROBLEMS.encode(), 512, 512, 10, C.byref(v), C.byref(found)) == 0
    assert found.value == 1 and v.value == temp
    # no matching line: the reference's fallback value 1.0
    assert L.tl_checking_value(PROBLEMS.encode(), 512, 512, 11, C.byref(v), C.byref(found)) == 0
    assert found.value == 0 and v.value == 1.0
    assert L.tl_checking_value(str(tmp_path / "none").encode(), 1, 1, 1, C.byref(v), C.byref(found)) == _lib.TL_E_IO
    bad = tmp_path / "bad.problems"
    bad.write_text("512 512 x 1.0\n")
    assert L.tl_checking_value(str(bad).encode(), 512, 512, 10, C.byref(v), C.byref(found)) == _lib.TL_E_PARSE
    d = C.c_double()
    assert L.tl_qa_check(temp, temp, C.byref(d)) == 1 and d.value == 0.0
    assert L.tl_qa_check(temp * (1 + 0.9e-5), temp, C.byref(d)) == 1
    assert L.tl_qa_check(temp * (1 + 1.1e-5), temp, C.byref(d)) == 0  # |100 a/cv - 100| >= 0.001 fails
    s = T.Settings.from_text(g["deck"])
    assert T.checkingvalue(s, PROBLEMS) == temp


def test_oracle_reproduces_fp64_rectangle_fixture():
    """The FP64 answer the mixed-precision solver is held to: the oracle
    reproduces it bit for bit, each step's residual meets 1e-12 and the
    conserved quantities stay fixed (vol, mass exactly; ie = sum vol*u to
    the residual)."""
    g = _load("rect_256_fp64.json")
    d = O.parse_deck(g["deck"])
    d.stop_rule, d.rel_tol, d.maxiters = O.STOP_RELATIVE, g["rel_tol"], 100000
    o = O.Oracle(d, threads=8)
    o.initialise()
    s, st = o.run(3)
    assert np.array_equal(s, np.array(g["summary"]))
    assert all(x["rel_resid"] <= 1e-12 for x in st)
    assert np.all(s[:, 0] == s[0, 0]) and np.all(s[:, 1] == s[0, 1])
    np.testing.assert_allclose(s[:, 2], s[0, 2], rtol=1e-9)


def test_c2_fixture_invariants():
    """The 4096^2 fixture (too costly to recompute here): vol and mass fixed,
    ie conserved (1^T A = 1^T under reflection), the three solvers agree
    within their tolerances at every step they share, and the first step's CG
    pre-steps of Chebyshev and PPCG are the same solve."""
    g = _load("c2_4096.json")
    ref = None
    for solver in ("cg", "cheby", "ppcg"):
        if solver not in g:
            continue
        s = np.array(g[solver]["summary"])
        assert np.all(s[:, 0] == s[0, 0]) and np.all(s[:, 1] == s[0, 1])
        np.testing.assert_allclose(s[:, 2], s[0, 2], rtol=1e-9)
        assert np.all(np.diff(s[:, 3]) != 0)
        if ref is None:
            ref = s
        else:
            k = min(len(ref), len(s))
            np.testing.assert_allclose(s[:k], ref[:k], rtol=1e-8)
    assert "cheby" in g and "ppcg" in g
    # the first step's CG phase is the same computation for both
    assert g["cheby"]["stats"][0]["cg_iters"] == g["ppcg"]["stats"][0]["cg_iters"]
