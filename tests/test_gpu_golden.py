"""GPU: the library (through its C ABI) against the committed golden
fixtures -- no oracle at run time -- and the command-line driver's QA line
against the committed tea.problems."""
import json
import os
import subprocess

import numpy as np
import pytest

import tealeaf_hip as T
from tealeaf_hip import decks

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
DECKS = os.path.join(ROOT, "tealeaf.jl_amd", "decks")
RTOL = 1e-10


def _load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_c1_512_cg_matches_golden_and_passes_qa():
    print("case: golden 512^2 reference deck, 10 CG steps", flush=True)
    g = _load("c1_512_cg.json")
    s = T.Settings.from_text(g["deck"])
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, 10)
    want = np.array(g["intended"]["summary"])
    for t in range(10):
        assert _rel(summ[t], want[t]) < RTOL, (t, summ[t], want[t])
        wi = g["intended"]["stats"][t]["iterations"]
        assert abs(st[t]["iterations"] - wi) <= max(3, wi // 50), (t, st[t]["iterations"], wi)
    qa = T.qa_check(c, s, os.path.join(DECKS, "tea.problems"))
    assert qa["passed"] and qa["qa_diff"] < 1e-6, qa


@pytest.mark.parametrize("solver", ["cg", "cheby", "ppcg", "jacobi"])
def test_solvers_64_match_golden(solver):
    print(f"case: golden 64^2 reference deck, 3 steps, {solver}", flush=True)
    g = _load("solvers_64.json")[solver]
    s = T.Settings.from_text(g["deck"])
    c = T.initialiseapp(s)
    summ, st = T.diffuse(c, s, 3)
    for t in range(3):
        assert _rel(summ[t], g["summary"][t]) < RTOL, (solver, t, summ[t], g["summary"][t])
        if solver in ("cheby", "ppcg"):
            assert st[t]["cg_iters"] == g["stats"][t]["cg_iters"]


def test_rect_32_u_field_matches_golden():
    print("case: golden 32^2 random rectangles, u after one step", flush=True)
    g = _load("rect_32_u.json")
    s = T.Settings.from_text(g["deck"])
    c = T.initialiseapp(s)
    T.diffuse(c, s, 1)
    u = c.interior("u")
    want = np.array(g["u_interior"])
    assert u.shape == want.shape
    assert np.max(np.abs(u - want)) <= RTOL * np.max(np.abs(want))


def test_cli_runs_deck_and_prints_passed():
    print("case: CLI tealeaf on the 512^2 deck with tea.problems", flush=True)
    exe = os.path.join(ROOT, "tealeaf.jl_amd", "lib", "tealeaf")
    out = subprocess.run([exe, "-i", os.path.join(DECKS, "tea_512.in"), "-p", os.path.join(DECKS, "tea.problems")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [json.loads(x) for x in out.stdout.splitlines() if x.startswith("{")]
    assert len(lines) == 11 and lines[-1]["check"] == "PASSED", lines[-1]
    g = _load("c1_512_cg.json")
    assert _rel(lines[9]["temp"], g["intended"]["summary"][9][3]) < RTOL


def test_debugrecord_layout(tmp_path):
    print("case: debugrecord dump of a 6x4 chunk", flush=True)
    text = decks.reference_deck(6, 1).replace("y_cells=6", "y_cells=4")
    s = T.Settings.from_text(text)
    c = T.initialiseapp(s)
    path = tmp_path / "debug.txt"
    T.debugrecord(c, str(path))
    lines = path.read_text().split("\n")
    assert lines[0] == "density0" and len(lines[1].split(" ")) == c.X
    i = lines.index("energy")
    rows = lines[i + 1:i + 1 + c.Y]
    got = np.array([[float(v) for v in r.split(" ")] for r in rows])
    assert np.array_equal(got, c.field("energy"))
    assert [ln for ln in lines if ln in T.FIELDS] == list(T.FIELDS)


def test_diffuse_writes_debugfile_each_step(tmp_path):
    """diffuse! with a debug file (TeaLeaf.jl:63-68): an old file is replaced
    and the chunk is dumped at the start of every timestep."""
    print("case: diffuse with debugfile, 8x8, 2 steps", flush=True)
    s = T.Settings.from_text(decks.reference_deck(8, 2, max_iters=100))
    path = tmp_path / "dbg.txt"
    path.write_text("stale\n")
    s.set(debugfile=str(path))
    c = T.initialiseapp(s)
    T.diffuse(c, s, 2)
    lines = path.read_text(encoding="utf-8").split("\n")
    assert "stale" not in lines
    assert lines.count("density0") == 2 and lines.count("chebyβ") == 2
