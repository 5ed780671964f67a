"""GPU: the 4096^2 reference deck (CG, Chebyshev with 8 CG pre-steps, PPCG
with 10 inner steps; dt 0.004, eps 1e-15, each solver's own stop rule)
against the oracle's per-step answers committed in tests/golden/c2_4096.json:
all four field-summary quantities within 1e-10 at every step the fixture
holds, the CG pre-step count of Chebyshev and PPCG exactly, and the
iteration counts as in the other parity tests."""
import json
import os

import numpy as np
import pytest

import tealeaf_hip as T

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_4096.json")
RTOL = 1e-10


def _fixture():
    with open(GOLD) as f:
        return json.load(f)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("solver", ["cg", "cheby", "ppcg"])
def test_reference_deck_4096_every_step_matches_oracle(solver):
    g = _fixture()[solver]
    n = len(g["summary"])
    print(f"case: {solver} 4096^2 reference deck, {n} steps vs the oracle fixture", flush=True)
    s = T.Settings.from_text(g["deck"])
    c = T.initialiseapp(s)
    for t in range(n):
        st = T.step(c, s)
        got = T.fieldsummary(c)
        want = np.array(g["summary"][t])
        ost = g["stats"][t]
        rel = np.max(np.abs(got - want) / np.abs(want))
        print(f"  step {t + 1}: {st['iterations']} iterations (oracle {ost['iterations']}), "
              f"cg {st['cg_iters']} (oracle {ost['cg_iters']}), max rel diff {rel:.2e}", flush=True)
        assert rel < RTOL, (t, got, want)
        if solver != "cg":
            assert st["cg_iters"] == ost["cg_iters"], (t, st["cg_iters"], ost["cg_iters"])
        assert abs(st["iterations"] - ost["iterations"]) <= max(3, ost["iterations"] // 50), (t, st, ost)
    c.close()
