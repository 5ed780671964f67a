"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, its deck parser agrees with the oracle's, and its host-side
scalar routines (eigenvalue bounds, Chebyshev coefficients, iteration
estimate) agree with the oracle and with textbook results."""
import ctypes as C
import os
import re

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O
import tealeaf_hip as T
from tealeaf_hip import _lib, decks

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, "include", "tealeaf_hip.h")).read()
    declared = set(re.findall(r"\b(tl_[a-z0-9_]+)\s*\(", hdr)) - {"tl_step_cb"}
    L = T.load()
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert declared <= set(_lib.EXPORTS) | {"tl_step_cb"}, declared - set(_lib.EXPORTS)


def test_constants_match_header():
    """Every integer #define of the header that the Python mirror names has the
    same value there; the field ids follow the mirror's field order."""
    hdr = open(os.path.join(ROOT, "include", "tealeaf_hip.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define\s+(TL_[A-Z0-9_]+)\s+(-?\d+)\b", hdr)}
    assert len(defs) > 50
    named = 0
    for k, v in defs.items():
        if k.startswith("TL_F_"):
            assert _lib.FIELDS[v] == k[len("TL_F_"):].lower(), (k, v)
            named += 1
            continue
        obj = getattr(_lib, k, getattr(T, k, None))
        if obj is not None:
            assert int(obj) == v, (k, v, obj)
            named += 1
    assert defs["TL_NFIELDS"] == len(_lib.FIELDS)
    assert defs["TL_K_COUNT"] == _lib.TL_K_COUNT and set(_lib.KERNEL_NAMES) >= set(range(_lib.TL_K_COUNT))
    assert named >= len(defs) - 3, named


def test_struct_sizes_match_header():
    # the ctypes mirrors must agree with the C layout the library was built with
    p = _lib.SolveParams()
    T.load().tl_default_params(C.byref(p))
    assert (p.solver, p.coefficient, p.max_iters, p.presteps, p.ppcg_inner_steps) == (0, 1, 10000, 30, 10)
    assert p.eps == 1e-15 and p.epslim == 1e-5 and p.rel_tol == 1e-10 and p.check_result == 1


def test_struct_layouts_match_header(tmp_path):
    # every member offset and struct size of the ctypes mirrors against the
    # header as a C compiler lays it out (gcc, the same ABI as the library)
    mirrors = {"tl_solve_params": _lib.SolveParams, "tl_solve_stats": _lib.SolveStats, "tl_state": _lib.State,
               "tl_deck": _lib.Deck, "tl_train_params": _lib.TrainParams}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "tealeaf_hip.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            cname, fname, v = ln.split()
            got[(cname, fname)] = int(v)
    for cname, py in mirrors.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


@pytest.mark.parametrize("text", [
    decks.reference_deck(512, 10),
    decks.reference_deck(4096, 20, solver="ppcg", presteps=8),
    decks.random_rect_deck(1001, n=8192, steps=20),
    decks.random_rect_deck(3, n=100, steps=2, nrect=5),
])
def test_parser_agrees_with_oracle(text):
    s = T.Settings.from_text(text)
    d = O.parse_deck(text)
    assert (s.xcells, s.ycells, s.halodepth) == (d.xcells, d.ycells, d.halodepth)
    assert s.dx == d.dx and s.dy == d.dy and s.dtinit == d.dtinit and s.eps == d.eps
    assert s.maxiters == d.maxiters and s.presteps == d.presteps and s.ppcginnersteps == d.ppcginnersteps
    assert s.solver.name == d.solver and s.endstep == d.endstep
    assert len(s.states) == len(d.states)
    for a, b in zip(s.states, d.states):
        for k in ("density", "energy", "xmin", "ymin", "xmax", "ymax", "radius", "geometry"):
            assert a[k] == getattr(b, k), (k, a[k], getattr(b, k))


def test_parser_build_keys_and_warnings():
    s = T.Settings.from_text(decks.random_rect_deck(1, n=64, steps=1, precision="fp32_ir", warm_start="learned")
                             + "bogus_key=1\nuse_chebyshev\n")
    assert s.precision == T.TL_FP32_IR and s.warm_start == T.TL_WARM_LEARNED
    assert s.stop_rule == T.TL_STOP_RELATIVE and s.rel_tol == 1e-10
    assert s.n_warnings == 2  # unknown key and unknown solver: warned, ignored (settings.jl:109,127)
    with pytest.raises(T.TeaLeafError):
        T.Settings.from_text("x_cells=4\n")  # no state 1


def test_parser_mixed_precision_keys():
    base = decks.random_rect_deck(1, n=64, steps=1)
    s = T.Settings.from_text(base)
    assert s.precision == T.TL_FP64 and s.ir_delta == 0.1  # defaults
    s = T.Settings.from_text(base + "precision = fp32_ir\nir_delta=0.3\nrel_tol=1e-9\nstop_rule=relative\n")
    assert s.precision == T.TL_FP32_IR and s.ir_delta == 0.3 and s.rel_tol == 1e-9
    assert s.stop_rule == T.TL_STOP_RELATIVE
    assert s._deck.params.fp32_kernels == _lib.TL_FP32_KERNELS_AUTO
    for v, code in (("two", _lib.TL_FP32_KERNELS_TWO), ("fused", _lib.TL_FP32_KERNELS_FUSED),
                    ("pairs", _lib.TL_FP32_KERNELS_PAIRS), ("auto", _lib.TL_FP32_KERNELS_AUTO)):
        assert T.Settings.from_text(base + f"fp32_kernels={v}\n")._deck.params.fp32_kernels == code
    assert s._deck.params.warm_policy == _lib.TL_WARM_POLICY_BACKOFF
    assert T.Settings.from_text(base + "warm_policy=every\n")._deck.params.warm_policy == _lib.TL_WARM_POLICY_EVERY
    assert s._deck.params.ppcg_pairs == _lib.TL_PPCG_PAIRS_AUTO
    assert T.Settings.from_text(base + "ppcg_pairs=off\n")._deck.params.ppcg_pairs == _lib.TL_PPCG_PAIRS_OFF
    assert s._deck.params.cheby_pairs == _lib.TL_CHEBY_PAIRS_AUTO
    assert T.Settings.from_text(base + "cheby_pairs=off\n")._deck.params.cheby_pairs == _lib.TL_CHEBY_PAIRS_OFF
    n0 = T.Settings.from_text(base).n_warnings
    for bad in ("precision=fp16\n", "ir_delta=abc\n", "stop_rule=sometimes\n", "warm_start=maybe\n",
                "fp32_kernels=three\n", "warm_policy=sometimes\n", "ppcg_pairs=3\n", "cheby_pairs=2\n"):
        assert T.Settings.from_text(base + bad).n_warnings == n0 + 1, bad  # warned and ignored


def test_eigen_bounds_match_scipy_and_oracle():
    rng = np.random.default_rng(2)
    L = T.load()
    for n in (1, 2, 7, 30):
        a = rng.uniform(0.05, 1.0, n)
        b = rng.uniform(0.05, 0.9, n)
        mn, mx = C.c_double(), C.c_double()
        assert L.tl_eigen_bounds(_dp(a), _dp(b), n, C.byref(mn), C.byref(mx)) == 0
        d = 1 / a + np.r_[0, b[:-1] / a[:-1]]
        e = np.sqrt(b[:-1]) / a[:-1]
        ev = sla.eigvalsh_tridiagonal(d, e) if n > 1 else d
        assert abs(mn.value - 0.95 * ev[0]) <= 1e-12 * ev[0]
        assert abs(mx.value - 1.05 * ev[-1]) <= 1e-12 * ev[-1]
        omn, omx = O.eigenvalues(np.r_[0, a], np.r_[0, b], n)
        assert abs(omn - mn.value) <= 1e-12 * omn and abs(omx - mx.value) <= 1e-12 * omx


def test_cheby_coef_and_calciter_match_oracle():
    L = T.load()
    n = 40
    a, b, th = np.zeros(n), np.zeros(n), C.c_double()
    L.tl_cheby_coef(1.3, 870.0, n, _dp(a), _dp(b), C.byref(th))
    oa, ob, oth = O.cheby_coef(1.3, 870.0, n)
    np.testing.assert_array_equal(a, oa[1:n + 1])
    np.testing.assert_array_equal(b, ob[1:n + 1])
    assert th.value == oth
    for args in ((1.2, 880.0, 3.3e-2, 4.1e3), (0.9, 55.0, 1e-3, 1e2), (1.0, 2.0e6, 5.0, 1e9)):
        assert L.tl_cheby_calciter(*args) == O.calciter(*args)
    # known answers (condition number 9: gamma = 1/2; eps bb / (4 err) = 2^-20 -> 10)
    assert L.tl_cheby_calciter(1.0, 9.0, 1.0, 2.0 ** 34) == 10
    assert L.tl_cheby_calciter(1.0, 9.0, 2.0 ** -6, 2.0 ** 34) == 7


def test_invalid_arguments_return_codes():
    L = T.load()
    ctx = C.c_void_p()
    assert L.tl_create(C.byref(ctx), 0, 10, 2, 0.0, 0.0, 1.0, 1.0, 0) == _lib.TL_E_INVALID
    assert L.tl_create(C.byref(ctx), 10, 10, 0, 0.0, 0.0, 1.0, 1.0, 0) == _lib.TL_E_INVALID
    assert L.tl_create(C.byref(ctx), 10, 10, 1, 0.0, 0.0, 1.0, 1.0, 0) == _lib.TL_E_INVALID  # depth-1 halo
    assert L.tl_create(C.byref(ctx), 10, 10, 2, 0.0, 0.0, -1.0, 1.0, 0) == _lib.TL_E_INVALID


def test_julia_float_format():
    # Julia's print(::Float64): shortest digits, ".0" on integral values,
    # "e" notation outside [1e-4, 1e6) (not checked against Julia: none here)
    from tealeaf_hip import _julia_float as j
    cases = {0.1: "0.1", 1e-5: "1.0e-5", 1e-4: "0.0001", 1e5: "100000.0", 1e6: "1.0e6", 1234567.0: "1.234567e6",
             2.5: "2.5", 100.0: "100.0", -3e-7: "-3.0e-7", 0.0: "0.0", 9.99e-5: "9.99e-5"}
    for x, want in cases.items():
        assert j(x) == want, (x, j(x), want)
    for x in np.random.default_rng(0).standard_normal(200) * 10.0 ** np.random.default_rng(1).integers(-8, 9, 200):
        assert float(j(float(x))) == x  # round trip


def test_kernel_names_follow_coefficient_format():
    # the FP32 kernels' last template argument is the 16-bit coefficient format
    assert _lib.kernel_name(_lib.TL_K_CG_UR32) == "k_cgg<4, true, 4, true, 1>"
    assert _lib.kernel_name(_lib.TL_K_CG_UR32, _lib.TL_COEF_FP16) == "k_cgg<4, true, 4, true, 1>"
    assert _lib.kernel_name(_lib.TL_K_CG_UR32, _lib.TL_COEF_BF16) == "k_cgg<4, true, 4, true, 0>"
    assert _lib.kernel_name(_lib.TL_K_CG_W32, _lib.TL_COEF_BF16) == "k_cgg<4, false, 2, true, 0>"
    assert _lib.kernel_name(_lib.TL_K_CG_UR64, _lib.TL_COEF_BF16) == _lib.KERNEL_NAMES[_lib.TL_K_CG_UR64]
    assert _lib.kernel_name(_lib.TL_K_CGF32) == "k_cgf<1, 1, 0>"
    assert _lib.kernel_name(_lib.TL_K_CGF32, _lib.TL_COEF_BF16) == "k_cgf<1, 0, 0>"


def test_benchmark_problem_is_fixed():
    # SplitMix64 known-answer vector (seed 0) and the benchmark deck itself:
    # a change of the generator or of the deck text changes the measured problem
    import hashlib
    g = decks.SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    text = decks.random_rect_deck(1001, n=8192, steps=1000, precision="fp32_ir", warm_start="learned")
    assert hashlib.sha256(text.encode()).hexdigest() == \
        "3a117da74cbbfd7a139452e2efe5f47c023044df222040def0fa82fae7366f86"
    s = T.Settings.from_text(text)
    assert (s.xcells, s.ycells) == (8192, 8192) and s.precision == T.TL_FP32_IR


def _check(solver=_lib.TL_CG, precision=_lib.TL_FP64, stop_rule=_lib.TL_STOP_RELATIVE, rel_tol=1e-10, eps=1e-15,
           iterations=100, converged=1, bb=4.0, u0sq=4.0, resid2=0.0):
    p = _lib.SolveParams()
    T.load().tl_default_params(C.byref(p))
    p.solver, p.precision, p.stop_rule, p.rel_tol, p.eps = solver, precision, stop_rule, rel_tol, eps
    st = _lib.SolveStats()
    st.iterations, st.converged = iterations, converged
    return T.load().tl_check_solution(C.byref(p), C.byref(st), bb, u0sq, resid2)


def test_convergence_check_rejects_stale_reductions():
    """tl_solve's last step: a convergence claim is checked against the FP64
    residual of the returned u recomputed by separate kernels.  A reduction
    that read zeros instead of its partials -- the solve's sum u0^2, the
    residual at CG-init taken for an exact start, or an FP64 refresh taken for
    convergence -- must come back as TL_E_CHECK, never as a converged solve."""
    ok, bad = _lib.TL_OK, _lib.TL_E_CHECK
    # relative rule, ||u0|| = 2: the threshold is 2e-10
    assert _check(resid2=(1.9e-10) ** 2) == ok
    assert _check(resid2=(2.0e-10) ** 2) == ok                         # on the threshold
    assert _check(resid2=(2.0e-10 * (1 + 1e-8)) ** 2) == ok            # summation-order slack
    assert _check(resid2=(2.1e-10) ** 2) == bad                        # claimed, not met
    assert _check(precision=_lib.TL_FP32_IR, resid2=(3e-10) ** 2) == bad
    assert _check(resid2=(3e-10) ** 2, converged=0) == ok               # max_iters: no claim
    # the solve's sum u0^2 read as zero (stale partials) while the field is not zero
    assert _check(bb=0.0, resid2=0.0) == bad
    assert _check(bb=4.0 * (1 + 1e-12)) == ok
    assert _check(bb=-1.0, solver=_lib.TL_JACOBI) == ok                # Jacobi forms no bb
    # exact start (0 iterations): the recomputed residual must meet the rule
    for solver in (_lib.TL_CG, _lib.TL_CHEBY, _lib.TL_PPCG):
        assert _check(solver=solver, stop_rule=_lib.TL_STOP_REFERENCE, iterations=0, resid2=0.0) == ok
        assert _check(solver=solver, stop_rule=_lib.TL_STOP_REFERENCE, iterations=0, resid2=1e-2) == bad
    assert _check(iterations=0, resid2=1e-2) == bad
    assert _check(iterations=0, resid2=(1e-10) ** 2) == ok
    # the reference rule trusts the recursive residual after iteration 0 (CG.jl:24)
    assert _check(stop_rule=_lib.TL_STOP_REFERENCE, iterations=5, resid2=1e-2) == ok
    # NaN sums never pass
    assert _check(resid2=float("nan")) == bad
    assert _check(u0sq=float("nan"), bb=-1.0) == bad
