"""Pins for the CPU oracle (runs without a GPU).

The reference ships no test vectors (no tests, tea.in / tea.problems are
git-ignored), so the oracle is pinned by what the mathematics fixes: closed
forms, conservation, symmetry, brute-force sparse solves, textbook
eigenvalue and Chebyshev routines, and the closed form of the reference's
as-written (literal) CG.  Each check is chosen so that a plausible slip --
a wrong halo index, a dropped stencil term, a transposed coefficient, an
off-by-one in the Lanczos matrix -- fails it.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from tealeaf_hip import decks


def _uniform_deck(n, rho, e, steps=3, solver="cg", extra=""):
    return (f"state 1 density={rho!r} energy={e!r}\nx_cells={n}\ny_cells={n}\nxmin=0.0\nymin=0.0\n"
            f"xmax=10.0\nymax=10.0\ninitial_timestep=0.004\nend_step={steps}\nuse_{solver}\n"
            f"presteps=30\nmax_iters=5000\n" + extra)


def _make(text, threads=4, **over):
    d = O.parse_deck(text, literal=over.get("literal", 0))
    for k, v in over.items():
        setattr(d, k, v)
    o = O.Oracle(d, threads=threads)
    o.initialise()
    return o


def _assemble(o):
    """Sparse A of the interior with the reflective (Neumann) boundary folded
    in: A = I + sum over interior faces of K (e_i - e_j)(e_i - e_j)^T."""
    h = o.hd
    ny, nx = o.Y - 2 * h, o.X - 2 * h
    kx, ky = o.field("kx"), o.field("ky")
    idx = lambda x, y: y * nx + x
    rows, cols, vals = [], [], []
    diag = np.ones(nx * ny)
    for y in range(ny):
        for x in range(nx):
            i = idx(x, y)
            if x + 1 < nx:
                k = kx[y + h, x + 1 + h]
                j = idx(x + 1, y)
                diag[i] += k
                diag[j] += k
                rows += [i, j]
                cols += [j, i]
                vals += [-k, -k]
            if y + 1 < ny:
                k = ky[y + 1 + h, x + h]
                j = idx(x, y + 1)
                diag[i] += k
                diag[j] += k
                rows += [i, j]
                cols += [j, i]
                vals += [-k, -k]
    A = sp.coo_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny)).tocsr() + sp.diags(diag)
    return A.tocsr()


# ------------------------------------------------------------------ deck


def test_parse_reference_deck():
    d = O.parse_deck(decks.reference_deck(512, 10))
    assert (d.xcells, d.ycells, d.halodepth, d.endstep, d.maxiters) == (512, 512, 2, 10, 10000)
    assert d.dx == 10 / 512 and d.dtinit == 0.004 and d.eps == 1e-15 and d.solver == "cg"
    assert len(d.states) == 3
    s2 = d.states[1]
    assert s2.xmin == 0.0 + d.dx / 100 and s2.xmax == 1.0 - d.dx / 100
    assert s2.ymin == 1.0 + d.dy / 100 and s2.ymax == 2.0 - d.dy / 100
    # keys the reference does not know are warned about and ignored
    d2 = O.parse_deck("state 1 density=1 energy=1\nx_cells=4\ny_cells=4\nbogus_key=3\nuse_chebyshev\ncheck_result\n")
    assert len(d2.warnings) == 2 and d2.solver == "cg"


def test_parse_literal_d3_offsets():
    d = O.parse_deck(decks.reference_deck(512, 10), literal=O.LIT_D3)
    assert d.states[1].xmin == 0.0 + 10 / 100  # dx still at its default 10 while parsing


# ------------------------------------------------------------ rasterise


def _block(o, density):
    m = o.interior("density") == density
    ys, xs = np.nonzero(m)
    return m.sum(), (xs.min(), xs.max(), ys.min(), ys.max()) if len(xs) else None


@pytest.mark.parametrize("literal,counts,temp", [
    (0, (51 * 52, 257 * 52), None),
    (O.LIT_D3, (42 * 42, 247 * 42), None),
])
def test_rasterise_counts_and_initial_temp(literal, counts, temp):
    o = _make(decks.reference_deck(512, 1), literal=literal)
    e0 = o.interior("energy0")
    n2 = int(((e0 == 25.0)).sum())
    n3 = int(((e0 == 0.1)).sum())
    assert (n2, n3) == counts
    # initial temp = vol * sum(rho * rho * e) from the counts alone
    vol = (10 / 512) ** 2
    nb = 512 * 512 - n2 - n3
    expect = vol * (nb * 100 * 100 * 1e-4 + n2 * 0.1 * 0.1 * 25.0 + n3 * 0.1 * 0.1 * 0.1)
    got = o.summary()[3]
    assert abs(got / expect - 1) < 1e-12
    if literal == 0:
        assert abs(got - 94.14839) < 1e-5


def test_rasterise_literal_d4_transposes():
    t = decks.reference_deck(64, 1)
    a = _make(t).field("density")
    b = _make(t, literal=O.LIT_D4).field("density")
    assert np.array_equal(a, b.T)
    assert not np.array_equal(a, b)


def test_circle_state():
    t = ("state 1 density=1 energy=1\nstate 2 density=5 energy=2 geometry=circular xmin=5.0 ymin=5.0 radius=2.0\n"
         "x_cells=50\ny_cells=50\nxmin=0\nymin=0\nxmax=10\nymax=10\n")
    o = _make(t)
    m = o.interior("density") == 5
    c = (np.arange(50) + 0.5) * 0.2
    xx, yy = np.meshgrid(c, c)
    expect = (xx - (5 + 0.2 / 100)) ** 2 + (yy - (5 + 0.2 / 100)) ** 2 <= 4.0
    assert np.array_equal(m, expect)


# -------------------------------------------------------------- operator


def test_uniform_deck_closed_form():
    rho, e = 2.5, 0.3
    o = _make(_uniform_deck(32, rho, e))
    s, st = o.run(3)
    for t in range(3):
        np.testing.assert_allclose(s[t], [100.0, 100 * rho, 100 * rho * e, 100 * rho * rho * e], rtol=1e-12)


def test_cosine_eigenmode():
    n, rho = 40, 2.0
    o = _make(_uniform_deck(n, rho, 1.0))
    rx, ry = o.rxry()
    o.cg_init()
    m, l = 3, 2
    c = (np.arange(n) + 0.5) / n
    b = np.outer(np.cos(np.pi * l * c), np.cos(np.pi * m * c))  # [y, x]
    h = o.hd
    o.field("p")[h:-h, h:-h] = b
    o.halo_update(O.X_P)
    o.smvp("p", "w")
    K = rx / rho
    lam = 1 + 2 * K * (1 - np.cos(np.pi * m / n)) + 2 * (ry / rho) * (1 - np.cos(np.pi * l / n))
    np.testing.assert_allclose(o.interior("w"), lam * b, atol=1e-12 * lam)
    # the CG solve of A u = b recovers b / lam
    o2 = _make(_uniform_deck(n, rho, 1.0))
    o2.field("energy")[h:-h, h:-h] = b / rho
    o2.halo_update(O.X_ENERGY)
    o2.step()
    np.testing.assert_allclose(o2.interior("u"), b / lam, atol=1e-12)


def test_operator_symmetric_unit_row_sums():
    o = _make(decks.random_rect_deck(3, n=36, steps=1))
    o.cg_init()
    rng = np.random.default_rng(0)
    h = o.hd
    x = rng.standard_normal((36, 36))
    y = rng.standard_normal((36, 36))
    o.field("p")[h:-h, h:-h] = x
    o.halo_update(O.X_P)
    o.smvp("p", "w")
    Ax = o.interior("w").copy()
    o.field("p")[h:-h, h:-h] = y
    o.halo_update(O.X_P)
    o.smvp("p", "w")
    Ay = o.interior("w").copy()
    assert abs(np.vdot(Ax, y) - np.vdot(x, Ay)) < 1e-12 * abs(np.vdot(Ax, y))
    o.field("p")[h:-h, h:-h] = 1.0
    o.halo_update(O.X_P)
    o.smvp("p", "w")
    np.testing.assert_allclose(o.interior("w"), 1.0, rtol=1e-12)


def test_residual_matches_assembled_operator():
    # residual!, kernels.jl:227-232: r = u0 - A u on the interior, A with the
    # reflective boundary; checked against the independently assembled matrix
    o = _make(decks.random_rect_deck(4, n=20, steps=1))
    o.cg_init()
    rng = np.random.default_rng(5)
    h = o.hd
    u = rng.standard_normal((20, 20))
    u0 = rng.standard_normal((20, 20))
    o.field("u")[h:-h, h:-h] = u
    o.field("u0")[h:-h, h:-h] = u0
    o.halo_update(O.X_U)
    o.residual()
    want = u0.ravel() - _assemble(o) @ u.ravel()
    np.testing.assert_allclose(o.interior("r").ravel(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_halo_intended_and_literal_indices():
    t = _uniform_deck(6, 1.0, 1.0)
    for lit in (0, O.LIT_D2):
        o = _make(t, literal=lit)
        g = o.field("sd")
        g[:] = np.arange(g.size, dtype=float).reshape(g.shape)
        ref = g.copy()
        o.halo_update(O.X_SD)
        h = 2
        X = o.X
        for j in range(h, o.Y - h):
            if lit == 0:
                assert g[j, h - 1] == ref[j, h] and g[j, X - h] == ref[j, X - h - 1]
                assert g[j, 0] == ref[j, 0]  # depth 1 leaves layer 2 alone
            else:
                # as written: left ghost hd-kk (1-based) = layer 2, right mirrors x-hd-1-kk
                assert g[j, h - 2] == ref[j, h] and g[j, h - 1] == ref[j, h - 1]
                assert g[j, X - h] == ref[j, X - h - 3]


# ---------------------------------------------------------- brute force


@pytest.mark.parametrize("solver", ["cg", "cheby", "ppcg", "jacobi"])
def test_solvers_match_sparse_direct_solve(solver):
    text = decks.random_rect_deck(9, n=24, steps=1).replace("use_cg", f"use_{solver}")
    o = _make(text, maxiters=20000, presteps=30)
    b = o.interior("energy") * o.interior("density")
    o.step()
    A = _assemble(o)
    u = spla.spsolve(A, b.ravel()).reshape(b.shape)
    np.testing.assert_allclose(o.interior("u"), u, rtol=1e-8, atol=1e-10 * np.abs(u).max())


def test_cross_solver_agreement_and_conservation():
    res = {}
    for solver in ("cg", "cheby", "ppcg", "jacobi"):
        text = decks.reference_deck(64, 3, solver=solver, presteps=30, max_iters=20000)
        s, st = O.run_deck(text, threads=4)[1]
        res[solver] = s
        # ie = sum vol*rho*energy = sum vol*u is conserved: 1^T A = 1^T
        np.testing.assert_allclose(s[:, 2], s[0, 2], rtol=1e-11)
    for solver in ("cheby", "ppcg", "jacobi"):
        np.testing.assert_allclose(res[solver], res["cg"], rtol=1e-9)


def test_ppcg_without_inner_steps_is_cg():
    # PPCG stops on |rr| < eps (PPCG.jl:50), CG on sqrt|rr| < eps (CG.jl:24):
    # with eps = 1e-15 PPCG leaves ||r|| up to ~3e-8, hence the 1e-9 tolerance
    a = O.run_deck(decks.reference_deck(48, 2, solver="ppcg", presteps=30, ppcg_inner_steps=0), threads=4)[1][0]
    b = O.run_deck(decks.reference_deck(48, 2, solver="cg"), threads=4)[1][0]
    np.testing.assert_allclose(a, b, rtol=1e-9)


def test_reductions_deterministic_across_threads():
    t = decks.reference_deck(96, 2)
    a = O.run_deck(t, threads=1)[1][0]
    b = O.run_deck(t, threads=8)[1][0]
    assert np.array_equal(a, b)


# ---------------------------------------------------- eigen / Chebyshev


def test_tqli_matches_scipy():
    rng = np.random.default_rng(1)
    for n in (1, 2, 5, 17, 40):
        d = rng.uniform(1, 100, n)
        e = rng.uniform(-5, 5, max(n - 1, 0))
        np.testing.assert_allclose(O.tqli(d, e), sla.eigvalsh_tridiagonal(d, e), rtol=1e-12, atol=1e-12)


def test_eigenvalues_from_cg_history_bound_the_spectrum():
    # Lanczos tridiagonal of CG's alpha/beta: Ritz values inside the spectrum of A
    n = 25
    o = _make(decks.random_rect_deck(4, n=20, steps=1), maxiters=n)
    st = o.step()
    assert st["iterations"] == n
    ca, cb = o.history(0, n), o.history(1, n)
    mn, mx = O.eigenvalues(ca, cb, n)
    # textbook: eigenvalues of the same tridiagonal (kernels.jl:20-31)
    d = 1 / ca[1:n + 1] + np.r_[0, cb[1:n] / ca[1:n]]
    e = np.sqrt(cb[1:n]) / ca[1:n]
    ev = sla.eigvalsh_tridiagonal(d, e)
    assert abs(mn - 0.95 * ev[0]) < 1e-10 * ev[0] and abs(mx - 1.05 * ev[-1]) < 1e-10 * ev[-1]
    lam = np.linalg.eigvalsh(_assemble(o).toarray())
    assert lam[0] * (1 - 1e-9) <= ev[0] and ev[-1] <= lam[-1] * (1 + 1e-9)


def test_cheby_coef_closed_form():
    emin, emax, n = 0.9, 420.0, 25
    a, b, th = O.cheby_coef(emin, emax, n)
    delta = (emax - emin) / 2
    sigma = th / delta
    T = lambda k: np.cosh(k * np.arccosh(sigma))
    for k in range(1, n + 1):
        assert abs(a[k] - T(k - 1) / T(k + 1)) < 1e-12
        assert abs(b[k] - 2 * T(k) / (delta * T(k + 1))) < 1e-12 * abs(b[k])
    assert th == (emin + emax) / 2


def test_calciter_known_answers():
    """Cheby.calciter (Cheby.jl:109-118) on hand-computed cases.  With
    condition number 9, gamma = (3 - 1) / (3 + 1) = 1/2; eps(Float64) = 2^-52,
    so bb = 2^34, err = 1 gives eps bb / (4 err) = 2^-20 and the estimate is
    log(2^-20) / (2 log(1/2)) = 10 in any log base.  A dropped factor 2 gives
    20, an inverted ratio -10, gamma taken as 1/gamma -10."""
    assert O.calciter(1.0, 9.0, 1.0, 2.0 ** 34) == 10
    assert O.calciter(0.5, 4.5, 1.0, 2.0 ** 34) == 10      # only the ratio eigmax/eigmin matters
    assert O.calciter(1.0, 9.0, 2.0 ** -6, 2.0 ** 34) == 7  # err 2^-6: 2^-14 -> 7
    # condition number 4: gamma = 1/3; eps bb / (4 err) = 3^-8 -> 4
    bb = 3.0 ** -8 * 4 / np.finfo(float).eps
    assert O.calciter(1.0, 4.0, 1.0, bb) == 4


def test_calciter_rounding_bound():
    """The estimate k is the nearest integer to x = ln(ita) / (2 ln gamma), so
    |2 k ln(gamma) - ln(ita)| = 2 |ln gamma| |k - x| <= |ln gamma| for every
    input (ita = eps bb / (4 err)); checked on random inputs spanning the
    solver's range of condition numbers, errors and norms."""
    rng = np.random.default_rng(12)
    eps = np.finfo(float).eps
    for _ in range(2000):
        emin = 10.0 ** rng.uniform(-3, 1)
        kappa = 10.0 ** rng.uniform(0.05, 7)
        err = 10.0 ** rng.uniform(-12, 3)
        bb = 10.0 ** rng.uniform(-3, 12)
        k = O.calciter(emin, emin * kappa, err, bb)
        g = (np.sqrt(kappa) - 1) / (np.sqrt(kappa) + 1)
        ita = eps * bb / (4 * err)
        assert abs(2 * k * np.log(g) - np.log(ita)) <= abs(np.log(g)) * (1 + 1e-9), (emin, kappa, err, bb, k)


# ------------------------------------------------- literal-mode closed form


def test_literal_d1_closed_form():
    """As written, smvp applies only D = 1 + kx[k+1] + kx[k] + ky[j+1] + ky[j],
    so each step's CG solves D u = b cell by cell: u_n = u_0 / D^n."""
    text = decks.reference_deck(40, 3)
    o = _make(text, literal=O.LIT_D1)
    u0 = o.interior("energy") * o.interior("density")
    rho = o.interior("density")
    s, st = o.run(3)
    h = o.hd
    kx, ky = o.field("kx"), o.field("ky")
    D = (1 + kx[h:-h, h + 1:o.X - h + 1] + kx[h:-h, h:-h] + ky[h + 1:o.Y - h + 1, h:-h] + ky[h:-h, h:-h])
    vol = o.deck.dx * o.deck.dy
    for t in range(3):
        expect = np.sum(vol * rho * u0 / D ** (t + 1))
        assert abs(s[t, 3] / expect - 1) < 1e-11
