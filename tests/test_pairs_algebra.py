"""The scalar algebra of the FP32 inner CG in pairs of iterations (k_cgp),
checked in FP64 against two plain CG iterations.

From r = r(t-1) and p' = p(t-1), the pair forms p = r + beta p' and takes
its scalars from six inner products of the basis p, A p, A^2 p, r, A r:
    alpha0 = gamma / (p, Ap),  gamma1 = gamma - 2 alpha0 (r, Ap) + alpha0^2 (Ap, Ap),
    beta0 = gamma1 / gamma,
    (p1, A p1) = (r, Ar) - 2 alpha0 (Ap, Ar) + alpha0^2 (Ap, A^2 p)
                 + 2 beta0 ((r, Ap) - alpha0 (Ap, Ap)) + beta0^2 (p, Ap),
    alpha1 = gamma1 / (p1, A p1).
Two CG iterations computed the textbook way, one vector at a time, must give
the same scalars and the same r(t+1), p(t+1), x to rounding.  The operator is
the 5-point conduction stencil with zero-flux edges and random face
coefficients, an SPD matrix like the solver's.  Also checked: the null second
step of a pair that stops after its first (alpha1 = 0, p(t) kept, gamma(t-1)
kept) leaves the state one CG iteration gives, and the pair after a refresh
from that state again equals two CG iterations."""
import numpy as np
import pytest


def _operator(nx, ny, seed):
    rng = np.random.default_rng(seed)
    kx = np.zeros((ny, nx + 1))
    ky = np.zeros((ny + 1, nx))
    kx[:, 1:nx] = rng.uniform(0.1, 50.0, (ny, nx - 1))
    ky[1:ny, :] = rng.uniform(0.1, 50.0, (ny - 1, nx))

    def A(v):
        v = v.reshape(ny, nx)
        out = v.copy()
        out[:, :-1] += kx[:, 1:nx] * (v[:, :-1] - v[:, 1:])
        out[:, 1:] += kx[:, 1:nx] * (v[:, 1:] - v[:, :-1])
        out[:-1, :] += ky[1:ny, :] * (v[:-1, :] - v[1:, :])
        out[1:, :] += ky[1:ny, :] * (v[1:, :] - v[:-1, :])
        return out.ravel()
    return A


def _pair(A, r, pp, beta):
    p = r + beta * pp
    Ap, Ar = A(p), A(r)
    A2p = A(Ap)
    gam = r @ r
    pap, rap, apap, rar, apar, apa2p = p @ Ap, r @ Ap, Ap @ Ap, r @ Ar, Ap @ Ar, Ap @ A2p
    a0 = gam / pap
    g1 = gam - 2 * a0 * rap + a0 * a0 * apap
    b0 = g1 / gam
    p1ap1 = (rar - 2 * a0 * apar + a0 * a0 * apa2p) + 2 * b0 * (rap - a0 * apap) + b0 * b0 * pap
    a1 = g1 / p1ap1
    r1 = r - a0 * Ap
    p1 = r1 + b0 * p
    ap1 = A(p1)
    return dict(a0=a0, g1=g1, b0=b0, a1=a1, p=p, p1=p1, r2=r1 - a1 * ap1, dx=a0 * p + a1 * p1, r1=r1)


def _cg_two(A, r, pp, beta):
    """Two textbook CG iterations from r and p' (p = r + beta p')."""
    p = r + beta * pp
    gam = r @ r
    q = A(p)
    a0 = gam / (p @ q)
    r1 = r - a0 * q
    g1 = r1 @ r1
    b0 = g1 / gam
    p1 = r1 + b0 * p
    q1 = A(p1)
    a1 = g1 / (p1 @ q1)
    return dict(a0=a0, g1=g1, b0=b0, a1=a1, p1=p1, r2=r1 - a1 * q1, dx=a0 * p + a1 * p1)


@pytest.mark.parametrize("nx,ny,seed", [(16, 12, 1), (33, 7, 2), (8, 40, 3)])
def test_pair_scalars_equal_two_cg_iterations(nx, ny, seed):
    A = _operator(nx, ny, seed)
    rng = np.random.default_rng(100 + seed)
    b = rng.standard_normal(nx * ny)
    # a CG state some iterations in: r, p' and beta from plain CG
    x = np.zeros_like(b)
    r = b.copy()
    p = r.copy()
    gam = r @ r
    for _ in range(5):
        q = A(p)
        a = gam / (p @ q)
        x += a * p
        r -= a * q
        g = r @ r
        beta = g / gam
        gam = g
        pp = p
        p = r + beta * p
    want = _cg_two(A, r, pp, beta)
    got = _pair(A, r, pp, beta)
    for k in ("a0", "g1", "b0", "a1"):
        assert got[k] == pytest.approx(want[k], rel=1e-9), (k, got[k], want[k])
    for k in ("p1", "r2", "dx"):
        scale = np.max(np.abs(want[k]))
        assert np.max(np.abs(got[k] - want[k])) <= 1e-9 * scale, k
    # the six-sum route is not a retyping of the two-step route: with a wrong
    # sign on the (Ap, Ar) term the second step's alpha is off
    p_ = r + beta * pp
    Ap, Ar = A(p_), A(r)
    bad = ((r @ Ar + 2 * got["a0"] * (Ap @ Ar) + got["a0"] ** 2 * (Ap @ A(Ap)))
           + 2 * got["b0"] * ((r @ Ap) - got["a0"] * (Ap @ Ap)) + got["b0"] ** 2 * (p_ @ Ap))
    assert abs(got["g1"] / bad - want["a1"]) > 1e-3 * abs(want["a1"])


def test_null_second_step_and_refresh_beta():
    """A pair that stops after its first step: alpha1 = 0, beta0 = 0, p(t)
    kept, gamma(t-1) kept.  The state equals one CG iteration (x, r) with
    p(t) as the direction; after a refresh that replaces r by b - A x, the
    next pair (beta = gamma_new / gamma(t-1), as the fused kernel forms it
    after a refresh at t) equals two CG iterations from there."""
    nx, ny = 20, 14
    A = _operator(nx, ny, 9)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(nx * ny)
    r = b.copy()
    pp = np.zeros_like(b)
    x = np.zeros_like(b)
    gam_prev = r @ r  # gamma(t-1)
    got = _pair(A, r, pp, 0.0)
    # null second step
    x_pair = x + got["a0"] * got["p"]
    r_pair = got["r1"]
    p_pair = got["p"]
    # one textbook iteration
    q = A(r)
    a = (r @ r) / (r @ q)
    assert np.allclose(x_pair, x + a * r, rtol=0, atol=1e-12 * np.max(np.abs(a * r)))
    assert np.allclose(r_pair, r - a * q, rtol=0, atol=1e-12 * np.max(np.abs(r)))
    assert np.array_equal(p_pair, r)
    # refresh: r = b - A x, and the beta formed from the kept gamma(t-1);
    # the next pair from that state equals two CG iterations from it
    r_new = b - A(x_pair)
    beta = (r_new @ r_new) / gam_prev
    want = _cg_two(A, r_new, p_pair, beta)
    nxt = _pair(A, r_new, p_pair, beta)
    for k in ("a0", "a1"):
        assert nxt[k] == pytest.approx(want[k], rel=1e-9)
