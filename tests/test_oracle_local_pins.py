"""Pins the oracle's LOCAL-mode (porous-infill) functions against the
mathematics: the Heaviside projection (TOP3D_XL.m:340, 380), the p-norm
local-volume constraint (335-336, 360), the sensitivity chain through both
PDE filters (341, 346, 350) and MMA (MMAseq/GenSub/SolveDIP, 1804-2049).

None of these checks restates the oracle's formulas: closed forms and limits
pin the values, central finite differences of exact (dense) compositions pin
every gradient, and MMA is pinned by the defining properties of its convex
subproblem (first-order agreement with the model at xval), the KKT conditions
of the returned primal-dual point, an independent nested-bisection solve of the
same subproblem, hand-computed asymptote updates and convergence to the analytic
optimum of a separable test problem.  A few checks also plant a plausible
mistake (on the CPU, in numpy) to show the bar would catch it.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import Oracle
from tests import bruteforce as bf

pytestmark = pytest.mark.usefixtures("oracle_lib")

ETA = 0.5  # etaPIO (234)


# ------------------------------------------------------------- Heaviside
@pytest.mark.parametrize("beta", [1.0, 4.0, 32.0, 128.0])
def test_heaviside_closed_forms(beta):
    xp, dx = Oracle.heaviside(beta, ETA, np.array([0.0, ETA, 1.0]))
    # the projection maps [0, 1] onto [0, 1] and the threshold to 1/2 (eta = 1/2)
    np.testing.assert_allclose(xp, [0.0, 0.5, 1.0], atol=1e-15)
    # slope at the threshold: beta / (tanh(beta eta) + tanh(beta (1 - eta))) = beta / (2 tanh(beta/2))
    assert abs(dx[1] - beta / (2 * math.tanh(beta / 2))) < 1e-13 * dx[1]
    for eta in (0.3, 0.7):
        xp2, _ = Oracle.heaviside(beta, eta, np.array([0.0, 1.0]))
        np.testing.assert_allclose(xp2, [0.0, 1.0], atol=1e-15)


def test_heaviside_limits_and_monotone():
    x = np.linspace(0, 1, 101)
    # beta -> 0: the identity
    xp, dx = Oracle.heaviside(1e-6, ETA, x)
    np.testing.assert_allclose(xp, x, atol=1e-11)
    np.testing.assert_allclose(dx, 1.0, atol=1e-11)
    # beta -> infinity: a step at eta
    xp, _ = Oracle.heaviside(1e4, ETA, x)
    assert np.all(xp[x < ETA - 0.01] < 1e-12) and np.all(xp[x > ETA + 0.01] > 1 - 1e-12)
    for beta in (1.0, 8.0, 64.0):
        xp, dx = Oracle.heaviside(beta, ETA, x)
        # non-decreasing (FP64 saturates far from eta for large beta), strictly near eta
        assert np.all(np.diff(xp) >= 0) and np.all(dx >= 0)
        mid = np.abs(x - ETA) < 0.1
        assert np.all(np.diff(xp[mid]) > 0) and np.all(dx[mid] > 0)


@pytest.mark.parametrize("beta", [1.0, 8.0, 64.0])
def test_heaviside_derivative_is_its_gradient(beta):
    rng = np.random.default_rng(3)
    x = rng.random(200)
    h = 1e-6
    xp_p, _ = Oracle.heaviside(beta, ETA, x + h)
    xp_m, _ = Oracle.heaviside(beta, ETA, x - h)
    _, dx = Oracle.heaviside(beta, ETA, x)
    fd = (xp_p - xp_m) / (2 * h)
    assert np.abs(fd - dx).max() < 1e-7 * max(1.0, np.abs(dx).max())


# ---------------------------------------------------------------- p-norm
def test_pnorm_closed_forms():
    n = 257
    for v in (0.4, 0.5):
        vm = np.full(n, v)
        for c in (0.3, 1.0, 1.7):
            f, _ = Oracle.pnorm(np.full(n, c * v), vm)
            assert abs(f - (c - 1.0)) < 1e-14
        # a single non-zero entry: (a^16 / N)^(1/16) / v - 1
        xh = np.zeros(n)
        xh[17] = 0.9
        f, _ = Oracle.pnorm(xh, vm)
        assert abs(f - (0.9 / v * n ** (-1.0 / 16) - 1.0)) < 1e-14
    # positively homogeneous of degree one in xhat / volMax
    rng = np.random.default_rng(5)
    xh, vm = rng.random(n), 0.3 + 0.4 * rng.random(n)
    f1, _ = Oracle.pnorm(xh, vm)
    f2, _ = Oracle.pnorm(2.5 * xh, vm)
    assert abs((f2 + 1) - 2.5 * (f1 + 1)) < 1e-14
    # a 16-norm mean lies between N^(-1/16) max and max
    m = (xh / vm).max()
    assert n ** (-1 / 16) * m <= f1 + 1 <= m


def test_pnorm_gradient_is_its_gradient():
    rng = np.random.default_rng(11)
    n = 300
    xh, vm = 0.2 + 0.6 * rng.random(n), np.full(n, 0.5)
    f, g = Oracle.pnorm(xh, vm)
    h = 1e-6
    idx = rng.choice(n, 40, replace=False)
    fd = np.empty(len(idx))
    for k, j in enumerate(idx):
        e = np.zeros(n)
        e[j] = h
        fd[k] = (Oracle.pnorm(xh + e, vm)[0] - Oracle.pnorm(xh - e, vm)[0]) / (2 * h)
    assert np.abs(fd - g[idx]).max() < 1e-7 * np.abs(g).max()
    # a plausible mistake (the exponent p - 1 dropped to p - 2) fails the bar
    wrong = g * (0.5 / xh)
    assert np.abs(fd - wrong[idx]).max() > 1e-2 * np.abs(g).max()


# ------------------------------------------------- sensitivity chains
def dense_filter(g, radius):
    """The PDE filter as a dense matrix, assembled from the Laplacian and mass
    matrices of the unit cube: element -> node T, (2 r^2 K_lap + M) solve,
    node -> element T^T, r = radius / (2 sqrt 3) (816-902)."""
    Kl, Mm = bf.laplace_mass_unit_cube()
    kern = 2 * (radius / (2 * math.sqrt(3))) ** 2 * Kl + Mm
    n, nE = g.nN[0], len(g.elems[0])
    Af = np.zeros((n, n))
    T = np.zeros((n, nE))
    for k, de in enumerate(g.elems[0]):
        nodes = g.node_index[g.elem_nodes(0, de)]
        Af[np.ix_(nodes, nodes)] += kern
        T[nodes, k] += 1.0 / 8
    return T.T @ np.linalg.solve(Af, T)


@pytest.fixture(scope="module")
def chain_model():
    vox = np.ones((4, 8, 4), bool)
    o = Oracle(vox)
    g = bf.Grid(vox)
    rmin, rhat = math.sqrt(3), 3.0
    o.filter_setup(0, rmin)
    o.filter_setup(1, rhat)
    return dict(o=o, g=g, Fmin=dense_filter(g, rmin), Fhat=dense_filter(g, rhat))


@pytest.mark.parametrize("beta", [1.0, 8.0])
def test_local_constraint_gradient_chain(chain_model, beta):
    """dfdx (346, 350) is the gradient of the p-norm constraint of
    xPhys = H(F_rMin(x)) filtered by F_rHat, as a function of x."""
    o, Fmin, Fhat = chain_model["o"], chain_model["Fmin"], chain_model["Fhat"]
    nE, ve0 = o.nE, 0.5
    vm = np.full(nE, ve0)

    def G(x):
        xp, _ = Oracle.heaviside(beta, ETA, Fmin @ x)
        return Oracle.pnorm(Fhat @ xp, vm)[0]

    rng = np.random.default_rng(17)
    x = 0.25 + 0.5 * rng.random(nE)
    xt = Fmin @ x
    xp, dx = Oracle.heaviside(beta, ETA, xt)
    f, _, dfdx, its = o.local_sens(beta, xp, xt, ve0, np.zeros(nE))
    assert its > 0
    assert abs(f - G(x)) < 1e-5
    h = 1e-5
    idx = rng.choice(nE, 24, replace=False)
    fd = np.array([(G(x + h * np.eye(nE)[j]) - G(x - h * np.eye(nE)[j])) / (2 * h) for j in idx])
    scale = np.abs(dfdx).max()
    # the filters are CG solves at tol 1e-4 (889-891): agreement to a few 1e-4
    assert np.abs(fd - dfdx[idx]).max() < 2e-3 * scale
    # planted mistakes fail the bar: the rMin filter of 350 left out, the rHat
    # filter of 346 left out, or dx left out
    _, dpde = Oracle.pnorm(Fhat @ xp, vm)
    for wrong in ((Fhat @ dpde) * dx, Fmin @ (dpde * dx)):
        assert np.abs(fd - wrong[idx]).max() > 2e-2 * scale
    if beta > 1:
        assert np.abs(fd - (Fmin @ (Fhat @ dpde))[idx]).max() > 2e-2 * scale


def test_local_compliance_gradient_chain(chain_model):
    """dc after 330 and 341 is the gradient of c / cSolid with respect to x
    for the SIMP model E = Emin + xPhys^3 (E0 - Emin) (107) of the projected,
    filtered design; U from a direct solve of the BC'd operator."""
    o, g, Fmin = chain_model["o"], chain_model["g"], chain_model["Fmin"]
    nE, beta, emin = o.nE, 4.0, 1e-6

    def compliance_of(xp):
        E = emin + xp ** 3 * (1 - emin)
        U = spla.spsolve(g.K_bc(E).tocsc(), g.F)
        return g.F @ U, U

    def C(x):
        xp, _ = Oracle.heaviside(beta, ETA, Fmin @ x)
        return compliance_of(xp)[0]

    rng = np.random.default_rng(23)
    x = 0.3 + 0.5 * rng.random(nE)
    xt = Fmin @ x
    xp, _ = Oracle.heaviside(beta, ETA, xt)
    c0, U = compliance_of(xp)
    ce = o.compliance(U)
    dc_raw = o.compliance_sens(xp, ce, 1.0)
    # first the sensitivity to xPhys itself
    h = 1e-6
    idx = rng.choice(nE, 12, replace=False)
    fd_p = np.array([(compliance_of(xp + h * np.eye(nE)[j])[0] - compliance_of(xp - h * np.eye(nE)[j])[0])
                     / (2 * h) for j in idx])
    assert np.abs(fd_p - dc_raw[idx]).max() < 1e-6 * np.abs(dc_raw).max()
    # then through the Heaviside projection and the rMin filter
    _, dc, _, _ = o.local_sens(beta, xp, xt, 0.5, dc_raw)
    h = 1e-5
    fd = np.array([(C(x + h * np.eye(nE)[j]) - C(x - h * np.eye(nE)[j])) / (2 * h) for j in idx])
    scale = np.abs(dc).max()
    assert np.abs(fd - dc[idx]).max() < 2e-3 * scale
    # cSolid normalises: dc scales as 1 / cSolid
    np.testing.assert_allclose(o.compliance_sens(xp, ce, 2.0), dc_raw / 2.0, rtol=1e-15)


# ------------------------------------------------------------------ MMA
def random_subproblem(rng, n, iter_=1, state=None):
    xval = 0.15 + 0.7 * rng.random(n)
    xmin, xmax = np.maximum(0.0, xval - 0.1), np.minimum(1.0, xval + 0.1)
    dfdx = rng.standard_normal(n)
    dgdx = 0.2 * rng.standard_normal(n) + 0.5 / n
    gx = float(rng.uniform(-0.05, 0.05))
    if state is None:
        state = dict(xold1=xval.copy(), xold2=xval.copy(), low=np.zeros(n), upp=np.zeros(n))
    return dict(iter_=iter_, xval=xval, xmin=xmin, xmax=xmax, dfdx=dfdx, gx=gx, dgdx=dgdx, **state)


def call_mma(p):
    return Oracle.mma_ex(p["iter_"], p["xval"], p["xmin"], p["xmax"], p["xold1"], p["xold2"], p["dfdx"],
                         p["gx"], p["dgdx"], p["low"], p["upp"])


def approx_terms(r, x):
    """f0~(x) - const and f1~(x) - const: the separable MMA approximations."""
    U, L = r["upp"], r["low"]
    f0 = np.sum(r["p0"] / (U - x) + r["q0"] / (x - L))
    f1 = np.sum(r["pij"] / (U - x) + r["qij"] / (x - L))
    return f0, f1


def test_mma_subproblem_is_a_first_order_model():
    """GenSub (1873-1928): the approximations are convex, match the objective
    and constraint gradients at xval, and the constraint value there is gx."""
    rng = np.random.default_rng(29)
    for _ in range(5):
        p = random_subproblem(rng, 64)
        r = call_mma(p)
        x0, U, L = p["xval"], r["upp"], r["low"]
        assert np.all(L < r["alpha"]) and np.all(r["alpha"] <= x0) and np.all(x0 <= r["beta"])
        assert np.all(r["beta"] < U)
        assert np.all(r["alpha"] >= p["xmin"]) and np.all(r["beta"] <= p["xmax"])
        for k in ("p0", "q0", "pij", "qij"):
            assert np.all(r[k] > 0)
        g0 = r["p0"] / (U - x0) ** 2 - r["q0"] / (x0 - L) ** 2
        g1 = r["pij"] / (U - x0) ** 2 - r["qij"] / (x0 - L) ** 2
        np.testing.assert_allclose(g0, p["dfdx"], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(g1, p["dgdx"], rtol=1e-12, atol=1e-14)
        _, f1 = approx_terms(r, x0)
        assert abs(f1 - r["b"] - p["gx"]) < 1e-12 * max(1.0, abs(f1))


def test_mma_asymptotes_hand_computed():
    """Asymptote initialisation (iter < 3) and adaptation (iter >= 3) of
    GenSub, 1880-1901, on values worked out by hand."""
    xval = np.array([0.5, 0.5, 0.5, 0.5, 0.05])
    xmin, xmax = np.array([0.4] * 4 + [0.0]), np.array([0.6] * 4 + [0.15])
    xold1 = np.array([0.4, 0.6, 0.5, 0.4, 0.1])
    xold2 = np.array([0.3, 0.3, 0.2, 0.3, 0.2])
    low = np.array([0.2, 0.2, 0.2, 0.495, -5.0])
    upp = np.array([0.6, 0.9, 0.7, 0.6, 5.0])
    z = np.zeros(5)
    r = Oracle.mma_ex(1, xval, xmin, xmax, xold1, xold2, z + 0.1, 0.0, z + 0.1, low, upp)
    # iter < 3: xval -/+ 0.5 (xmax - xmin)
    np.testing.assert_allclose(r["low"], [0.4, 0.4, 0.4, 0.4, -0.025], atol=1e-15)
    np.testing.assert_allclose(r["upp"], [0.6, 0.6, 0.6, 0.6, 0.125], atol=1e-15)
    r = Oracle.mma_ex(3, xval, xmin, xmax, xold1, xold2, z + 0.1, 0.0, z + 0.1, low, upp)
    # element 0: (x - x1)(x1 - x2) = 0.1 * 0.1 > 0 -> gamma 1.2:
    #   low = 0.5 - 1.2 (0.4 - 0.2) = 0.26, upp = 0.5 + 1.2 (0.6 - 0.4) = 0.74
    # element 1: (-0.1)(0.3) < 0 -> gamma 0.7:
    #   low = 0.5 - 0.7 (0.6 - 0.2) = 0.22, upp = 0.5 + 0.7 (0.9 - 0.6) = 0.71
    # element 2: x = x1 -> gamma 1: low = 0.2, upp = 0.7
    # element 3: gamma 1.2: low = 0.5 - 1.2 (0.4 - 0.495) = 0.614 -> clamped to
    #   lowmax = 0.5 - 0.01 * 0.2 = 0.498; upp = 0.5 + 1.2 * 0.2 = 0.74
    # element 4: (-0.05)(-0.1) > 0 -> gamma 1.2: low = 0.05 - 1.2 (0.1 + 5) = -6.07
    #   -> clamped to lowmin = 0.05 - 10 * 0.15 = -1.45; upp = 0.05 + 1.2 * 4.9 = 5.93
    #   -> clamped to uppmax = 0.05 + 1.5 = 1.55
    np.testing.assert_allclose(r["low"], [0.26, 0.22, 0.2, 0.498, -1.45], atol=1e-14)
    np.testing.assert_allclose(r["upp"], [0.74, 0.71, 0.7, 0.74, 1.55], atol=1e-14)
    # Update (1867-1868): xold2 <- xold1, xold1 <- xval
    np.testing.assert_array_equal(r["xold2"], xold1)
    np.testing.assert_array_equal(r["xold1"], xval)


def kkt_residual(r, c=1000.0):
    """KKT residual of (x, lambda) for min f0~ + c y s.t. f1~ - y <= b,
    alpha <= x <= beta, y >= 0 (a = 0, d = 0, 1814-1816)."""
    x, lam, U, L = r["x"], r["lam"], r["upp"], r["low"]
    grad = (r["p0"] + lam * r["pij"]) / (U - x) ** 2 - (r["q0"] + lam * r["qij"]) / (x - L) ** 2
    scale = np.abs((r["p0"] + lam * r["pij"]) / (U - x) ** 2).max()
    at_lo = np.isclose(x, r["alpha"], rtol=0, atol=1e-12)
    at_hi = np.isclose(x, r["beta"], rtol=0, atol=1e-12)
    inner = ~(at_lo | at_hi)
    stat = max(np.abs(grad[inner]).max(initial=0.0), np.maximum(0, -grad[at_lo]).max(initial=0.0),
               np.maximum(0, grad[at_hi]).max(initial=0.0)) / scale
    y = max(0.0, lam - c)
    _, f1 = approx_terms(r, x)
    viol = f1 - y - r["b"]
    return stat, max(viol, 0.0), abs(lam * viol), lam


def test_mma_kkt_of_returned_point():
    """SolveDIP (1930-1962) returns a KKT point of the subproblem."""
    rng = np.random.default_rng(31)
    active = 0
    for _ in range(8):
        r = call_mma(random_subproblem(rng, 80))
        stat, viol, comp, lam = kkt_residual(r)
        assert stat < 1e-8, stat
        # SolveDIP's last barrier parameter is the first power of ten below
        # 1e-9 sqrt(m + n) times ten: mu lambda ~ that value, within 0.9 of it
        assert viol < 1e-8 and comp < 2e-8 * math.sqrt(81), (viol, comp)
        assert lam >= 0
        active += lam > 1e-6
    assert active >= 2  # the constraint was binding in some of the problems


def solve_subproblem_by_bisection(r, c=1000.0):
    """Independent solve of the MMA subproblem: for a multiplier lambda each
    x_j minimises its strictly convex term on [alpha_j, beta_j] (bisection on
    the sign of the derivative, not the closed-form stationary point), and
    lambda is the root of the decreasing dual gradient (bisection again),
    with the artificial variable y = max(0, lambda - c)."""
    U, L, a, b = r["upp"], r["low"], r["alpha"], r["beta"]

    def x_of(lam):
        P, Q = r["p0"] + lam * r["pij"], r["q0"] + lam * r["qij"]
        lo, hi = a.copy(), b.copy()
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            d = P / (U - mid) ** 2 - Q / (mid - L) ** 2
            hi = np.where(d > 0, mid, hi)
            lo = np.where(d > 0, lo, mid)
        return 0.5 * (lo + hi)

    def h(lam):
        x = x_of(lam)
        return np.sum(r["pij"] / (U - x) + r["qij"] / (x - L)) - r["b"] - max(0.0, lam - c), x

    h0, x0 = h(0.0)
    if h0 <= 0:
        return x0, 0.0
    lo, hi = 0.0, 1.0
    while h(hi)[0] > 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if h(mid)[0] > 0:
            lo = mid
        else:
            hi = mid
    lam = 0.5 * (lo + hi)
    return x_of(lam), lam


def test_mma_matches_independent_subproblem_solve():
    """The same subproblem solved by nested bisection on the primal-dual
    optimality conditions."""
    rng = np.random.default_rng(37)
    binding = 0
    for _ in range(6):
        p = random_subproblem(rng, 30)
        r = call_mma(p)
        x, lam = solve_subproblem_by_bisection(r)
        np.testing.assert_allclose(r["x"], x, atol=1e-7)
        assert abs(r["lam"] - lam) <= 1e-6 * max(1.0, lam)
        binding += lam > 0
    assert binding >= 2


def test_mma_converges_to_analytic_optimum():
    """Sequential MMA with the LOCAL loop's move limit 0.1 (356, 364-365) on
    min sum w_j / x_j  s.t.  mean(x) <= V: the optimum is x_j proportional
    to sqrt(w_j) (Lagrange), reached to 1e-6."""
    rng = np.random.default_rng(41)
    n, V = 50, 0.4
    w = 0.5 + rng.random(n)
    x_opt = n * V * np.sqrt(w) / np.sqrt(w).sum()
    assert x_opt.min() > 0 and x_opt.max() < 1
    x = np.full(n, V)
    xold1, xold2, low, upp = x.copy(), x.copy(), np.zeros(n), np.zeros(n)
    for it in range(1, 61):
        dfdx = -w / x ** 2
        gx = x.mean() / V - 1.0
        dgdx = np.full(n, 1.0 / (n * V))
        xmin, xmax = np.maximum(0.0, x - 0.1), np.minimum(1.0, x + 0.1)
        r = Oracle.mma(it, x, xmin, xmax, xold1, xold2, dfdx, gx, dgdx, low, upp)
        x, xold1, xold2, low, upp = r["x"], r["xold1"], r["xold2"], r["low"], r["upp"]
    np.testing.assert_allclose(x, x_opt, atol=1e-6)
