"""Pins the FP64 exact-cone oracle (oracle/exact_cone.cpp), the exact
Coulomb baseline the temporalCone solve is measured against (CPU only):

  * the reference's known answer (apps/reference_conic_probe.cpp:9-100): one
    contact, nv = 3, mu = 0.7, solved in two tangent bases 0.731 rad apart;
    optimality residual <= 1e-11, cone violation <= 1e-12, the post-impulse
    velocity independent of the tangent basis to 1e-9, and a warm start from
    the solution returning the same velocity (1e-10) in no more iterations;
  * closed forms: frictionless contact (lambda_n = max(0, -l_n / H_nn)), a
    sticking contact whose unconstrained minimiser lies inside the cone
    (lambda = -H^-1 l), a separating contact (lambda = 0);
  * brute force: small random elliptic-cone programs solved by SciPy's SLSQP
    with the cone written as inequality constraints.
"""

import numpy as np
import pytest

import oracle_lib as ol

M_KAT = np.array([[1.3, 0.2, 0.1], [0.2, 0.8, 0.05], [0.1, 0.05, 1.7]])
V_KAT = np.array([2.0, -1.0, 0.5])


def kat_problem(angle):
    c, s = np.cos(angle), np.sin(angle)
    J = np.array([[0.0, 1.0, 0.0], [c, 0.0, s], [-s, 0.0, c]])
    return J, np.zeros(3), np.full(3, 1.0e-5), np.array([[0.7, 0.7]])


def test_reference_known_answer_basis_invariance_and_warm_start():
    J0, t0, r0, mu0 = kat_problem(0.0)
    J1, t1, r1, mu1 = kat_problem(0.731)
    c0, l0, v0, s0 = ol.exact_cone(M_KAT, V_KAT, J0, t0, r0, mu0)
    c1, l1, v1, s1 = ol.exact_cone(M_KAT, V_KAT, J1, t1, r1, mu1)
    assert c0 == 0 and c1 == 0
    assert s0["residual"] <= 1.0e-11 and s1["residual"] <= 1.0e-11
    assert s0["violation"] <= 1.0e-12 and s1["violation"] <= 1.0e-12
    assert np.linalg.norm(v0 - v1) <= 1.0e-9
    cw, lw, vw, sw = ol.exact_cone(M_KAT, V_KAT, J0, t0, r0, mu0, warm=l0)
    assert cw == 0 and sw["used_warm"]
    assert np.linalg.norm(vw - v0) <= 1.0e-10
    assert sw["iterations"] <= s0["iterations"]
    # this problem slides: the tangential impulse sits on the cone boundary
    assert l0[0] > 0.0 and abs(np.hypot(l0[1], l0[2]) - 0.7 * l0[0]) <= 1.0e-9


def hessian(M, J, reg):
    return J @ M @ J.T + np.diag(reg)


def test_frictionless_contact_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(20):
        A = rng.normal(size=(4, 4))
        M = A @ A.T + 0.5 * np.eye(4)
        J = rng.normal(size=(3, 4))
        vf = rng.normal(size=4)
        reg = np.full(3, 1.0e-6)
        code, lam, vel, st = ol.exact_cone(M, vf, J, np.zeros(3), reg, np.array([[0.0, 0.0]]))
        assert code == 0
        H = hessian(M, J, reg)
        ln = J[0] @ vf
        np.testing.assert_allclose(lam, [max(0.0, -ln / H[0, 0]), 0.0, 0.0], atol=1.0e-9, rtol=1.0e-8)
        np.testing.assert_allclose(vel, vf + M @ J.T @ lam, atol=1.0e-12)


def test_sticking_contact_is_the_unconstrained_minimiser():
    J = np.eye(3)
    M = np.diag([1.0, 2.0, 3.0])
    vf = np.array([-1.0, 0.05, -0.02])  # approaching, small slip: stays inside a mu = 1 cone
    reg = np.full(3, 1.0e-6)
    code, lam, vel, st = ol.exact_cone(M, vf, J, np.zeros(3), reg, np.array([[1.0, 1.0]]))
    assert code == 0
    H = hessian(M, J, reg)
    expect = -np.linalg.solve(H, J @ vf)
    assert np.hypot(expect[1], expect[2]) < expect[0]
    np.testing.assert_allclose(lam, expect, rtol=1.0e-8, atol=1.0e-10)
    np.testing.assert_allclose(vel, 0.0, atol=1.0e-5)


def test_separating_contact_has_no_impulse():
    code, lam, vel, st = ol.exact_cone(np.eye(3), np.array([0.5, 0.0, 0.0]), np.eye(3), np.zeros(3),
                                       np.full(3, 1.0e-6), np.array([[0.8, 0.8]]))
    assert code == 0 and np.all(lam == 0.0)


@pytest.mark.parametrize("seed", range(6))
def test_matches_brute_force_qp(seed):
    from scipy.optimize import minimize

    rng = np.random.default_rng(100 + seed)
    nv, nc = 5, 1 + seed % 2
    A = rng.normal(size=(nv, nv))
    M = A @ A.T + 0.3 * np.eye(nv)
    J = rng.normal(size=(3 * nc, nv))
    vf = rng.normal(size=nv)
    tg = rng.normal(scale=0.1, size=3 * nc)
    reg = np.full(3 * nc, 1.0e-4)
    mu = rng.uniform(0.3, 1.2, size=(nc, 2))
    code, lam, vel, st = ol.exact_cone(M, vf, J, tg, reg, mu)
    assert code == 0
    H = hessian(M, J, reg)
    lin = J @ vf - tg
    f = lambda x: 0.5 * x @ H @ x + lin @ x  # noqa: E731
    cons = []
    for c in range(nc):
        cons.append({"type": "ineq", "fun": lambda x, c=c: x[3 * c]})
        cons.append({"type": "ineq", "fun": lambda x, c=c: x[3 * c] ** 2 - (x[3 * c + 1] / mu[c, 0]) ** 2
                     - (x[3 * c + 2] / mu[c, 1]) ** 2})
    best = None
    for start in (np.zeros(3 * nc), lam + 1.0e-3):
        res = minimize(f, start, jac=lambda x: H @ x + lin, constraints=cons, method="SLSQP",
                       options={"ftol": 1e-14, "maxiter": 2000})
        if best is None or res.fun < best.fun:
            best = res
    # SLSQP may end marginally outside the cone (lower objective by its
    # constraint slack), so the oracle's feasible optimum must be within 1e-6
    assert f(lam) <= best.fun + 1.0e-6 * (1.0 + abs(best.fun))
    assert st["residual"] <= 1.0e-11
    for c in range(nc):
        n_, u_, v_ = lam[3 * c:3 * c + 3]
        assert n_ >= -1.0e-12
        assert np.hypot(u_ / mu[c, 0], v_ / mu[c, 1]) <= n_ * (1.0 + 1.0e-9) + 1.0e-12
