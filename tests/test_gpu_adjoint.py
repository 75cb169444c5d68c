"""Adjoint gradient of a shape loss with respect to growth-field coefficients
(64 Gaussian bumps on the C1 mesh, 20 steps), checked against central finite
differences of the forward solve."""
import numpy as np
import pytest

import problems

pytestmark = pytest.mark.gpu


def _loss(c, u):
    x = u.reshape(-1, 3)[c["apical"]]
    return float(((x - c["target"]) ** 2).sum())


def _forward(orc, s, c, coef, tol=1e-12, rtol=1e-13):
    rate = c["phi"] @ coef
    inv, dinv = problems.growth_steps(orc, rate, c["dt"], c["n_steps"])
    u0 = c["mesh"]["nodes"].reshape(-1)
    u, st = s.adj_forward(u0, inv, c["dt"], c["place"], c["etype"], c["spans"], tol=tol,
                          maxit=25, krylov_rtol=rtol)
    return u, inv, dinv, st


@pytest.fixture(scope="module")
def setup(orc):
    from tissueorigami_amd import nr
    c = problems.c5_config()
    m = c["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    yield c, s
    s.close()


def test_adjoint_gradient_matches_central_differences(orc, setup):
    c, s = setup
    u, inv, dinv, stf = _forward(orc, s, c, c["coef"])
    dL_du = np.zeros_like(u)
    x = u.reshape(-1, 3)
    dL = dL_du.reshape(-1, 3)
    dL[c["apical"]] = 2 * (x[c["apical"]] - c["target"])
    dL_dr, stb = s.adj_backward(dinv, dL_du, krylov_rtol=1e-13)
    grad = (c["phi"].T @ dL_dr) / 3600.0          # rate per hour -> per second
    h = 1e-6
    idx = np.random.default_rng(1).choice(64, 8, replace=False)
    for k in idx:
        cp = c["coef"].copy(); cm = c["coef"].copy()
        cp[k] += h; cm[k] -= h
        lp = _loss(c, _forward(orc, s, c, cp)[0])
        lm = _loss(c, _forward(orc, s, c, cm)[0])
        fd = (lp - lm) / (2 * h)
        assert abs(grad[k] - fd) <= 1e-5 * abs(fd), (k, grad[k], fd)


def test_adjoint_matches_oracle_forward(orc, setup):
    """The adjoint's forward checkpoints are the plain NR step: compare the
    final positions with the oracle's dense step sequence."""
    c, s = setup
    u, inv, dinv, st = _forward(orc, s, c, c["coef"], tol=1e-8, rtol=1e-12)
    m = c["mesh"]
    lam, mu = orc.lame(c["E"], c["nu"])
    uo = m["nodes"].reshape(-1).copy()
    for k in range(c["n_steps"]):
        A = orc.visc_surfaces(len(m["nodes"]), m["conn"], uo, c["place"], c["etype"], c["spans"])
        P = orc.Problem(m["conn"], m["refpos"], inv[k], lam, mu, c["eta"], c["ext_visc"], A)
        uo, du, conv = P.nr_step_dense(uo, c["dt"], 1e-8, 20)
    assert np.abs(u - uo).max() <= 1e-9 * np.abs(uo).max()


def test_adjoint_cost_is_within_three_forwards(orc, setup):
    c, s = setup
    u, inv, dinv, stf = _forward(orc, s, c, c["coef"], tol=1e-8, rtol=1e-10)
    dL_du = np.ones_like(u)
    _, stb = s.adj_backward(dinv, dL_du, krylov_rtol=1e-10)
    # device time of the backward sweep against the forward solve
    assert stb["ms_total"] <= 3.0 * stf["ms_total"], (stb["ms_total"], stf["ms_total"])
