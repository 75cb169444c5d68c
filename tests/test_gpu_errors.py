"""Error conventions of the C ABI, forced from the host (no kernel faults or
hangs): every entry returns a status instead of the reference's exit() calls
(NewtonRaphsonSolver.cpp:659-727), a NaN in the residual is reported instead
of printed (NewtonRaphsonSolver.cpp:453-455), and a Newton loop that hits its
cap still writes its last iterate, as the reference accepts it
(Simulation.cpp:6724-6727).  Every test here needs an MI355X."""
import ctypes

import numpy as np
import pytest

import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _solver(nr, p):
    m = p["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], p["refpos"])
    s.set_material(p["E"], p["nu"], p["eta"], p["ext_visc"], p["fixed"])
    s.set_growth(p["invFg"])
    s.set_visc_surface(p["area"])
    return s


def _raw_step(nr, s, u_n, dt, maxit, precision, rtol):
    """tf_nr_step through ctypes, returning (status, u_out, stats per tissue)."""
    L = nr.load()
    u_n = np.ascontiguousarray(u_n, dtype=np.float64).reshape(-1)
    out = np.full_like(u_n, np.nan)
    st = (nr.StepStats * s.n_tissues)()
    code = L.tf_nr_step(s._h, u_n.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(dt),
                        1e-8, int(maxit), nr.PREC[precision], float(rtol),
                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), st)
    return code, out, [st[t].as_dict() for t in range(s.n_tissues)]


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-10), ("mixed", 1e-3)])
def test_nan_in_growth_gives_tf_err_nan(nr, precision, rtol):
    """A NaN in one prism's InvFg makes the residual NaN: TF_ERR_NAN, with the
    NaN residual caught before any fallback solve."""
    p = problems.perturbed_problem(seed=21, visc=True, fixed=True)
    inv = p["invFg"].copy()
    inv[3, 4] = np.nan
    p["invFg"] = inv
    s = _solver(nr, p)
    code, u, st = _raw_step(nr, s, p["un"], p["dt"], 20, precision, rtol)
    assert code == nr.TF_ERR_NAN, (code, nr.load().tf_last_error())
    assert st[0]["status"] == nr.TF_ERR_NAN
    assert st[0]["nr_iterations"] == 1
    assert st[0]["fp64_fallbacks"] == 0 and st[0]["gmres_fallbacks"] == 0
    assert b"NaN" in nr.load().tf_last_error()
    s.close()


def test_nan_in_one_tissue_of_a_batch_leaves_the_others(nr):
    """In a batch only the tissue with the NaN reports TF_ERR_NAN; the other
    tissue converges to the result it has alone, bit for bit."""
    p0 = problems.perturbed_problem(seed=22, visc=True, fixed=True)
    p1 = problems.perturbed_problem(seed=23, visc=True, fixed=True)
    p1 = dict(p1)
    inv = p1["invFg"].copy()
    inv[0, 0] = np.inf
    p1["invFg"] = inv
    b = nr.TissueBatch([(len(p["mesh"]["nodes"]), p["mesh"]["conn"], p["refpos"])
                        for p in (p0, p1)])
    cat = lambda k: np.concatenate([p[k] for p in (p0, p1)])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"), cat("fixed"))
    b.set_growth(cat("invFg"))
    b.set_visc_surface(cat("area"))
    b.set_tile_size(2)
    un = cat("un")
    code, u, st = _raw_step(nr, b, un, p0["dt"], 20, "fp64", 1e-10)
    assert code == nr.TF_ERR_NAN
    assert st[0]["status"] == nr.TF_OK and st[0]["converged"] == 1
    assert st[1]["status"] == nr.TF_ERR_NAN
    a = _solver(nr, p0)
    a.set_tile_size(2)
    u0, st0 = a.step(p0["un"], p0["dt"], 1e-8, 20, "fp64", 1e-10)
    n0 = 3 * len(p0["mesh"]["nodes"])
    assert np.array_equal(u[:n0], u0)
    assert st0["krylov_iterations"] == st[0]["krylov_iterations"]
    a.close()
    b.close()


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-10), ("mixed", 1e-3)])
def test_newton_cap_writes_the_iterate(orc, nr, precision, rtol):
    """maxit = 1: two NR iterations run (the reference's cap test is
    iteratorK > maxIteration after the increment), the status is
    TF_ERR_NOT_CONVERGED and u_out holds the second iterate.  In FP64 it equals
    the oracle's two-iteration iterate.  The mixed step solves each Newton
    system only to its forcing term, so its iterates take another path far
    from the root; its iterate is checked for being written, reproducible and
    a reduction of the FP64 residual."""
    p = problems.perturbed_problem(seed=24, visc=True, fixed=True)
    s = _solver(nr, p)
    code, u, st = _raw_step(nr, s, p["un"], p["dt"], 1, precision, rtol)
    assert code == nr.TF_ERR_NOT_CONVERGED, (code, nr.load().tf_last_error())
    assert st[0]["nr_iterations"] == 2 and st[0]["converged"] == 0
    assert st[0]["du_norms"][-1] > 1e-8
    assert np.all(np.isfinite(u)) and not np.array_equal(u, p["un"])
    if precision == "fp64":
        lam, mu = orc.lame(p["E"], p["nu"])
        P = orc.Problem(p["mesh"]["conn"], p["refpos"], p["invFg"], lam, mu, p["eta"],
                        p["ext_visc"], p["area"], p["fixed"])
        ref, du, conv = P.nr_step_dense(p["un"], p["dt"], 1e-8, 1)
        assert not conv and len(du) == 2
        assert du[-1] > 1e-8  # the cap, not convergence, ended the oracle's loop too
        assert np.abs(u - ref).max() <= 1e-9 * np.abs(ref).max()
    else:
        code2, u2, _ = _raw_step(nr, s, p["un"], p["dt"], 1, precision, rtol)
        assert code2 == nr.TF_ERR_NOT_CONVERGED and np.array_equal(u, u2)
        g0, _ = s.assemble(p["un"], p["un"], p["dt"])
        g2, _ = s.assemble(u, p["un"], p["dt"])
        assert np.linalg.norm(g2) < np.linalg.norm(g0)
    s.close()


def test_solve_iteration_cap_gives_tf_err_krylov(nr):
    """tf_nr_solve with maxit = 1 cannot reach 1e-10: TF_ERR_KRYLOV, the
    iterate still written."""
    p = problems.perturbed_problem(seed=25, visc=True, fixed=True)
    s = _solver(nr, p)
    g, _ = s.assemble(p["uk"], p["un"], p["dt"])
    L = nr.load()
    x = np.full_like(g, np.nan)
    it = np.zeros(1, np.int32)
    code = L.tf_nr_solve(s._h, g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                         x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1e-10, 1,
                         it.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert code == nr.TF_ERR_KRYLOV
    assert b"iteration cap" in L.tf_last_error()
    assert it[0] == 1 and np.all(np.isfinite(x))
    # the same solve with room converges
    x2, it2 = s.solve(g, 1e-10, 20000)
    assert it2 > 1
    s.close()


def test_invalid_arguments_give_tf_err_arg(nr):
    p = problems.perturbed_problem(seed=26)
    s = _solver(nr, p)
    L = nr.load()
    un = np.ascontiguousarray(p["un"])
    out = np.empty_like(un)
    st = (nr.StepStats * 1)()
    dptr = ctypes.POINTER(ctypes.c_double)
    # maxit beyond the statistics capacity, an unknown precision, dt <= 0
    assert L.tf_nr_step(s._h, un.ctypes.data_as(dptr), 1.0, 1e-8, 64, 0, 1e-10,
                        out.ctypes.data_as(dptr), st) == nr.TF_ERR_ARG
    assert L.tf_nr_step(s._h, un.ctypes.data_as(dptr), 1.0, 1e-8, 20, 7, 1e-10,
                        out.ctypes.data_as(dptr), st) == nr.TF_ERR_ARG
    assert L.tf_nr_step(s._h, un.ctypes.data_as(dptr), 0.0, 1e-8, 20, 0, 1e-10,
                        out.ctypes.data_as(dptr), st) == nr.TF_ERR_ARG
    s.close()
