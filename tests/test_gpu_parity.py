"""Parity of the HIP path (through the C ABI) with the CPU oracle and with the
reference's own Pardiso fixtures.  Every test here needs an MI355X."""
import numpy as np
import pytest

import problems
import replay

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _gpu_solver(nr, p, pairs=None):
    m = p["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], p["refpos"])
    s.set_material(p["E"], p["nu"], p["eta"], p["ext_visc"], p["fixed"], pairs)
    s.set_growth(p["invFg"])
    s.set_visc_surface(p["area"])
    return s


def _oracle_problem(orc, p, pairs=None):
    lam, mu = orc.lame(p["E"], p["nu"])
    return orc.Problem(p["mesh"]["conn"], p["refpos"], p["invFg"], lam, mu, p["eta"],
                       p["ext_visc"], p["area"], p["fixed"], pairs)


@pytest.mark.parametrize("visc", [False, True])
def test_element_kernel_matches_oracle(orc, nr, visc):
    p = problems.perturbed_problem(seed=1, visc=visc)
    s = _gpu_solver(nr, p)
    s.assemble(p["uk"], p["un"], p["dt"])
    ge, gv, Ke, Kv = s.element_outputs()
    m = p["mesh"]
    lam, mu = orc.lame(p["E"], p["nu"])
    pos = p["uk"].reshape(-1, 3)[m["conn"]].reshape(-1, 18)
    vel = ((p["uk"] - p["un"]) * (1.0 / p["dt"])).reshape(-1, 3)[m["conn"]].reshape(-1, 18)
    oge, ogv, oKe, oKv = orc.elements(pos, p["refpos"], p["invFg"], lam, mu, p["eta"], vel,
                                      p["dt"], 1)
    for a, b in ((ge, oge), (gv, ogv), (Ke, oKe), (Kv, oKv)):
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 1e-13 * scale
    # the FP64 kernel follows the oracle's operation order: almost every value
    # is bitwise identical (only the device log() may differ in the last bit)
    assert np.mean(Ke == oKe) > 0.5
    s.close()


@pytest.mark.parametrize("visc", [False, True])
def test_mixed_element_tangents_match_oracle(orc, nr, visc):
    """The FP32 tangents of the mixed path (closed forms, FP32 arithmetic from
    FP64 kinematics) equal the oracle's FP64 tangents to FP32 rounding, and
    the forces stay FP64."""
    p = problems.perturbed_problem(seed=1, visc=visc)
    s = _gpu_solver(nr, p)
    s.assemble(p["uk"], p["un"], p["dt"], precision="mixed")
    ge, gv, Ke, Kv = s.element_outputs()
    m = p["mesh"]
    lam, mu = orc.lame(p["E"], p["nu"])
    pos = p["uk"].reshape(-1, 3)[m["conn"]].reshape(-1, 18)
    vel = ((p["uk"] - p["un"]) * (1.0 / p["dt"])).reshape(-1, 3)[m["conn"]].reshape(-1, 18)
    oge, ogv, oKe, oKv = orc.elements(pos, p["refpos"], p["invFg"], lam, mu, p["eta"], vel,
                                      p["dt"], 1)
    for a, b in ((ge, oge), (gv, ogv)):
        assert np.abs(a - b).max() <= 1e-13 * max(np.abs(b).max(), 1e-300)
    for a, b in ((Ke, oKe), (Kv, oKv)):
        scale = max(np.abs(b).max(), 1e-300)
        assert np.abs(a - b).max() <= 2e-6 * scale
    s.close()


@pytest.mark.parametrize("visc,fixed", [(False, False), (True, False), (True, True)])
def test_assembled_jacobian_and_residual(orc, nr, visc, fixed):
    p = problems.perturbed_problem(seed=2, visc=visc, fixed=fixed)
    s = _gpu_solver(nr, p)
    g, gi = s.assemble(p["uk"], p["un"], p["dt"])
    K, og, ogi = _oracle_problem(orc, p).assemble_dense(p["uk"], p["un"], p["dt"])
    n = K.shape[0]
    ia, ja, a = s.get_csr(threshold=-1.0)  # every structural entry
    G = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(ia))
    G[rows, ja] = a
    scale = np.abs(K).max()
    assert np.abs(G - K).max() <= 1e-13 * scale
    assert np.abs(g - og).max() <= 1e-13 * np.abs(og).max()
    # structural pattern is a superset of the reference's thresholded one
    sia, sja = s.csr_pattern()
    S = np.zeros((n, n), bool)
    S[np.repeat(np.arange(n), np.diff(sia)), sja] = True
    assert np.all(S[np.abs(K) > 1e-13])
    # thresholded pattern equals constructiaForPardiso on the oracle's dense K
    band = (np.abs(K) > 1e-16) & (np.abs(K) < 1e-10)
    assert not band.any(), "test problem has entries near the threshold"
    tia, tja, ta = s.get_csr(threshold=1e-13)
    oia, oja, oa = orc.csr_threshold(K)
    assert np.array_equal(tia, oia) and np.array_equal(tja, oja)
    assert np.abs(ta - oa).max() <= 1e-13 * scale
    s.close()


@pytest.mark.parametrize("visc,fixed", [(False, False), (True, False), (True, True)])
def test_mixed_operator_matches_oracle(orc, nr, visc, fixed):
    """The FP32 operator of the mixed assembly equals the oracle's FP64 K to
    FP32 rounding.  The element kernel sums each prism's elastic block and the
    viscous block that the reference's transposed storage places at the same
    K block; a misplaced viscous block would be off by its asymmetry (5e-5 of
    the largest entry here).  The residual stays FP64."""
    p = problems.perturbed_problem(seed=2, visc=visc, fixed=fixed)
    s = _gpu_solver(nr, p)
    g, gi = s.assemble(p["uk"], p["un"], p["dt"], precision="mixed")
    K, og, ogi = _oracle_problem(orc, p).assemble_dense(p["uk"], p["un"], p["dt"])
    n = K.shape[0]
    ia, ja, a = s.get_csr(threshold=-1.0)
    G = np.zeros((n, n))
    G[np.repeat(np.arange(n), np.diff(ia)), ja] = a
    scale = np.abs(K).max()
    assert np.abs(G - K).max() <= 4e-6 * scale
    assert np.abs(g - og).max() <= 1e-13 * np.abs(og).max()
    s.close()


def test_assembly_is_bitwise_deterministic(nr):
    p = problems.perturbed_problem(seed=3, visc=True)
    s = _gpu_solver(nr, p)
    g1, _ = s.assemble(p["uk"], p["un"], p["dt"])
    _, _, a1 = s.get_csr(-1.0)
    g2, _ = s.assemble(p["uk"], p["un"], p["dt"])
    _, _, a2 = s.get_csr(-1.0)
    assert np.array_equal(g1, g2) and np.array_equal(a1, a2)
    s.close()


@pytest.mark.parametrize("visc", [False, True])
def test_krylov_solve_matches_dense_lu(orc, nr, visc):
    p = problems.perturbed_problem(seed=4, visc=visc, fixed=True)
    s = _gpu_solver(nr, p)
    g, _ = s.assemble(p["uk"], p["un"], p["dt"])
    K, og, _ = _oracle_problem(orc, p).assemble_dense(p["uk"], p["un"], p["dt"])
    x, it = s.solve(g, rtol=1e-12, maxit=5000)
    xd = orc.dense_solve_transposed(K, og)
    J = K.T
    assert np.linalg.norm(J @ x - og) <= 1e-11 * np.linalg.norm(og)
    assert np.abs(x - xd).max() <= 1e-9 * np.abs(xd).max()
    assert it > 0
    s.close()


def test_binding_step_matches_oracle(orc, nr):
    p = problems.perturbed_problem(seed=5, visc=True)
    nn = len(p["mesh"]["nodes"])
    # bind x and y of a few nodes to other nodes' DOFs (slave, master)
    pairs = np.array([[3 * 7 + 0, 3 * 2 + 0], [3 * 7 + 1, 3 * 2 + 1], [3 * 11 + 2, 3 * 3 + 2],
                      [3 * (nn - 2) + 0, 3 * 4 + 0]], np.int32)
    s = _gpu_solver(nr, p, pairs)
    op = _oracle_problem(orc, p, pairs)
    g, _ = s.assemble(p["uk"], p["un"], p["dt"])
    K, og, _ = op.assemble_dense(p["uk"], p["un"], p["dt"])
    assert np.abs(g - og).max() <= 1e-13 * np.abs(og).max()
    ia, ja, a = s.get_csr(threshold=-1.0)
    n = K.shape[0]
    G = np.zeros((n, n))
    G[np.repeat(np.arange(n), np.diff(ia)), ja] = a
    assert np.abs(G - K).max() <= 1e-13 * np.abs(K).max()
    u_gpu, st = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-12)
    u_orc, du, conv = op.nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert st["converged"] and conv
    assert np.abs(u_gpu - u_orc).max() <= 1e-9 * np.abs(u_orc).max()
    for sl, ms in pairs:
        assert u_gpu[sl] - p["un"][sl] == pytest.approx(u_gpu[ms] - p["un"][ms], abs=1e-12)
    s.close()


@pytest.mark.parametrize("precision,method,rtol", [("fp64", "bicgstab", 1e-12),
                                                   ("mixed", "cg", 1e-3),
                                                   ("mixed", "bicgstab", 1e-5)])
def test_bound_system_on_tiled_solvers_matches_oracle(orc, nr, precision, method, rtol):
    """Binding as the explicit operator N^T J N + I_slave: the tiled solvers
    (FP64, and mixed precision with CG or BiCGStab) reach the oracle's dense
    bound step, and slaves move with their masters."""
    p = problems.perturbed_problem(seed=5, visc=True)
    nn = len(p["mesh"]["nodes"])
    pairs = np.array([[3 * 7 + 0, 3 * 2 + 0], [3 * 7 + 1, 3 * 2 + 1], [3 * 11 + 2, 3 * 3 + 2],
                      [3 * (nn - 2) + 0, 3 * 4 + 0], [3 * 9 + 0, 3 * 2 + 0]], np.int32)
    s = _gpu_solver(nr, p, pairs)
    s.set_krylov(precision, method)
    u_gpu, st = s.step(p["un"], p["dt"], 1e-8, 20, precision, rtol)
    u_orc, du, conv = _oracle_problem(orc, p, pairs).nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert st["converged"] and conv
    tol = 1e-9 if precision == "fp64" else 1e-8
    assert np.abs(u_gpu - u_orc).max() <= tol * np.abs(u_orc).max()
    for sl, ms in pairs:
        assert u_gpu[sl] - p["un"][sl] == pytest.approx(u_gpu[ms] - p["un"][ms], abs=1e-12)
    s.close()


def test_c1_rectangle_50_steps_match_oracle(orc, nr):
    """Plain FP64 small rectangle, 50 steps of 10 s against the oracle's
    dense-LU (Pardiso stand-in) path: node positions within 1e-9 relative."""
    c = problems.c1_config()
    m = c["mesh"]
    ne = len(m["conn"])
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    lam, mu = orc.lame(c["E"], c["nu"])
    u_g = m["nodes"].reshape(-1).copy()
    u_o = u_g.copy()
    Fg = np.tile(np.eye(3), (ne, 1, 1))
    from tissueorigami_amd import tissue
    inc = tissue.growth_increment(c["rate"], c["dt"])
    worst = 0.0
    for step in range(c["n_steps"]):
        Fg = inc @ Fg
        invFg = orc.inv3_many(Fg)
        s.set_growth(invFg)
        A_g = s.update_visc_surface(u_g, c["place"], c["etype"], c["spans"])
        A_o = orc.visc_surfaces(len(m["nodes"]), m["conn"], u_o, c["place"], c["etype"],
                                c["spans"])
        assert np.abs(A_g - A_o).max() <= 1e-14 * np.abs(A_o).max()
        u_g, st = s.step(u_g, c["dt"], 1e-8, 20, "fp64", 1e-12)
        P = orc.Problem(m["conn"], m["refpos"], invFg, lam, mu, c["eta"], c["ext_visc"], A_o)
        u_o, du, conv = P.nr_step_dense(u_o, c["dt"], 1e-8, 20)
        assert st["converged"] and conv
        assert st["nr_iterations"] == len(du)
        worst = max(worst, np.abs(u_g - u_o).max() / np.abs(u_o).max())
    assert worst <= 1e-9, worst
    s.close()


REPLAY = [("run07007", k) for k in range(10)] + [("run07008", k) for k in range(-1, 10)]
# the oracle's worst replay errors with a 1.5x margin (positions) and the
# propagated Save_Force rounding, as in test_oracle_fixtures.py
POS_TOL = {"run07007": 1.5e-9, "run07008": 7e-8}
FORCE_TOL = {"run07007": 2e-5, "run07008": 5e-3}


@pytest.mark.parametrize("run,k", REPLAY)
def test_pardiso_run_replay_on_gpu(orc, nr, run, k):
    """Each step of the reference's Pardiso runs run07007 and run07008 (y fixed
    on the symmetry line, circumference x/y bound to the basal node) replayed
    in FP64 on the device: positions, |du| log and Save_Force."""
    d = replay.load(run)
    s_in = replay.step_inputs(d, k, orc)
    nn = d["frames"].shape[1]
    s = nr.NewtonRaphsonSolver(nn, s_in["conn"], s_in["refpos"])
    s.set_material(s_in["E"], s_in["nu"], s_in["eta"], s_in["ext_visc"], s_in["fixed"],
                   s_in["pairs"])
    s.set_growth(s_in["invFg"])
    s.set_visc_surface(s_in["visc_surf"])
    u, st = s.step(s_in["un"], s_in["dt"], 1e-8, 20, "fp64", 1e-13)
    assert st["converged"]
    assert np.abs(u - s_in["expected"]).max() < POS_TOL[run]
    du = np.array(st["du_norms"])
    assert len(du) == len(s_in["du_log"])
    big = s_in["du_log"] > 1e-10
    assert np.allclose(du[big], s_in["du_log"][big], rtol=1e-5, atol=0)
    if k >= 0:
        g, gi = s.assemble(s_in["expected"], s_in["un"], s_in["dt"])
        assert np.abs(gi - s_in["forces"]).max() < FORCE_TOL[run]
    s.close()


@pytest.mark.parametrize("k", [-1, 4, 9])
@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_run07008_mixed_precision_replay(orc, nr, k, method):
    """run07008 in mixed precision (FP32 tangents and bound Krylov operator,
    forcing term 1e-3): the step lands on the reference's frame as the FP64
    step does."""
    d = replay.load("run07008")
    s_in = replay.step_inputs(d, k, orc)
    s = nr.NewtonRaphsonSolver(d["frames"].shape[1], s_in["conn"], s_in["refpos"])
    s.set_material(s_in["E"], s_in["nu"], s_in["eta"], s_in["ext_visc"], s_in["fixed"],
                   s_in["pairs"])
    s.set_growth(s_in["invFg"])
    s.set_visc_surface(s_in["visc_surf"])
    s.set_krylov("mixed", method)
    u, st = s.step(s_in["un"], s_in["dt"], 1e-8, 20, "mixed", 1e-3)
    assert st["converged"]
    assert np.abs(u - s_in["expected"]).max() < POS_TOL["run07008"]
    dd = u - s_in["un"]
    p = s_in["pairs"]
    assert np.abs(dd[p[:, 0]] - dd[p[:, 1]]).max() <= 1e-12
    assert np.all(u[s_in["fixed"] == 1] == s_in["un"][s_in["fixed"] == 1])
    s.close()


def test_step_is_bitwise_reproducible(nr):
    p = problems.perturbed_problem(seed=6, visc=True)
    s = _gpu_solver(nr, p)
    u1, st1 = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-10)
    u2, st2 = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-10)
    assert np.array_equal(u1, u2)
    assert st1["krylov_iterations"] == st2["krylov_iterations"]
    s.close()


@pytest.mark.parametrize("krylov_rtol", [1e-5, 1e-10])
def test_mixed_precision_step_matches_fp64(orc, nr, krylov_rtol):
    """FP32 tangents / SpMV / block Jacobi with the FP64 residual converge to
    the FP64 root: positions within 1e-8 relative of the FP64 step and of the
    oracle's dense step."""
    p = problems.perturbed_problem(seed=9, visc=True, fixed=True, lx=6.0, ly=5.0, layers=3)
    s = _gpu_solver(nr, p)
    u64, st64 = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-10)
    u32, st32 = s.step(p["un"], p["dt"], 1e-8, 20, "mixed", krylov_rtol)
    uo, du, conv = _oracle_problem(orc, p).nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert st64["converged"] and st32["converged"] and conv
    scale = np.abs(uo).max()
    assert np.abs(u32 - u64).max() <= 1e-8 * scale
    assert np.abs(u32 - uo).max() <= 1e-8 * scale
    s.close()


def test_mixed_c1_rectangle_matches_oracle(orc, nr):
    c = problems.c1_config()
    m = c["mesh"]
    ne = len(m["conn"])
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    lam, mu = orc.lame(c["E"], c["nu"])
    from tissueorigami_amd import tissue
    inc = tissue.growth_increment(c["rate"], c["dt"])
    u_g = m["nodes"].reshape(-1).copy()
    u_o = u_g.copy()
    Fg = np.tile(np.eye(3), (ne, 1, 1))
    for step in range(10):
        Fg = inc @ Fg
        invFg = orc.inv3_many(Fg)
        s.set_growth(invFg)
        s.update_visc_surface(u_g, c["place"], c["etype"], c["spans"])
        A_o = orc.visc_surfaces(len(m["nodes"]), m["conn"], u_o, c["place"], c["etype"],
                                c["spans"])
        u_g, st = s.step(u_g, c["dt"], 1e-8, 20, "mixed", 1e-5)
        P = orc.Problem(m["conn"], m["refpos"], invFg, lam, mu, c["eta"], c["ext_visc"], A_o)
        u_o, du, conv = P.nr_step_dense(u_o, c["dt"], 1e-8, 20)
        assert st["converged"] and conv
        assert np.abs(u_g - u_o).max() <= 1e-8 * np.abs(u_o).max()
    s.close()


def test_device_resident_advance_matches_per_step_path(orc, nr):
    """tf_nr_advance (growth and viscosity surfaces on the device) reproduces
    the host-driven sequence set_growth / update_visc_surface / step."""
    c = problems.c1_config()
    m = c["mesh"]
    ne = len(m["conn"])
    from tissueorigami_amd import tissue
    rates = np.zeros((ne, 3))
    rates[:, 0] = rates[:, 1] = c["rate"] / 3600.0
    a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    a.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    a.set_growth_rates(rates)
    a.set_surface_flags(c["place"], c["etype"], c["spans"])
    u0 = m["nodes"].reshape(-1).copy()
    ua, stats = a.advance(12, c["dt"], u0=u0, krylov_rtol=1e-12)
    Fg_dev = a.growth_tensor()
    b = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    b.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    inc = tissue.growth_increment(c["rate"], c["dt"])
    Fg = np.tile(np.eye(3), (ne, 1, 1))
    ub = u0.copy()
    for k in range(12):
        Fg = inc @ Fg
        b.set_growth(orc.inv3_many(Fg))
        b.update_visc_surface(ub, c["place"], c["etype"], c["spans"])
        ub, st = b.step(ub, c["dt"], 1e-8, 20, "fp64", 1e-12)
        assert st["nr_iterations"] == stats[k]["nr_iterations"]
    assert np.abs(Fg_dev - Fg.reshape(ne, 9)).max() <= 1e-15
    assert np.abs(ua - ub).max() <= 1e-12 * np.abs(ub).max()
    a.close(); b.close()
