"""Parity at the benchmark's sizes, through properties that do not depend on
size: C2 (51 k prisms) against the oracle's sparse step, C3 (1.62 M prisms)
through agreement of the two precisions, the residual at the converged
positions and bitwise reproducibility.  Every test here needs an MI355X."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _solver(nr, w):
    m = w["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"])
    return s


def test_c2_step_matches_oracle_sparse_step(orc, nr):
    """One C2 time step (ellipse 100 x 70 um, 4 layers, viscous, random growth
    field): FP64 device step against the oracle's BSR + BiCGStab step, both
    with a 1e-12 linear solve; positions within 1e-9 relative, the same NR
    iteration count and the same |du| sequence (relative 1e-6 where |du| is
    above the linear-solve noise)."""
    from tissueorigami_amd.workloads import make_workload
    w = make_workload("c2")
    m = w["mesh"]
    nn = len(m["nodes"])
    inv = orc.inv3_many(w["inc"])
    u0 = m["nodes"].reshape(-1).copy()
    s = _solver(nr, w)
    s.set_growth(inv)
    A = s.update_visc_surface(u0, w["place"], w["etype"], w["spans"])
    A_o = orc.visc_surfaces(nn, m["conn"], u0, w["place"], w["etype"], w["spans"])
    assert np.abs(A - A_o).max() <= 1e-12 * np.abs(A_o).max()
    ug, st = s.step(u0, w["dt"], 1e-8, 20, "fp64", 1e-12)
    lam, mu = orc.lame(w["E"], w["nu"])
    P = orc.Problem(m["conn"], m["refpos"], inv, lam, mu, w["eta"], w["ext"], A_o)
    uo, du, kit, conv = P.nr_step_sparse(u0, w["dt"], 1e-8, 20, 1e-12, 50000)
    assert st["converged"] and conv
    assert st["nr_iterations"] == len(du)
    big = du > 1e-6
    assert np.allclose(np.array(st["du_norms"])[big], du[big], rtol=1e-6)
    assert np.abs(ug - uo).max() <= 1e-9 * np.abs(uo).max()
    s.close()


def test_c2_jacobian_equals_oracle_bsr(orc, nr):
    """The assembled C2 Jacobian (570 k blocks, viscous) at a perturbed state,
    exported through tf_nr_get_csr, equals the oracle's BSR assembly entry by
    entry (relative 1e-13 of the largest entry), and so does the residual."""
    from tissueorigami_amd.workloads import make_workload
    import scipy.sparse as sp
    w = make_workload("c2")
    m = w["mesh"]
    nn = len(m["nodes"])
    inv = orc.inv3_many(w["inc"])
    u0 = m["nodes"].reshape(-1).copy()
    uk = u0 + np.random.default_rng(5).normal(0, 0.05, u0.shape)
    A = orc.visc_surfaces(nn, m["conn"], u0, w["place"], w["etype"], w["spans"])
    lam, mu = orc.lame(w["E"], w["nu"])
    P = orc.Problem(m["conn"], m["refpos"], inv, lam, mu, w["eta"], w["ext"], A)
    rowptr, col, val, go = P.assemble_bsr(uk, u0, w["dt"])
    Jo = sp.bsr_matrix((val, col, rowptr), shape=(3 * nn, 3 * nn)).tocsr()
    s = _solver(nr, w)
    s.set_growth(inv)
    s.set_visc_surface(A)
    g, _ = s.assemble(uk, u0, w["dt"])
    ia, ja, a = s.get_csr(threshold=-1.0)      # every structural entry, K orientation
    K = sp.csr_matrix((a, ja, ia), shape=(3 * nn, 3 * nn))
    D = (K.T.tocsr() - Jo).tocoo()
    scale = np.abs(val).max()
    assert (np.abs(D.data).max() if D.nnz else 0.0) <= 1e-13 * scale
    assert np.abs(g - go).max() <= 1e-13 * np.abs(go).max()
    s.close()


def test_c2_circumference_binding_matches_oracle_sparse(orc, nr):
    """C2 with the reference's circumference binding (x and y of every
    circumference node bound to the basal node of its column, as
    Simulation::bindCircumferenceXY builds it): the bound FP64 step (explicit
    N^T J N + I on the tiled solver) against the oracle's sparse bound step,
    and the mixed step within 1e-8 of it."""
    from tissueorigami_amd.workloads import make_workload
    w = make_workload("c2")
    m = w["mesh"]
    nn, n2 = len(m["nodes"]), m["n2d"]
    circ = np.where(m["at_circumference"][:n2] == 1)[0]
    pairs = np.array([[3 * (L * n2 + k) + d, 3 * k + d] for k in circ
                      for L in range(1, m["n_layers"] + 1) for d in (0, 1)], np.int32)
    inv = orc.inv3_many(w["inc"])
    u0 = m["nodes"].reshape(-1).copy()
    A = orc.visc_surfaces(nn, m["conn"], u0, w["place"], w["etype"], w["spans"])
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"], None, pairs)
    s.set_growth(inv)
    s.set_visc_surface(A)
    ug, st = s.step(u0, w["dt"], 1e-8, 20, "fp64", 1e-12)
    um, stm = s.step(u0, w["dt"], 1e-8, 20, "mixed", 1e-3)
    lam, mu = orc.lame(w["E"], w["nu"])
    P = orc.Problem(m["conn"], m["refpos"], inv, lam, mu, w["eta"], w["ext"], A, None, pairs)
    uo, du, kit, conv = P.nr_step_sparse(u0, w["dt"], 1e-8, 20, 1e-12, 50000)
    assert st["converged"] and stm["converged"] and conv
    scale = np.abs(uo).max()
    assert np.abs(ug - uo).max() <= 1e-9 * scale
    assert np.abs(um - uo).max() <= 1e-8 * scale
    d = ug - u0
    assert np.abs(d[pairs[:, 0]] - d[pairs[:, 1]]).max() <= 1e-12
    s.close()


@pytest.fixture(scope="module")
def c3(nr):
    from tissueorigami_amd.workloads import make_workload
    w = make_workload("c3")
    s = _solver(nr, w)
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    yield w, s
    s.close()


def _c3_step(w, s, precision, rtol):
    ne = len(w["mesh"]["conn"])
    s.growth_tensor(np.tile(np.eye(3).reshape(9), (ne, 1)))
    u, st = s.advance(1, w["dt"], u0=w["mesh"]["nodes"].reshape(-1), precision=precision,
                      krylov_rtol=rtol)
    return u, st[0]


def test_c3_mixed_equals_fp64_and_is_reproducible(c3):
    """1.62 M prisms: the mixed step (FP32 tangents and Krylov, inexact
    Newton, FP64 residual) lands within 1e-8 relative of the FP64 step, and
    repeating each gives the same bits."""
    w, s = c3
    um, stm = _c3_step(w, s, "mixed", 1e-3)
    ud, std = _c3_step(w, s, "fp64", 1e-10)
    assert stm["converged"] and std["converged"]
    assert np.abs(um - ud).max() <= 1e-8 * np.abs(ud).max()
    um2, stm2 = _c3_step(w, s, "mixed", 1e-3)
    assert np.array_equal(um, um2)
    assert stm["krylov_iterations"] == stm2["krylov_iterations"]


def test_c3_trajectory_with_solve_policies_matches_fp64(c3):
    """1.62 M prisms, four steps: the mixed step with the warm start and the
    forcing-term safeguard (the bench's configuration) ends within 1e-8
    relative of the FP64 BiCGStab trajectory at 1e-10, and the later steps'
    first solves take fewer CG iterations than the first step's."""
    w, s = c3
    ne = len(w["mesh"]["conn"])
    u0 = w["mesh"]["nodes"].reshape(-1)
    fg0 = np.tile(np.eye(3).reshape(9), (ne, 1))
    s.growth_tensor(fg0)
    um, stm = s.advance(4, w["dt"], u0=u0, precision="mixed", krylov_rtol=1e-3)
    s.growth_tensor(fg0)
    ud, std = s.advance(4, w["dt"], u0=u0, precision="fp64", krylov_rtol=1e-10)
    assert all(st["converged"] for st in stm + std)
    assert np.abs(um - ud).max() <= 1e-8 * np.abs(ud).max()
    first = [st["krylov_iterations"][0] for st in stm]
    assert max(first[2:]) < first[0], first


def test_c3_converged_positions_are_a_root(c3):
    """The FP64 residual assembled at the converged positions of the mixed
    step is below 1e-6 of the residual at the start of the step."""
    w, s = c3
    u0 = w["mesh"]["nodes"].reshape(-1)
    um, st = _c3_step(w, s, "mixed", 1e-3)
    # the growth tensor and the surfaces of that step are the handle's state now
    g0, _ = s.assemble(u0, u0, w["dt"])
    g1, _ = s.assemble(um, u0, w["dt"])
    assert np.linalg.norm(g1) <= 1e-6 * np.linalg.norm(g0)


@pytest.fixture(scope="module")
def c2(nr):
    from tissueorigami_amd.workloads import make_workload
    return make_workload("c2")


def _c2_mixed_step(nr, orc, w, cell=-1.0, graphs=True):
    m = w["mesh"]
    s = _solver(nr, w)
    s.set_growth(orc.inv3_many(w["inc"]))
    u0 = m["nodes"].reshape(-1).copy()
    s.update_visc_surface(u0, w["place"], w["etype"], w["spans"])
    s.set_coarse(cell)
    s.set_graphs(graphs)
    u, st = s.step(u0, w["dt"], 1e-8, 20, "mixed", 1e-3)
    info = s.coarse_info()
    s.close()
    return u, st, info


def test_c2_coarse_correction_cuts_krylov_iterations(orc, nr, c2):
    """C2 mixed step with the two-level CG preconditioner (rigid-body modes of
    column aggregates) against block Jacobi alone: the same root (both within
    1e-8 of the FP64 step), at most 70 % of the Krylov iterations."""
    u1, st1, info1 = _c2_mixed_step(nr, orc, c2)
    u0, st0, info0 = _c2_mixed_step(nr, orc, c2, cell=0.0)
    assert info1[0] > 100 and info1[1] and not info0[1]
    s = _solver(nr, c2)
    s.set_growth(orc.inv3_many(c2["inc"]))
    ustart = c2["mesh"]["nodes"].reshape(-1).copy()
    s.update_visc_surface(ustart, c2["place"], c2["etype"], c2["spans"])
    ud, std = s.step(ustart, c2["dt"], 1e-8, 20, "fp64", 1e-10)
    s.close()
    assert st1["converged"] and st0["converged"] and std["converged"]
    scale = np.abs(ud).max()
    assert np.abs(u1 - ud).max() <= 1e-8 * scale
    assert np.abs(u0 - ud).max() <= 1e-8 * scale
    k1 = st1["krylov_iterations_total"] / st1["nr_iterations"]
    k0 = st0["krylov_iterations_total"] / st0["nr_iterations"]
    assert k1 <= 0.7 * k0, (k1, k0)


def test_c2_coarse_solve_is_reproducible_and_graph_safe(orc, nr, c2):
    """The two-level solve gives the same bits twice, and with the iteration
    chunks replayed through captured graphs or launched directly."""
    ua, sta, _ = _c2_mixed_step(nr, orc, c2)
    ub, stb, _ = _c2_mixed_step(nr, orc, c2)
    uc, stc, _ = _c2_mixed_step(nr, orc, c2, graphs=False)
    assert np.array_equal(ua, ub) and np.array_equal(ua, uc)
    assert sta["krylov_iterations"] == stb["krylov_iterations"] == stc["krylov_iterations"]
