"""Laser ablation on the device (tf_nr_set_ablation) against the oracle's
restatement: an ablated prism contributes no forces, tangents, viscosity
surfaces or growth, and an ablated node whose Jacobian diagonal entry is zero
gets 1 there (checkJacobianForAblatedNodes, NewtonRaphsonSolver.cpp:425-439).
Every test here needs an MI355X."""
import numpy as np
import pytest
import scipy.sparse as sp

import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _case(seed=41, visc=True):
    """A perturbed tissue with a few ablated prisms, and one corner node whose
    prisms are all ablated and which has no drag: its diagonal needs the rule."""
    p = problems.perturbed_problem(seed=seed, visc=visc, fixed=True)
    conn = p["mesh"]["conn"]
    nn = len(p["mesh"]["nodes"])
    node = int(conn[0, 0])
    flags = np.zeros(len(conn), np.uint8)
    flags[np.where((conn == node).any(axis=1))[0]] = 1
    flags[[len(conn) // 2, len(conn) - 2]] = 1
    ext = p["ext_visc"].copy(); ext[node] = 0.0
    area = p["area"].copy(); area[node] = 0.0
    fixed = p["fixed"].copy(); fixed[node] = 0
    p = dict(p, ext_visc=ext, area=area, fixed=fixed)
    return p, flags, np.array([node, 5], np.int32)


def _gpu(nr, p, flags, nodes):
    m = p["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], p["refpos"])
    s.set_material(p["E"], p["nu"], p["eta"], p["ext_visc"], p["fixed"])
    s.set_growth(p["invFg"])
    s.set_visc_surface(p["area"])
    s.set_ablation(flags, nodes)
    return s


def _orc(orc, p):
    lam, mu = orc.lame(p["E"], p["nu"])
    return orc.Problem(p["mesh"]["conn"], p["refpos"], p["invFg"], lam, mu, p["eta"],
                       p["ext_visc"], p["area"], p["fixed"])


@pytest.mark.parametrize("visc", [False, True])
def test_ablated_assembly_matches_oracle(orc, nr, visc):
    p, flags, nodes = _case(visc=visc)
    s = _gpu(nr, p, flags, nodes)
    g, gi = s.assemble(p["uk"], p["un"], p["dt"])
    ia, ja, a = s.get_csr(threshold=-1.0)
    n = 3 * len(p["mesh"]["nodes"])
    K = sp.csr_matrix((a, ja, ia), shape=(n, n)).toarray()
    with orc.ablation(prisms=flags, nodes=nodes):
        Ko, go, gio = _orc(orc, p).assemble_dense(p["uk"], p["un"], p["dt"])
    scale = np.abs(Ko).max()
    assert np.abs(K - Ko).max() <= 1e-13 * scale
    assert np.abs(g - go).max() <= 1e-13 * np.abs(go).max()
    assert np.abs(gi - gio).max() <= 1e-13 * np.abs(gio).max()
    d = 3 * int(nodes[0])
    assert np.array_equal(K[d:d + 3, d:d + 3], np.eye(3))
    # the element outputs of an ablated prism are zero
    ge, gv, Ke, Kv = s.element_outputs()
    e = int(np.where(flags)[0][0])
    assert not ge[e].any() and not gv[e].any() and not Ke[e].any()
    s.close()


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-12), ("mixed", 1e-3)])
def test_ablated_step_matches_oracle(orc, nr, precision, rtol):
    p, flags, nodes = _case(seed=44)
    s = _gpu(nr, p, flags, nodes)
    u, st = s.step(p["un"], p["dt"], 1e-8, 20, precision, rtol)
    with orc.ablation(prisms=flags, nodes=nodes):
        uo, du, conv = _orc(orc, p).nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert st["converged"] and conv
    tol = 1e-9 if precision == "fp64" else 1e-8
    assert np.abs(u - uo).max() <= tol * np.abs(uo).max()
    if precision == "fp64":
        assert st["nr_iterations"] == len(du)
    d = 3 * int(nodes[0])
    assert np.array_equal(u[d:d + 3], p["un"][d:d + 3])
    s.close()


def test_ablated_surfaces_match_oracle(orc, nr):
    from tissueorigami_amd import tissue
    m = tissue.ellipse(10.0, 7.0, 1.0, 5.0, 2)
    ne, nn = len(m["conn"]), len(m["nodes"])
    u = m["nodes"].reshape(-1)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    flags = np.zeros(ne, np.uint8)
    flags[::7] = 1
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_ablation(flags)
    A = s.update_visc_surface(u, place, etype, spans)
    with orc.ablation(prisms=flags):
        Ao = orc.visc_surfaces(nn, m["conn"], u, place, etype, spans)
    assert np.abs(A - Ao).max() <= 1e-12 * np.abs(Ao).max()
    assert not np.array_equal(Ao, orc.visc_surfaces(nn, m["conn"], u, place, etype, spans))
    s.close()


def test_ablated_tissue_in_a_batch_equals_alone(nr):
    p0, f0, n0 = _case(seed=46)
    p1 = problems.perturbed_problem(seed=47, visc=True, fixed=True)
    f1 = np.zeros(len(p1["mesh"]["conn"]), np.uint8)
    b = nr.TissueBatch([(len(p["mesh"]["nodes"]), p["mesh"]["conn"], p["refpos"])
                        for p in (p0, p1)])
    cat = lambda k: np.concatenate([p0[k], p1[k]])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"), cat("fixed"))
    b.set_growth(cat("invFg"))
    b.set_visc_surface(cat("area"))
    b.set_ablation(np.concatenate([f0, f1]), n0)
    b.set_tile_size(2)
    ub, stb = b.step(cat("un"), p0["dt"], 1e-8, 20, "fp64", 1e-10)
    a = _gpu(nr, p0, f0, n0)
    a.set_tile_size(2)
    ua, sta = a.step(p0["un"], p0["dt"], 1e-8, 20, "fp64", 1e-10)
    n = 3 * len(p0["mesh"]["nodes"])
    assert np.array_equal(ub[:n], ua)
    assert stb[0]["krylov_iterations"] == sta["krylov_iterations"]
    a.close()
    b.close()


def test_ablated_prisms_do_not_grow(nr):
    from tissueorigami_amd import tissue
    m = tissue.ellipse(8.0, 6.0, 1.0, 5.0, 2)
    ne, nn = len(m["conn"]), len(m["nodes"])
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    ext = np.zeros((nn, 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(np.full(ne, 1000.0), np.full(ne, 0.3), np.full(ne, 10.0), ext)
    s.set_growth_rates(tissue.rates_xyz(np.full(ne, 0.05)))
    s.set_surface_flags(place, etype, spans)
    flags = np.zeros(ne, np.uint8)
    flags[3] = 1
    s.set_ablation(flags)
    u, st = s.advance(2, 60.0, u0=m["nodes"].reshape(-1), precision="fp64")
    Fg = s.growth_tensor()
    assert np.array_equal(Fg[3], np.eye(3).reshape(9))
    assert Fg[4, 0] > 1.0
    assert all(x["converged"] for x in st)
    s.close()


def test_ablation_rejected_with_packing_and_adjoint(nr):
    p, flags, nodes = _case(seed=48)
    s = _gpu(nr, p, flags, nodes)
    m = p["mesh"]
    with pytest.raises(nr.TFError) as e:
        s.set_packing(m["node_placement"])
    assert e.value.code == nr.TF_ERR_ARG
    ne = len(m["conn"])
    with pytest.raises(nr.TFError) as e:
        s.adj_forward(p["un"], np.tile(p["invFg"][None], (1, 1, 1)), p["dt"],
                      np.zeros(ne, np.int32), np.zeros(ne, np.int32), np.zeros(ne, np.uint8))
    assert e.value.code == nr.TF_ERR_ARG
    s.close()
