"""Batched tissues (tf_batch_create): one launch per solver stage for many
independent meshes of uneven size.  Each tissue keeps its own Krylov scalars,
convergence flag and NR test, so a tissue stepped in a batch must give the
result of the same tissue stepped alone with the same tile size, bit for bit.
That holds for the two-level CG of the mixed step too: every tissue of a
batch has a coarse space of its own, the one it has alone.
Every test here needs an MI355X."""
import numpy as np
import pytest

import problems

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _ellipses(n=6, seed=7):
    """Six physical tissues of 324-1716 prisms (one to 22 slices of 64 rows)."""
    from tissueorigami_amd import tissue
    return tissue.batch_ellipses(n, (200, 3000), seed=seed)


def _growth_inputs(orc, c, dt):
    from tissueorigami_amd import tissue
    inc = tissue.growth_increment(c["rate"], dt)
    return orc.inv3_many(inc)


def _alone_physical(nr, orc, c, dt, spt):
    m = c["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    s.set_growth(_growth_inputs(orc, c, dt))
    s.update_visc_surface(m["nodes"].reshape(-1), c["place"], c["etype"], c["spans"])
    s.set_tile_size(spt)
    return s


def _batch_physical(nr, orc, cs, dt, spt):
    b = nr.TissueBatch([(len(c["mesh"]["nodes"]), c["mesh"]["conn"], c["mesh"]["refpos"])
                        for c in cs])
    cat = lambda k: np.concatenate([c[k] for c in cs])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
    b.set_growth(np.concatenate([_growth_inputs(orc, c, dt) for c in cs]))
    u0 = np.concatenate([c["mesh"]["nodes"].reshape(-1) for c in cs])
    b.update_visc_surface(u0, cat("place"), cat("etype"), cat("spans"))
    b.set_tile_size(spt)
    return b, u0


@pytest.mark.parametrize("precision,rtol,spt,method", [
    ("fp64", 1e-10, 8, "bicgstab"), ("fp64", 1e-10, 2, "bicgstab"), ("fp64", 1e-10, 4, "cg"),
    ("fp64", 1e-10, 4, "gmres"),
    ("mixed", 1e-3, 8, "cg"), ("mixed", 1e-5, 8, "bicgstab"), ("mixed", 1e-10, 4, "cg")])
def test_batch_step_is_bitwise_each_tissue_alone(orc, nr, precision, rtol, spt, method):
    cs = _ellipses()
    dt = 60.0
    b, u0 = _batch_physical(nr, orc, cs, dt, spt)
    b.set_krylov(precision, method)
    ub, stats = b.step(u0, dt, 1e-8, 20, precision, rtol)
    assert len(stats) == len(cs)
    parts = b.split_nodes(ub)
    # each tissue's strain is that of its own last NR iterate
    strain = b.element_strain()
    po = b.prism_offsets
    print("NR iterations per tissue:", [st["nr_iterations"] for st in stats])
    for t, c in enumerate(cs):
        s = _alone_physical(nr, orc, c, dt, spt)
        s.set_krylov(precision, method)
        ua, st = s.step(c["mesh"]["nodes"].reshape(-1), dt, 1e-8, 20, precision, rtol)
        assert st["converged"] and stats[t]["converged"]
        assert st["krylov_iterations"] == stats[t]["krylov_iterations"], t
        assert st["du_norms"] == stats[t]["du_norms"], t
        assert np.array_equal(parts[t].reshape(-1), ua), t
        assert np.array_equal(strain[po[t]:po[t + 1]], s.element_strain()), t
        s.close()
    b.close()


# Hard tissues: randomly perturbed, badly conditioned meshes (60 to 650
# nodes) on which the FP32 inner solve of some tissues diverges.
HARD = [dict(seed=11, lx=4.0, ly=3.0, layers=2, visc=True, fixed=False),
        dict(seed=13, lx=6.0, ly=5.0, layers=3, visc=False, fixed=True),
        dict(seed=14, lx=9.0, ly=2.0, layers=1, visc=True, fixed=False),
        dict(seed=15, lx=12.0, ly=9.0, layers=4, visc=True, fixed=False)]


def _alone(nr, p, spt):
    m = p["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], p["refpos"])
    s.set_material(p["E"], p["nu"], p["eta"], p["ext_visc"], p["fixed"])
    s.set_growth(p["invFg"])
    s.set_visc_surface(p["area"])
    s.set_tile_size(spt)
    return s


def _batch(nr, ps, spt):
    b = nr.TissueBatch([(len(p["mesh"]["nodes"]), p["mesh"]["conn"], p["refpos"]) for p in ps])
    cat = lambda k: np.concatenate([np.asarray(p[k]).reshape(len(p[k]), -1) for p in ps])
    fixed = np.concatenate([np.zeros((len(p["mesh"]["nodes"]), 3), np.uint8) if p["fixed"] is None
                            else p["fixed"] for p in ps])
    b.set_material(np.concatenate([p["E"] for p in ps]), np.concatenate([p["nu"] for p in ps]),
                   np.concatenate([p["eta"] for p in ps]), cat("ext_visc"), fixed)
    b.set_growth(cat("invFg"))
    b.set_visc_surface(np.concatenate([p["area"] for p in ps]))
    b.set_tile_size(spt)
    return b


def test_mixed_falls_back_to_fp64_per_tissue(nr):
    """Where the CG or the FP32 solve of a tissue fails, that tissue alone is
    solved again with BiCGStab, then with the FP64 operator: every tissue
    converges to its FP64 root, the batch still equals each tissue alone, and
    the fallbacks are counted."""
    ps = [problems.perturbed_problem(**h) for h in HARD]
    b = _batch(nr, ps, 8)
    un = np.concatenate([p["un"] for p in ps])
    ub, stats = b.step(un, 7.0, 1e-8, 20, "mixed", 1e-5)
    parts = b.split_nodes(ub)
    assert sum(st["fp64_fallbacks"] + st["cg_fallbacks"] for st in stats) > 0
    for t, p in enumerate(ps):
        s = _alone(nr, p, 8)
        ua, st = s.step(p["un"], 7.0, 1e-8, 20, "mixed", 1e-5)
        u64, st64 = s.step(p["un"], 7.0, 1e-8, 20, "fp64", 1e-10)
        assert st["converged"] and stats[t]["converged"] and st64["converged"]
        assert st["fp64_fallbacks"] == stats[t]["fp64_fallbacks"]
        assert st["cg_fallbacks"] == stats[t]["cg_fallbacks"]
        assert np.array_equal(parts[t].reshape(-1), ua), t
        assert np.abs(ua - u64).max() <= 1e-8 * np.abs(u64).max()
        s.close()
    b.close()


def test_batch_advance_is_bitwise_each_tissue_alone(nr):
    """Device-resident stepping (growth and surfaces on the device) of three
    C1-like rectangles with different growth rates and stiffness."""
    from tissueorigami_amd import tissue
    cfgs = []
    for k, (lx, ly, layers) in enumerate([(10.0, 5.0, 2), (6.0, 4.0, 3), (14.0, 8.0, 2)]):
        m = tissue.rectangle(lx, ly, 1.0, 2.0, layers)
        ne, nn = len(m["conn"]), len(m["nodes"])
        ext = np.zeros((nn, 3))
        ext[m["node_placement"] == 1] = 16000.0
        ext[m["node_placement"] == 0] = 10.0
        rates = np.zeros((ne, 3))
        rates[:, 0] = rates[:, 1] = np.where(m["layer"] == 0, 0.01 * (k + 1), 0.0) / 3600.0
        place, spans = tissue.element_placement(m["conn"], m["node_placement"])
        etype = tissue.element_type(m["conn"], m["node_type"])
        cfgs.append(dict(m=m, E=np.full(ne, 500.0 * (k + 1)), nu=np.full(ne, 0.3),
                         eta=np.full(ne, 5.0 * k), ext=ext, rates=rates, place=place,
                         spans=spans, etype=etype))
    b = nr.TissueBatch([(len(c["m"]["nodes"]), c["m"]["conn"], c["m"]["refpos"]) for c in cfgs])
    cat = lambda k: np.concatenate([c[k] for c in cfgs])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext"))
    b.set_growth_rates(cat("rates"))
    b.set_surface_flags(cat("place"), cat("etype"), cat("spans"))
    u0 = np.concatenate([c["m"]["nodes"].reshape(-1) for c in cfgs])
    ub, bstats = b.advance(8, 10.0, u0=u0, precision="mixed", krylov_rtol=1e-5)
    parts = b.split_nodes(ub)
    for t, c in enumerate(cfgs):
        m = c["m"]
        a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        a.set_material(c["E"], c["nu"], c["eta"], c["ext"])
        a.set_growth_rates(c["rates"])
        a.set_surface_flags(c["place"], c["etype"], c["spans"])
        ua, astats = a.advance(8, 10.0, u0=m["nodes"].reshape(-1), precision="mixed",
                               krylov_rtol=1e-5)
        for k in range(8):
            assert astats[k]["converged"] and bstats[k][t]["converged"]
            assert astats[k]["krylov_iterations"] == bstats[k][t]["krylov_iterations"]
        assert np.array_equal(parts[t].reshape(-1), ua), t
        a.close()
    b.close()


def test_batch_coarse_space_per_tissue(orc, nr):
    # the per-tissue coarse correction is on by default and cuts the CG
    # iterations of the mixed step against block Jacobi alone
    cs = _ellipses()
    dt = 60.0
    its = {}
    for cell in (-1.0, 0.0):
        b, u0 = _batch_physical(nr, orc, cs, dt, 8)
        b.set_coarse(cell)
        ub, stats = b.step(u0, dt, 1e-8, 20, "mixed", 1e-5)
        agg, active = b.coarse_info()
        assert all(s["converged"] for s in stats)
        if cell < 0:
            assert agg >= len(cs) and active == 1
        else:
            assert agg == 0 and active == 0
        its[cell] = [s["krylov_iterations"] for s in stats]
        b.close()
    big = int(np.argmax([len(c["mesh"]["nodes"]) for c in cs]))
    assert its[-1.0][big] < its[0.0][big], its


def test_autotune_picks_a_candidate_and_keeps_parity(orc, nr):
    cs = _ellipses()
    dt = 60.0
    b, u0 = _batch_physical(nr, orc, cs, dt, 8)
    b.step(u0, dt, 1e-8, 20, "mixed", 1e-5)
    spt, ms = b.autotune("mixed")
    assert spt in (2, 4, 8, 16, 32) and ms > 0.0
    ub, stats = b.step(u0, dt, 1e-8, 20, "mixed", 1e-5)
    parts = b.split_nodes(ub)
    s = _alone_physical(nr, orc, cs[2], dt, spt)
    ua, st = s.step(cs[2]["mesh"]["nodes"].reshape(-1), dt, 1e-8, 20, "mixed", 1e-5)
    assert np.array_equal(parts[2].reshape(-1), ua)
    s.close(); b.close()


def test_batch_rejects_single_tissue_entry_points(orc, nr):
    cs = _ellipses(2)
    b, u0 = _batch_physical(nr, orc, cs, 60.0, 8)
    with pytest.raises(nr.TFError):
        b.assemble(u0, u0, 60.0)
    with pytest.raises(nr.TFError):
        b.csr_pattern()
    b.close()


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_graph_replays_equal_direct_launches(orc, nr, precision, rtol):
    """Krylov chunks replayed from captured hipGraphs give, step after step,
    the bits of the same chunks launched one kernel at a time (host work and
    synchronisations run between the replays)."""
    cs = _ellipses(3, seed=9)
    dt = 60.0
    ga, u0 = _batch_physical(nr, orc, cs, dt, 4)
    da, _ = _batch_physical(nr, orc, cs, dt, 4)
    da.set_graphs(False)
    ug, ud = u0.copy(), u0.copy()
    for k in range(4):
        ug, sg = ga.step(ug, dt, 1e-8, 20, precision, rtol)
        ud, sd = da.step(ud, dt, 1e-8, 20, precision, rtol)
        assert np.array_equal(ug, ud), k
        assert [s["krylov_iterations"] for s in sg] == [s["krylov_iterations"] for s in sd]
        ga.update_visc_surface(ug, np.concatenate([c["place"] for c in cs]),
                               np.concatenate([c["etype"] for c in cs]),
                               np.concatenate([c["spans"] for c in cs]))
        da.update_visc_surface(ud, np.concatenate([c["place"] for c in cs]),
                               np.concatenate([c["etype"] for c in cs]),
                               np.concatenate([c["spans"] for c in cs]))
    ga.close(); da.close()


def test_fp64_cg_matches_fp64_bicgstab(orc, nr):
    """FP64 CG on the nearly symmetric Jacobian reaches the BiCGStab step's
    positions (both linear solves to 1e-10) within 1e-9 relative, and the
    dense oracle step."""
    p = problems.perturbed_problem(seed=9, visc=True, fixed=True, lx=6.0, ly=5.0, layers=3)
    s = _alone(nr, p, 8)
    ub, stb = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-10)
    s.set_krylov("fp64", "cg")
    uc, stc = s.step(p["un"], p["dt"], 1e-8, 20, "fp64", 1e-10)
    uo, du, conv = orc.Problem(p["mesh"]["conn"], p["refpos"], p["invFg"],
                               *orc.lame(p["E"], p["nu"]), p["eta"], p["ext_visc"], p["area"],
                               p["fixed"]).nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert stb["converged"] and stc["converged"] and conv
    scale = np.abs(uo).max()
    assert np.abs(uc - ub).max() <= 1e-9 * scale
    assert np.abs(uc - uo).max() <= 1e-9 * scale
    s.close()


def _circumference_pairs(m):
    """x and y of every circumference column bound to its basal node, as
    Simulation::bindCircumferenceXY builds them (node ids local to the mesh)."""
    n2 = m["n2d"]
    circ = np.where(m["at_circumference"][:n2] == 1)[0]
    return np.array([[3 * (L * n2 + k) + d, 3 * k + d] for k in circ
                     for L in range(1, m["n_layers"] + 1) for d in (0, 1)], np.int32)


@pytest.mark.parametrize("precision,rtol,method", [
    ("fp64", 1e-10, "bicgstab"), ("mixed", 1e-3, "cg"), ("mixed", 1e-5, "bicgstab")])
def test_batch_binding_is_bitwise_each_tissue_alone(orc, nr, precision, rtol, method):
    """Circumference binding in every tissue of a batch (the pairs of tissue t
    numbered in the concatenated node order): each tissue's step equals the
    same bound tissue stepped alone, bit for bit, and its slaves move with
    their masters."""
    cs = _ellipses(4, seed=9)
    dt = 60.0
    b, u0 = _batch_physical(nr, orc, cs, dt, 8)
    offs = np.concatenate([[0], np.cumsum([len(c["mesh"]["nodes"]) for c in cs])])
    pairs = [_circumference_pairs(c["mesh"]) for c in cs]
    cat = lambda k: np.concatenate([c[k] for c in cs])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"), None,
                   np.concatenate([p + 3 * offs[t] for t, p in enumerate(pairs)]))
    b.set_krylov(precision, method)
    ub, stats = b.step(u0, dt, 1e-8, 20, precision, rtol)
    parts = b.split_nodes(ub)
    for t, c in enumerate(cs):
        s = _alone_physical(nr, orc, c, dt, 8)
        s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"], None, pairs[t])
        s.set_krylov(precision, method)
        u_start = c["mesh"]["nodes"].reshape(-1)
        ua, st = s.step(u_start, dt, 1e-8, 20, precision, rtol)
        assert st["converged"] and stats[t]["converged"]
        assert st["krylov_iterations"] == stats[t]["krylov_iterations"], t
        assert st["du_norms"] == stats[t]["du_norms"], t
        assert np.array_equal(parts[t].reshape(-1), ua), t
        d = ua - u_start
        assert np.abs(d[pairs[t][:, 0]] - d[pairs[t][:, 1]]).max() <= 1e-12
        s.close()
    b.close()


def test_batch_rejects_a_pair_across_tissues(orc, nr):
    cs = _ellipses(2, seed=9)
    b, u0 = _batch_physical(nr, orc, cs, 60.0, 8)
    n0 = len(cs[0]["mesh"]["nodes"])
    cat = lambda k: np.concatenate([c[k] for c in cs])
    with pytest.raises(nr.TFError, match="binds two tissues"):
        b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"), None,
                       np.array([[3 * n0, 0]], np.int32))
    b.close()
