"""The FP64 two-level CG (FP64 operator, vectors and coarse inverse, the
coarse space of the mixed step): the same root as FP64 BiCGStab, fewer
iterations than without the coarse space, and a tissue in a batch solved as
alone.  Every test here needs an MI355X."""
import numpy as np
import pytest

from tissueorigami_amd import tissue, workloads

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
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    return s


def test_fp64_cg_with_coarse_space_finds_the_same_root(nr):
    w = workloads.make_workload("c2")
    u0 = w["mesh"]["nodes"].reshape(-1)
    ref = _solver(nr, w)
    ub, sb = ref.advance(2, w["dt"], u0=u0, precision="fp64", krylov_rtol=1e-10)
    cg = _solver(nr, w)
    cg.set_krylov("fp64", "cg")
    uc, sc = cg.advance(2, w["dt"], u0=u0, precision="fp64", krylov_rtol=1e-10)
    n_agg, active = cg.coarse_info()
    assert active and n_agg > 0
    assert all(x["converged"] for x in sb + sc)
    assert all(x["cg_fallbacks"] == 0 for x in sc)
    scale = np.abs(ub).max()
    assert np.abs(uc - ub).max() <= 1e-9 * scale
    # inexact Newton with the mixed step's forcing term: the same root
    inexact = _solver(nr, w)
    inexact.set_krylov("fp64", "cg")
    ui, si = inexact.advance(2, w["dt"], u0=u0, precision="fp64", krylov_rtol=1e-3)
    assert all(x["converged"] for x in si)
    assert np.abs(ui - ub).max() <= 1e-8 * scale
    inexact.close()
    # the coarse space cuts the iterations
    plain = _solver(nr, w)
    plain.set_krylov("fp64", "cg")
    plain.set_coarse(0.0)
    up, sp = plain.advance(2, w["dt"], u0=u0, precision="fp64", krylov_rtol=1e-10)
    assert np.abs(up - ub).max() <= 1e-9 * scale
    kc = sum(x["krylov_iterations_total"] for x in sc)
    kp = sum(x["krylov_iterations_total"] for x in sp)
    assert kc < 0.6 * kp, (kc, kp)
    for h in (ref, cg, plain):
        h.close()


def test_fp64_cg_batch_equals_each_tissue_alone(nr):
    ts = tissue.batch_ellipses(3, (800, 3000), seed=6)
    dt = 60.0

    def setup(h, ms):
        cat = lambda k: np.concatenate([t[k] for t in ms])
        h.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
        inc = np.concatenate([tissue.growth_increment(t["rate"], dt) for t in ms])
        h.set_growth(np.linalg.inv(inc.reshape(-1, 3, 3)).reshape(-1, 9))
        h.set_visc_surface(np.concatenate([np.ones(len(t["mesh"]["nodes"])) for t in ms]))
        h.set_krylov("fp64", "cg")
        h.set_tile_size(2)

    b = nr.TissueBatch([(len(t["mesh"]["nodes"]), t["mesh"]["conn"], t["mesh"]["refpos"])
                        for t in ts])
    setup(b, ts)
    u0 = np.concatenate([t["mesh"]["nodes"].reshape(-1) for t in ts])
    ub, sb = b.step(u0, dt, 1e-8, 20, "fp64", 1e-10)
    off = 0
    for i, t in enumerate(ts):
        m = t["mesh"]
        a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        setup(a, [t])
        ua, sa = a.step(m["nodes"].reshape(-1), dt, 1e-8, 20, "fp64", 1e-10)
        n = 3 * len(m["nodes"])
        assert sa["converged"] and sb[i]["converged"]
        assert np.array_equal(ub[off:off + n], ua)
        assert sb[i]["krylov_iterations"] == sa["krylov_iterations"]
        off += n
        a.close()
    b.close()
