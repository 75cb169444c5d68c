"""Young's modulus time series on the device (tf_nr_add_stiffness_series,
tf_nr_apply_stiffness_series, and tf_nr_advance, which applies the series at
the start of every step) against the oracle's restatement of
YoungsModulusModifier.  Every test here needs an MI355X."""
import numpy as np
import pytest

import problems
from tissueorigami_amd import tissue

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _setup(nr, c):
    m = c["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    return s


def _series(c, seed):
    m = c["mesh"]
    ne = len(m["conn"])
    rng = np.random.default_rng(seed)
    times = np.array([0.0, 240.0, 600.0, 900.0])
    grids = rng.uniform(0.5, 2.0, (4, 5, 7))
    rel = tissue.relative_positions(m["nodes"], m["conn"])
    # a few prisms outside the box on either side exercise the clamps
    rel[:5] = rng.uniform(-0.3, 1.3, (5, 2))
    app = (rng.uniform(size=ne) < 0.8).astype(np.uint8)
    return times, grids, app, rel


@pytest.mark.parametrize("linear", [True, False])
@pytest.mark.parametrize("method", [0, 1, 2])
def test_multiplier_matches_oracle(orc, nr, linear, method):
    c = problems.c1_config()
    s = _setup(nr, c)
    times, grids, app, rel = _series(c, 7 + method)
    s.add_stiffness_series(times, grids, linear, method, app, rel)
    # a second series over the same prisms, applied after the first
    g2 = np.stack([np.ones((3, 3)), np.full((3, 3), 1.2)])
    s.add_stiffness_series([0.0, 600.0], g2, True, 1, app, rel)
    o1 = orc.StiffnessSeries(times, grids, linear, method, app, rel)
    o2 = orc.StiffnessSeries([0.0, 600.0], g2, True, 1, app, rel)
    ne = len(c["mesh"]["conn"])
    E = np.asarray(c["E"], float) * np.ones(ne)
    nu = np.asarray(c["nu"], float) * np.ones(ne)
    stm = np.ones(ne)
    lam, mu = orc.lame(E, nu)
    dt = 60.0
    for k in range(18):   # crosses two interval boundaries and the end of the second series
        o1.apply(k * dt, dt, E, nu, stm, lam, mu)
        o2.apply(k * dt, dt, E, nu, stm, lam, mu)
        s.apply_stiffness_series(k * dt, dt)
    got = s.stiffness_multiplier()
    if linear:
        assert np.array_equal(got, stm)
    else:   # the device's exp and the host's may differ in the last bit
        assert np.allclose(got, stm, rtol=1e-14, atol=0)
    assert np.all(got[app == 0] == 1.0)
    s.close()


def test_step_after_series_equals_step_with_scaled_modulus(orc, nr):
    # the NR step sees multiplier x E: a handle stepped after the series gives
    # the result of a fresh handle whose E is the scaled one, bit for bit
    c = problems.c1_config()
    m = c["mesh"]
    ne = len(m["conn"])
    a = _setup(nr, c)
    times, grids, app, rel = _series(c, 3)
    a.add_stiffness_series(times, grids, True, 1, app, rel)
    for k in range(5):
        a.apply_stiffness_series(k * 60.0, 60.0)
    stm = a.stiffness_multiplier()
    assert np.any(stm != 1.0)
    b = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    b.set_material(stm * np.asarray(c["E"], float) * np.ones(ne), c["nu"], c["eta"], c["ext_visc"])
    inc = tissue.growth_increment(c["rate"], c["dt"])
    invFg = orc.inv3_many(inc)
    u0 = m["nodes"].reshape(-1)
    for h in (a, b):
        h.set_growth(invFg)
        h.update_visc_surface(u0, c["place"], c["etype"], c["spans"])
    ua, sa = a.step(u0, c["dt"], 1e-8, 20, "fp64", 1e-12)
    ub, sb = b.step(u0, c["dt"], 1e-8, 20, "fp64", 1e-12)
    assert sa["converged"] and sa["du_norms"] == sb["du_norms"]
    assert np.array_equal(ua, ub)
    a.close(); b.close()


def test_advance_applies_series_each_step(orc, nr):
    c = problems.c1_config()
    m = c["mesh"]
    ne = len(m["conn"])
    rates = np.zeros((ne, 3))
    rates[:, 0] = rates[:, 1] = c["rate"] / 3600.0
    times, grids, app, rel = _series(c, 5)
    a = _setup(nr, c)
    a.set_growth_rates(rates)
    a.set_surface_flags(c["place"], c["etype"], c["spans"])
    a.add_stiffness_series(times, grids, False, 1, app, rel)
    a.set_time(120.0)
    dt = c["dt"]
    ua, stats = a.advance(4, dt, u0=m["nodes"].reshape(-1), krylov_rtol=1e-12)
    assert all(s["converged"] for s in stats)
    assert a.time() == 120.0 + 4 * dt
    o = orc.StiffnessSeries(times, grids, False, 1, app, rel)
    E = np.asarray(c["E"], float) * np.ones(ne)
    nu = np.asarray(c["nu"], float) * np.ones(ne)
    stm = np.ones(ne)
    lam, mu = orc.lame(E, nu)
    for k in range(4):
        o.apply(120.0 + k * dt, dt, E, nu, stm, lam, mu)
    assert np.allclose(a.stiffness_multiplier(), stm, rtol=1e-14, atol=0)
    a.close()


def test_batch_series_matches_each_tissue_alone(orc, nr):
    # a series over a batch handle (prisms concatenated) gives every tissue the
    # multipliers and the step it gets alone, bit for bit (linear series)
    ts = tissue.batch_ellipses(3, (300, 1500), seed=5)
    dt = 60.0
    times = np.array([0.0, 300.0, 900.0])
    rng = np.random.default_rng(11)
    grids = rng.uniform(0.6, 1.8, (3, 4, 6))
    per = []
    for c in ts:
        m = c["mesh"]
        ne = len(m["conn"])
        per.append((tissue.relative_positions(m["nodes"], m["conn"]),
                    (rng.uniform(size=ne) < 0.7).astype(np.uint8)))
    b = nr.TissueBatch([(len(c["mesh"]["nodes"]), c["mesh"]["conn"], c["mesh"]["refpos"])
                        for c in ts])
    cat = lambda k: np.concatenate([c[k] for c in ts])
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
    b.set_tile_size(8)
    b.add_stiffness_series(times, grids, True, 1, np.concatenate([a for _, a in per]),
                           np.concatenate([r for r, _ in per]))
    inc = [orc.inv3_many(tissue.growth_increment(c["rate"], dt)) for c in ts]
    b.set_growth(np.concatenate(inc))
    u0 = np.concatenate([c["mesh"]["nodes"].reshape(-1) for c in ts])
    b.update_visc_surface(u0, cat("place"), cat("etype"), cat("spans"))
    for k in range(6):
        b.apply_stiffness_series(k * dt, dt)
    ub, bst = b.step(u0, dt, 1e-8, 20, "fp64", 1e-10)
    mb = b.stiffness_multiplier()
    parts = b.split_nodes(ub)
    off = 0
    for t, c in enumerate(ts):
        m = c["mesh"]
        ne = len(m["conn"])
        a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        a.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        a.set_tile_size(8)
        a.add_stiffness_series(times, grids, True, 1, per[t][1], per[t][0])
        a.set_growth(inc[t])
        a.update_visc_surface(m["nodes"].reshape(-1), c["place"], c["etype"], c["spans"])
        for k in range(6):
            a.apply_stiffness_series(k * dt, dt)
        ua, ast = a.step(m["nodes"].reshape(-1), dt, 1e-8, 20, "fp64", 1e-10)
        assert np.array_equal(mb[off:off + ne], a.stiffness_multiplier()), t
        assert ast["du_norms"] == bst[t]["du_norms"], t
        assert np.array_equal(parts[t].reshape(-1), ua), t
        off += ne
        a.close()
    b.close()


def test_series_rejects_bad_arguments(nr):
    c = problems.c1_config()
    s = _setup(nr, c)
    ne = len(c["mesh"]["conn"])
    rel = np.zeros((ne, 2)); app = np.ones(ne, np.uint8)
    with pytest.raises(nr.TFError):
        s.add_stiffness_series([0.0], np.ones((1, 3, 3)), True, 0, app, rel)   # one time point
    with pytest.raises(nr.TFError):
        s.add_stiffness_series([0.0, 1.0], np.ones((2, 1, 3)), True, 0, app, rel)  # ny < 2
    with pytest.raises(nr.TFError):
        s.add_stiffness_series([0.0, 1.0], np.ones((2, 3, 3)), True, 3, app, rel)  # method
    s.close()
