"""Solve policies of the inexact Newton loop.

Warm start of a step's first Krylov solve (tf_nr_set_warm_start): the
increment still meets rtol |g|, so a trajectory stepped with the warm start
lands where the same trajectory stepped from zero lands, within the Newton
tolerance; the first solves of later steps take fewer Krylov iterations;
positions handed in start a trajectory without history, so repeating a call
repeats its bits.

Forcing-term safeguard (tf_nr_set_forcing_safeguard): a solve predicted to be
a step's last runs to a looser tolerance; the step converges at the same NR
iteration to within a tenth of its last increment, and tight solves are left
alone.  Every test here needs an MI355X."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


@pytest.fixture(scope="module")
def c2():
    from tissueorigami_amd.workloads import make_workload
    return make_workload("c2")


def _solver(nr, w, warm, safeguard=True):
    m = w["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    s.set_material(w["E"], w["nu"], w["eta"], w["ext"])
    s.set_growth_rates(w["rates"])
    s.set_surface_flags(w["place"], w["etype"], w["spans"])
    s.set_warm_start(warm)
    s.set_forcing_safeguard(safeguard)
    return s


def _trajectory(nr, w, warm, precision, rtol, steps=4, safeguard=True):
    s = _solver(nr, w, warm, safeguard)
    u, st = s.advance(steps, w["dt"], u0=w["mesh"]["nodes"].reshape(-1), precision=precision,
                      krylov_rtol=rtol)
    s.close()
    return u, st


@pytest.mark.parametrize("precision,rtol,frac", [("mixed", 1e-3, 0.8), ("fp64", 1e-10, 0.95)])
def test_warm_trajectory_lands_on_the_cold_one(nr, c2, precision, rtol, frac):
    """C2, four steps: with and without the warm start both converge every
    step and end within 1e-9 relative of each other (the NR tolerance on each
    step's increment is 1e-8 absolute, about 1e-10 of the positions); the
    first solves of steps 2-4 need at most frac of the cold start's Krylov
    iterations in all (the guess removes a fixed factor of the residual, a
    larger share of a 1e-3 solve than of a 1e-10 one, whose BiCGStab counts
    also vary by some ten per cent from step to step)."""
    uw, sw = _trajectory(nr, c2, True, precision, rtol)
    uc, sc = _trajectory(nr, c2, False, precision, rtol)
    print(precision, "warm", [st["krylov_iterations"] for st in sw])
    print(precision, "cold", [st["krylov_iterations"] for st in sc])
    assert all(st["converged"] for st in sw) and all(st["converged"] for st in sc)
    scale = np.abs(uc).max()
    assert np.abs(uw - uc).max() <= 1e-9 * scale
    # the first step has no history: identical
    assert sw[0]["krylov_iterations"] == sc[0]["krylov_iterations"]
    warm = sum(sw[k]["krylov_iterations"][0] for k in range(1, 4))
    cold = sum(sc[k]["krylov_iterations"][0] for k in range(1, 4))
    assert warm <= frac * cold, (warm, cold)


def test_warm_trajectory_is_reproducible_and_positions_reset_history(nr, c2):
    """A trajectory repeated from the same positions repeats its bits (the
    positions handed in drop the history of the first run), and a call that
    continues the trajectory (no positions) uses the history."""
    s = _solver(nr, c2, True)
    u0 = c2["mesh"]["nodes"].reshape(-1)
    fg0 = np.tile(np.eye(3).reshape(9), (len(c2["mesh"]["conn"]), 1))
    s.growth_tensor(fg0)
    ua, sa = s.advance(3, c2["dt"], u0=u0, precision="mixed", krylov_rtol=1e-3)
    fg3 = s.growth_tensor()
    s.growth_tensor(fg0)
    ub, sb = s.advance(3, c2["dt"], u0=u0, precision="mixed", krylov_rtol=1e-3)
    assert np.array_equal(ua, ub)
    assert [st["krylov_iterations"] for st in sa] == [st["krylov_iterations"] for st in sb]
    # continuing: the next step starts warm, a restarted one cold
    _, sc = s.advance(1, c2["dt"], precision="mixed", krylov_rtol=1e-3)
    s.close()
    s2 = _solver(nr, c2, True)
    s2.growth_tensor(fg3)
    _, sd = s2.advance(1, c2["dt"], u0=ub, precision="mixed", krylov_rtol=1e-3)
    s2.close()
    assert sc[0]["krylov_iterations"][0] < sd[0]["krylov_iterations"][0]


def test_batch_warm_start_is_bitwise_each_tissue_alone(nr):
    """Three uneven tissues stepped four times in one batch handle give, per
    tissue, the bits of the tissue stepped alone with its own history."""
    from tissueorigami_amd import tissue
    ts = tissue.batch_ellipses(3, prism_range=(1500, 6000), seed=11)
    dt = 60.0
    meshes = [(len(c["mesh"]["nodes"]), c["mesh"]["conn"], c["mesh"]["refpos"]) for c in ts]
    cat = lambda k: np.concatenate([c[k] for c in ts])
    b = nr.TissueBatch(meshes)
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
    b.set_growth_rates(np.concatenate([tissue.rates_xyz(c["rate"]) for c in ts]))
    b.set_surface_flags(cat("place"), cat("etype"), cat("spans"))
    u0 = np.concatenate([c["mesh"]["nodes"].reshape(-1) for c in ts])
    ub, _ = b.advance(4, dt, u0=u0, precision="mixed", krylov_rtol=1e-3)
    parts = b.split_nodes(ub)
    b.close()
    for c, ubt in zip(ts, parts):
        m = c["mesh"]
        a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        a.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        a.set_growth_rates(tissue.rates_xyz(c["rate"]))
        a.set_surface_flags(c["place"], c["etype"], c["spans"])
        ua, _ = a.advance(4, dt, u0=m["nodes"].reshape(-1), precision="mixed", krylov_rtol=1e-3)
        a.close()
        assert np.array_equal(ua, ubt.reshape(-1))


def test_forcing_safeguard_loosens_only_the_last_solve(nr, c2):
    """C2 mixed (forcing term 1e-3), three steps from zero with and without
    the safeguard: the same NR iteration counts, the same first two solves,
    no solve with more Krylov iterations and a last solve with clearly fewer,
    and positions within 1e-9 relative; an FP64 step at 1e-10 (a tight solve)
    gives the same bits."""
    uo, so = _trajectory(nr, c2, False, "mixed", 1e-3, steps=3, safeguard=True)
    uf, sf = _trajectory(nr, c2, False, "mixed", 1e-3, steps=3, safeguard=False)
    print("safeguard on ", [st["krylov_iterations"] for st in so])
    print("safeguard off", [st["krylov_iterations"] for st in sf])
    assert all(st["converged"] for st in so + sf)
    assert np.abs(uo - uf).max() <= 1e-9 * np.abs(uf).max()
    for a, b in zip(so, sf):
        assert a["nr_iterations"] == b["nr_iterations"]
        ka, kb = a["krylov_iterations"], b["krylov_iterations"]
        assert ka[:2] == kb[:2]
        assert all(x <= y for x, y in zip(ka, kb))
        assert ka[-1] <= 0.6 * kb[-1]
    ud, sd = _trajectory(nr, c2, False, "fp64", 1e-10, steps=1, safeguard=True)
    ue, se = _trajectory(nr, c2, False, "fp64", 1e-10, steps=1, safeguard=False)
    assert np.array_equal(ud, ue)
    assert sd[0]["krylov_iterations"] == se[0]["krylov_iterations"]


def test_policies_c1_advance_matches_oracle_steps(orc, nr):
    """C1 (plain rectangle), ten device-resident mixed steps with the warm
    start and the safeguard on (forcing term 1e-3) against the oracle's dense
    NR steps with the same growth and surfaces: every step within 1e-8
    relative."""
    import sys, os
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import problems
    from tissueorigami_amd import tissue
    c = problems.c1_config()
    m = c["mesh"]
    ne, nn = len(m["conn"]), len(m["nodes"])
    rates = np.zeros((ne, 3))
    rates[:, 0] = rates[:, 1] = c["rate"] / 3600.0
    s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    s.set_growth_rates(rates)
    s.set_surface_flags(c["place"], c["etype"], c["spans"])
    u0 = m["nodes"].reshape(-1).copy()
    lam, mu = orc.lame(c["E"], c["nu"])
    inc = tissue.growth_increment(c["rate"], c["dt"])
    Fg = np.tile(np.eye(3), (ne, 1, 1))
    u_o = u0.copy()
    u_g = None
    for k in range(10):
        u_g, st = s.advance(1, c["dt"], u0=u0 if k == 0 else None, precision="mixed",
                            krylov_rtol=1e-3)
        Fg = inc @ Fg
        A_o = orc.visc_surfaces(nn, m["conn"], u_o, c["place"], c["etype"], c["spans"])
        P = orc.Problem(m["conn"], m["refpos"], orc.inv3_many(Fg), lam, mu, c["eta"],
                        c["ext_visc"], A_o)
        u_o, du, conv = P.nr_step_dense(u_o, c["dt"], 1e-8, 20)
        assert st[0]["converged"] and conv
        assert np.abs(u_g - u_o).max() <= 1e-8 * np.abs(u_o).max(), k
    s.close()
