"""The batched sweep workload itself: 64 elliptical tissues of 1.7 k-37 k
prisms (prism counts log-uniform in 2 k-40 k, 2-6 layers, E uniform in
100-2000 Pa, seeded growth fields), 853 k prisms in one batch handle, stepped
in mixed precision with the autotuned tile size.  Every tissue must converge
to a root of its FP64 residual; four of them (the smallest, the largest and
two between) must equal the tissue stepped alone bit for bit, and their FP64
step must match the oracle's sparse step.  Every test here needs an MI355X."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 60.0
N_STEPS = 2


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


@pytest.fixture(scope="module")
def c4(nr):
    """The batch stepped N_STEPS mixed steps from the meshes' reference
    positions with the autotuned tile, positions and growth kept per step."""
    from tissueorigami_amd import tissue
    ts = tissue.batch_ellipses(64, seed=4)
    meshes = [(len(c["mesh"]["nodes"]), c["mesh"]["conn"], c["mesh"]["refpos"]) for c in ts]
    cat = lambda k: np.concatenate([c[k] for c in ts])
    b = nr.TissueBatch(meshes)
    b.set_material(cat("E"), cat("nu"), cat("eta"), cat("ext_visc"))
    b.set_growth_rates(np.concatenate([tissue.rates_xyz(c["rate"]) for c in ts]))
    b.set_surface_flags(cat("place"), cat("etype"), cat("spans"))
    b.set_krylov("mixed", "cg")
    u0 = np.concatenate([c["mesh"]["nodes"].reshape(-1) for c in ts])
    # one step to have an assembled operator for the tuner, then start over
    b.advance(1, DT, u0=u0, precision="mixed", krylov_rtol=1e-3, want_positions=False)
    spt, _ = b.autotune("mixed")
    b.set_tile_size(spt)
    ne = b.n_prisms
    b.growth_tensor(np.tile(np.eye(3).reshape(9), (ne, 1)))
    b.set_time(0.0)
    us, Fgs, stats = [u0], [], []
    for k in range(N_STEPS):
        u, st = b.advance(1, DT, u0=us[-1] if k == 0 else None, precision="mixed",
                          krylov_rtol=1e-3)
        us.append(u)
        Fgs.append(b.growth_tensor())
        stats.append(st[0])
    b.close()
    return dict(ts=ts, spt=spt, us=us, Fgs=Fgs, stats=stats, node_off=b.node_offsets,
                prism_off=b.prism_offsets)


def _part(a, off, t, k):
    a = np.asarray(a).reshape(-1, k)
    return a[off[t]:off[t + 1]].reshape(-1)


def _sample(c4):
    sizes = [len(c["mesh"]["conn"]) for c in c4["ts"]]
    order = np.argsort(sizes)
    return [int(order[0]), int(order[len(order) // 3]), int(order[2 * len(order) // 3]),
            int(order[-1])]


def test_c4_workload_shape(c4):
    ts = c4["ts"]
    sizes = np.array([len(c["mesh"]["conn"]) for c in ts])
    assert len(ts) == 64 and 800_000 < sizes.sum() < 900_000
    assert sizes.min() >= 1500 and sizes.max() <= 42_000
    assert {c["layers"] for c in ts} <= {2, 3, 4, 5, 6}
    assert c4["spt"] in (2, 4, 8, 16, 32)


def test_c4_every_tissue_converges_to_its_fp64_root(nr, c4):
    """Every tissue of every step converged, and the FP64 residual assembled at
    its converged positions (its own growth tensor and surfaces) is at most
    1e-6 of the residual at the start of the step."""
    ts, off, poff = c4["ts"], c4["node_off"], c4["prism_off"]
    for k, st in enumerate(c4["stats"]):
        assert all(s["converged"] and s["status"] == nr.TF_OK for s in st), k
    k = N_STEPS - 1  # the last step: start us[k], result us[k + 1]
    worst = 0.0
    for t, c in enumerate(ts):
        m = c["mesh"]
        s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        Fg = _part(c4["Fgs"][k], poff, t, 9).reshape(-1, 3, 3)
        s.set_growth(np.linalg.inv(Fg).reshape(-1, 9))
        ustart = _part(c4["us"][k], off, t, 3)
        uend = _part(c4["us"][k + 1], off, t, 3)
        s.update_visc_surface(ustart, c["place"], c["etype"], c["spans"])
        g0, _ = s.assemble(ustart, ustart, DT)
        g1, _ = s.assemble(uend, ustart, DT)
        s.close()
        r = np.linalg.norm(g1) / np.linalg.norm(g0)
        worst = max(worst, r)
        assert r <= 1e-6, (t, r)
    print(f"C4 worst |g(u)| / |g(u_n)| over 64 tissues: {worst:.3e}")


def test_c4_sampled_tissues_equal_the_tissue_alone(nr, c4):
    """The smallest, the largest and two tissues between, each stepped alone
    with the same inputs and tile size: the same bits and Krylov counts."""
    from tissueorigami_amd import tissue
    ts, off = c4["ts"], c4["node_off"]
    for t in _sample(c4):
        c = ts[t]
        m = c["mesh"]
        s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        s.set_growth_rates(tissue.rates_xyz(c["rate"]))
        s.set_surface_flags(c["place"], c["etype"], c["spans"])
        s.set_krylov("mixed", "cg")
        u = m["nodes"].reshape(-1)
        # the batch's sequence: one step, then back to the start state
        s.advance(1, DT, u0=u, precision="mixed", krylov_rtol=1e-3, want_positions=False)
        s.set_tile_size(c4["spt"])
        s.growth_tensor(np.tile(np.eye(3).reshape(9), (len(m["conn"]), 1)))
        s.set_time(0.0)
        for k in range(N_STEPS):
            u, st = s.advance(1, DT, u0=u if k == 0 else None, precision="mixed",
                              krylov_rtol=1e-3)
            assert np.array_equal(u, _part(c4["us"][k + 1], off, t, 3)), (t, k)
            sb = c4["stats"][k][t]
            assert st[0]["krylov_iterations"] == sb["krylov_iterations"], (t, k)
            assert st[0]["nr_iterations"] == sb["nr_iterations"], (t, k)
        s.close()


def test_c4_sampled_tissues_fp64_step_matches_oracle(orc, nr, c4):
    """The first step of the four sampled tissues in FP64 on the device against
    the oracle's BSR + BiCGStab step (both linear solves to 1e-12) from the
    same growth tensor and surfaces: positions within 1e-9 relative, the same
    NR iteration count; and the batch's mixed result within 1e-8 of both."""
    ts, off, poff = c4["ts"], c4["node_off"], c4["prism_off"]
    for t in _sample(c4):
        c = ts[t]
        m = c["mesh"]
        nn = len(m["nodes"])
        Fg = _part(c4["Fgs"][0], poff, t, 9)
        inv = orc.inv3_many(Fg)
        u0 = m["nodes"].reshape(-1).copy()
        A = orc.visc_surfaces(nn, m["conn"], u0, c["place"], c["etype"], c["spans"])
        s = nr.NewtonRaphsonSolver(nn, m["conn"], m["refpos"])
        s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
        s.set_growth(inv)
        s.set_visc_surface(A)
        ug, st = s.step(u0, DT, 1e-8, 20, "fp64", 1e-12)
        s.close()
        lam, mu = orc.lame(c["E"], c["nu"])
        P = orc.Problem(m["conn"], m["refpos"], inv, lam, mu, c["eta"], c["ext_visc"], A)
        uo, du, kit, conv = P.nr_step_sparse(u0, DT, 1e-8, 20, 1e-12, 50000)
        assert st["converged"] and conv
        assert st["nr_iterations"] == len(du), t
        scale = np.abs(uo).max()
        assert np.abs(ug - uo).max() <= 1e-9 * scale, t
        um = _part(c4["us"][1], off, t, 3)
        assert np.abs(um - uo).max() <= 1e-8 * scale, t
