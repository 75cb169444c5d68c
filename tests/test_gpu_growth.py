"""Growth functions on the device (tf_nr_set_growth_prisms, tf_nr_add_uniform_growth,
tf_nr_add_grid_growth, tf_nr_set_column_volume_conservation,
tf_nr_update_relative_positions, tf_nr_grow, and tf_nr_advance running them
every step) against the oracle's restatement, on the reference's regression
runs and on synthetic tissues.  Every test here needs an MI355X."""
import math

import numpy as np
import pytest

import replay
from tissueorigami_amd import tissue, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _ident(n):
    return np.tile(np.eye(3).reshape(9), (n, 1))


def _load(s, P, fns, ex, interp, pinned=False, pin_times=()):
    s.set_growth_prisms(P["type"], P["place"], P["flags"], P["rel_z"], P.get("weights"),
                        ex.get("apical_of"), ex.get("excluded"), ex.get("symmetric_y", False),
                        P.get("explicit_ecm", False))
    for f in fns:
        if f["type"] == 1:
            s.add_uniform_growth(f["t"][0], f["t"][1], f["rate"], f["angle"], f["apply"],
                                 f.get("z", (0.0, 1.0)))
        else:
            s.add_grid_growth(f["t"][0], f["t"][1], f["rates"], f["angles_deg"], f["apply"],
                              f.get("z", (0.0, 1.0)))
    s.set_growth_options(interp, pinned, pin_times)
    if ex.get("columns") is not None:
        s.set_column_volume_conservation(ex["columns"])


@pytest.mark.parametrize("run", ["run07007", "run07008"])
def test_fixture_growth_matches_oracle(orc, nr, run):
    """Every growing step of the regression run, from the saved F_g and the
    aligned start positions: the device's F_g against the oracle's (which is
    pinned to Save_Growth), and the device's relative positions bit for bit."""
    d = replay.load(run)
    P, fns, ex = replay.growth_setup(d)
    conn = d["conn"]
    ne, nn = len(conn), len(d["mesh_nodes"])
    s = nr.NewtonRaphsonSolver(nn, conn, d["refpos"])
    _load(s, P, fns, ex, ex["interp"], pinned=ex["pinned"])
    u0 = replay.aligned_start(d, 0)
    s.update_relative_positions(u0.reshape(-1))          # also the pinned positions
    pin = orc.relative_positions(u0, conn, ex["excluded"], ex["symmetric_y"], ex["apical_of"])
    dt = float(d["dt"])
    for k in range(6):
        u = replay.aligned_start(d, k)
        Fprev = _ident(ne) if k == 0 else d["Fg"][k - 1]
        s.growth_tensor(Fprev)
        s.update_relative_positions(u.reshape(-1))
        cur, pinned = s.relative_positions()
        rel = orc.relative_positions(u, conn, ex["excluded"], ex["symmetric_y"], ex["apical_of"])
        assert np.array_equal(cur, rel) and np.array_equal(pinned, pin)
        s.grow(k * dt, dt, u.reshape(-1))
        got = s.growth_tensor()
        Fo, _ = orc.growth_step(conn, d["refpos"], u, P, fns, pin if ex["pinned"] else rel,
                                k * dt, dt, ex["interp"], Fprev, columns=ex["columns"])
        # the device's exp / cos / sin and the host's may differ in the last bit
        assert np.abs(got - Fo).max() <= 1e-14, (k, np.abs(got - Fo).max())
        keep = ~ex["ecm"]
        tol = 1e-9 if run == "run07008" else 1e-13   # run07008: 10-digit positions of turned prisms
        assert np.abs(got - d["Fg"][k])[keep].max() <= tol, k
    s.close()


def _synthetic(seed=3, layers=3):
    m = tissue.ellipse(10.0, 7.0, 1.0, 3.0, layers)
    conn = m["conn"]
    ne = len(conn)
    rng = np.random.default_rng(seed)
    place, _ = tissue.element_placement(conn, m["node_placement"])
    etype = tissue.element_type(conn, m["node_type"]).astype(np.int32)
    etype[rng.uniform(size=ne) < 0.15] = 2          # some linker prisms
    flags = np.zeros(ne, np.uint8)
    flags[place == 1] = 4                           # actin apical layer
    _, L = tissue.tissue_height(m["nodes"], conn, m["node_type"], m["node_placement"])
    stacks = tissue.element_stacks(conn, etype, place, L)
    P = dict(type=etype, place=place.astype(np.int32), flags=flags,
             rel_z=tissue.initial_z_positions(m["nodes"], conn, m["node_type"], m["node_placement"]),
             weights=rng.uniform(0.2, 0.9, (ne, 2)), explicit_ecm=False)
    grid_r = rng.uniform(-2e-5, 6e-5, (6, 5, 3))
    grid_a = rng.uniform(-90.0, 270.0, (6, 5))
    grid_r[1, 2, :2] = (1e-5, 1.02e-5)              # an eliminated angle
    fns = [dict(type=3, t=(0.0, 1e9), apply=3, z=(0.0, 1.0), rates=grid_r, angles_deg=grid_a),
           dict(type=1, t=(60.0, 1e9), apply=1, z=(0.2, 0.9), rate=(1e-5, -5e-6, 2e-6),
                angle=0.7)]
    ex = dict(apical_of=tissue.apical_positioning(stacks, etype, place, ne, L),
              excluded=None, symmetric_y=False,
              columns=tissue.conserved_columns(stacks, etype, place, np.zeros(ne, bool),
                                               flags == 4, L))
    return m, P, fns, ex


@pytest.mark.parametrize("interp", [0, 1])
@pytest.mark.parametrize("conserve", [False, True])
def test_synthetic_growth_matches_oracle(orc, nr, interp, conserve):
    """Grid and uniform functions, linker weights, actin, z range, a tissue
    turned by 0.5 rad (every prism's growth rotated), with and without column
    volume conservation, two steps."""
    m, P, fns, ex = _synthetic()
    if not conserve:
        ex = dict(ex, columns=None)
    conn = m["conn"]
    ne = len(conn)
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), conn, m["refpos"])
    _load(s, P, fns, ex, interp)
    c, sn = math.cos(0.5), math.sin(0.5)
    R = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1.0]])
    u = m["nodes"] @ R.T
    Fo = _ident(ne)
    for k, t in enumerate((0.0, 60.0)):
        uk = u * (1.0 + 0.01 * k)
        s.update_relative_positions(uk.reshape(-1))
        s.grow(t, 60.0, uk.reshape(-1))
        rel = orc.relative_positions(uk, conn, None, False, ex["apical_of"])
        Fo, _ = orc.growth_step(conn, m["refpos"], uk, P, fns, rel, t, 60.0, interp, Fo,
                                columns=ex["columns"])
    got = s.growth_tensor()
    assert np.abs(got - Fo).max() <= 1e-13
    assert np.abs(got - _ident(ne)).max() > 1e-3
    if conserve:
        V = orc.reference_volumes(m["refpos"])
        det = np.linalg.det(got.reshape(-1, 3, 3))
        for col in ex["columns"]:
            col = col[col >= 0]
            assert abs(det[col] @ V[col] - V[col].sum()) < 1e-12 * V[col].sum()
    s.close()


def test_batch_growth_equals_each_tissue_alone(nr):
    """Two tissues of different shape in a batch, each in its own bounding box:
    F_g bit for bit as each alone."""
    parts = [_synthetic(seed=3), _synthetic(seed=4, layers=2)]
    b = nr.TissueBatch([(len(m["nodes"]), m["conn"], m["refpos"]) for m, _, _, _ in parts])
    off = np.cumsum([0] + [len(m["conn"]) for m, _, _, _ in parts])
    cat = lambda key: np.concatenate([P[key] for _, P, _, _ in parts])
    P = dict(type=cat("type"), place=cat("place"), flags=cat("flags"), rel_z=cat("rel_z"),
             weights=cat("weights"))
    ap = np.concatenate([np.where(ex["apical_of"] >= 0, ex["apical_of"] + off[i], -1)
                         for i, (_, _, _, ex) in enumerate(parts)])
    cols = [np.where(ex["columns"] >= 0, ex["columns"] + off[i], -1)
            for i, (_, _, _, ex) in enumerate(parts)]
    width = max(c.shape[1] for c in cols)
    cols = np.concatenate([np.pad(c, ((0, 0), (0, width - c.shape[1])), constant_values=-1)
                           for c in cols])
    fns = parts[0][2]
    _load(b, P, fns, dict(apical_of=ap, columns=cols), 1)
    u = np.concatenate([m["nodes"] for m, _, _, _ in parts])
    b.update_relative_positions(u.reshape(-1))
    b.grow(100.0, 60.0, u.reshape(-1))
    Fb = b.growth_tensor()
    for i, (m, Pi, _, ex) in enumerate(parts):
        a = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        _load(a, Pi, fns, ex, 1)
        a.update_relative_positions(m["nodes"].reshape(-1))
        a.grow(100.0, 60.0, m["nodes"].reshape(-1))
        assert np.array_equal(Fb[off[i]:off[i + 1]], a.growth_tensor())
        a.close()
    b.close()


def test_advance_refreshes_relative_positions_every_step(orc, nr):
    """tf_nr_advance with a grid growth function and a bilinear stiffness
    series on the growing tissue: per step the relative positions at the step's
    start positions, the series, the growth, the surfaces and the NR step, as the
    oracle replays them step by step.  The same replay with the positions of the
    first step kept throughout ends elsewhere, so stale positions would fail."""
    c = workloads.c1_config()
    m = c["mesh"]
    conn = m["conn"]
    ne, nn = len(conn), len(m["nodes"])
    place = c["place"].astype(np.int32)
    etype = c["etype"].astype(np.int32)
    P = dict(type=etype, place=place, flags=np.zeros(ne, np.uint8),
             rel_z=tissue.initial_z_positions(m["nodes"], conn, m["node_type"], m["node_placement"]),
             weights=None, explicit_ecm=False)
    rng = np.random.default_rng(11)
    grid = dict(type=3, t=(0.0, 1e9), apply=1, z=(0.0, 1.0),
                rates=rng.uniform(0.0, 4e-4, (5, 3, 3)) * np.array([1.0, 1.0, 0.05]),
                angles_deg=rng.uniform(0.0, 180.0, (5, 3)))
    times = np.array([0.0, 1000.0])
    grids = rng.uniform(0.5, 2.0, (2, 4, 9))
    app = np.ones(ne, np.uint8)
    s = nr.NewtonRaphsonSolver(nn, conn, m["refpos"])
    s.set_material(c["E"], c["nu"], c["eta"], c["ext_visc"])
    s.set_surface_flags(c["place"], c["etype"], c["spans"])
    _load(s, P, [grid], dict(apical_of=None, excluded=None, columns=None), 1)
    s.add_stiffness_series(times, grids, True, 1, app)
    dt, n = 60.0, 4
    u0 = m["nodes"].reshape(-1)
    ua, stats = s.advance(n, dt, u0=u0, krylov_rtol=1e-12)
    assert all(x["converged"] for x in stats)

    def replay_oracle(stale):
        u = u0.copy()
        Fg = _ident(ne)
        E = np.asarray(c["E"], float)
        nu = np.asarray(c["nu"], float)
        stm = np.ones(ne)
        lam, mu = orc.lame(E, nu)
        rel0 = orc.relative_positions(u0, conn)
        for k in range(n):
            t = k * dt
            rel = rel0 if stale else orc.relative_positions(u, conn)
            orc.StiffnessSeries(times, grids, True, 1, app, rel).apply(t, dt, E, nu, stm, lam, mu)
            Fg, inv = orc.growth_step(conn, m["refpos"], u, P, [grid], rel, t, dt, 1, Fg)
            A = orc.visc_surfaces(nn, conn, u, c["place"], c["etype"], c["spans"])
            prob = orc.Problem(conn, m["refpos"], inv, lam, mu, c["eta"], c["ext_visc"], A)
            u, _, _, conv = prob.nr_step_sparse(u, dt, 1e-8, 20, krylov_rtol=1e-13)
            assert conv
        return u, stm, Fg

    uo, stm, Fg = replay_oracle(False)
    assert np.allclose(s.stiffness_multiplier(), stm, rtol=1e-14, atol=0)
    assert np.abs(s.growth_tensor() - Fg).max() <= 1e-12
    scale = np.abs(uo).max()
    assert np.abs(ua - uo).max() <= 1e-9 * scale
    us, stm_s, _ = replay_oracle(True)
    assert np.abs(us - uo).max() > 1e3 * np.abs(ua - uo).max()
    assert np.abs(stm_s - stm).max() > 1e-6
    s.close()


def test_growth_argument_errors(nr):
    m, P, fns, ex = _synthetic()
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
    with pytest.raises(nr.TFError) as e:      # functions need the prisms first
        s.add_uniform_growth(0.0, 1.0, (1e-5, 0, 0), 0.0)
    assert e.value.code == nr.TF_ERR_ARG
    with pytest.raises(nr.TFError):           # positions needed before growth
        s.grow(0.0, 60.0)
    _load(s, P, fns, dict(ex, columns=None), 1)
    with pytest.raises(nr.TFError):           # no relative positions yet
        s.grow(0.0, 60.0, m["nodes"].reshape(-1))
    bad = P["type"].copy(); bad[0] = 5
    with pytest.raises(nr.TFError):
        s.set_growth_prisms(bad, P["place"], P["flags"], P["rel_z"])
    with pytest.raises(nr.TFError):
        s.add_grid_growth(0.0, 1.0, np.zeros((1, 3, 3)), np.zeros((1, 3)))
    cols = ex["columns"].copy()
    first = np.where(cols[:, 0] >= 0)[0]
    cols[first[1], 0] = cols[first[0], 0]
    with pytest.raises(nr.TFError):           # a prism in two columns
        s.set_column_volume_conservation(cols)
    s.close()
