"""Pins the oracle's growth producers (oracle/oracle_growth.cpp) against facts
that do not come from its own formulas: numpy's SVD for the polar factor,
known answers for the compatible grid corners, closed forms for the growth
step, and the reference's own Save_Growth output of run07008 (uniform growth,
rotated prisms) and run07007 (grid growth with column volume conservation).
Also checks the host set-up helpers of tissueorigami_amd.tissue (tissue
height, stacks, apical positioning) on the fixtures' meshes.  CPU only."""
import math

import numpy as np
import pytest

import replay
from tissueorigami_amd import tissue


def rz(th):
    c, s = math.cos(th), math.sin(th)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


# ---- polar factor and growth rotation

def test_polar_factor_matches_numpy_svd(orc):
    rng = np.random.default_rng(5)
    for _ in range(200):
        F = np.eye(3) + 0.4 * rng.standard_normal((3, 3))
        if np.linalg.det(F) <= 0:
            continue
        U, S, Vt = np.linalg.svd(F)
        R = orc.polar_rotation(F)
        # F^T F squares the condition number of F
        assert np.abs(R - U @ Vt).max() < 1e-15 * (S[0] / S[2]) ** 2 + 1e-14
    assert orc.polar_rotation(np.diag([1.0, 1.0, -1.0])) is None


def _prism():
    ref = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    return ref


@pytest.mark.parametrize("th", [0.3, -0.8, 2.5])
def test_growth_rotation_of_a_turned_prism(orc, th):
    ref = _prism()
    S = np.diag([1.3, 0.8, 1.1])                   # a stretch, then a turn about z
    pos = ref @ (rz(th) @ S).T + np.array([2.0, -1.0, 0.5])
    g = orc.growth_rotation(ref, pos)
    assert g is not None
    # theta_z = atan2(R01, R00) = -th; GrowthStrainsRotMat = R_z(-th)^T = R_z(th)
    assert np.abs(g - rz(th)).max() < 1e-13


def test_small_turns_are_not_rotations(orc):
    ref = _prism()
    assert orc.growth_rotation(ref, ref @ rz(0.015).T) is None
    assert orc.growth_rotation(ref, ref @ np.diag([1.2, 0.9, 1.0]).T) is None


# ---- compatible grid corners

def _grid(rates, angles):
    """2 x 2 grid from per-corner rates / angles in corner order (0,0) (1,0) (0,1) (1,1)."""
    r = np.zeros((2, 2, 3)); a = np.zeros((2, 2))
    for c, (i, j) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1)]):
        r[i, j] = rates[c]; a[i, j] = angles[c]
    return r, a


def test_grid_corners_within_45_degrees(orc):
    rates = [(2.0, 1.0, 0.1), (3.0, 1.0, 0.2), (4.0, 1.0, 0.3), (5.0, 1.0, 0.4)]
    r, a = _grid(rates, [10.0, 80.0, 100.0, 170.0])
    g, an, el = orc.grid_compatible(r, a)
    d = np.degrees(an[0])
    # 80 -> -10 and 100 -> 10 turn by 90 degrees and swap x / y; 170 -> -10 by 180
    assert np.allclose(d, [10.0, -10.0, 10.0, -10.0], atol=1e-12)
    assert not el.any()
    assert np.array_equal(g[0, 0], rates[0]) and np.array_equal(g[0, 3], rates[3])
    assert np.array_equal(g[0, 1], [1.0, 3.0, 0.2]) and np.array_equal(g[0, 2], [1.0, 4.0, 0.3])


def test_grid_corners_eliminated_by_aspect_ratio(orc):
    rates = [(1.0, 1.05, 0.0), (2.0, 1.0, 0.0), (1.0, 1.0, 0.0), (1.0, 2.0, 0.0)]
    r, a = _grid(rates, [30.0, 40.0, 50.0, 160.0])
    g, an, el = orc.grid_compatible(r, a)
    assert el[0].tolist() == [True, False, True, False]
    # base: corner 1 (40 degrees); an eliminated corner keeps angle 0, and
    # corner 0 lies before the base, so it is not moved; 160 is 120 degrees
    # from the base: turned by -90 to 70 with x / y swapped
    d = np.degrees(an[0])
    assert d[0] == 0.0 and abs(d[1] - 40.0) < 1e-12 and abs(d[3] - 70.0) < 1e-12
    # corner 2 (eliminated) lies after the base: 0 - 40 is within 45 degrees, unchanged
    assert d[2] == 0.0
    assert np.array_equal(g[0, :3], np.array(rates[:3]))
    assert np.array_equal(g[0, 3], [2.0, 1.0, 0.0])


# ---- relative positions

def test_relative_positions_match_the_host_helper(orc):
    m = tissue.ellipse(9.0, 6.0, 1.0, 3.0, 3)
    rng = np.random.default_rng(2)
    u = m["nodes"] + 0.05 * rng.standard_normal(m["nodes"].shape)
    ex = np.zeros(len(u), np.uint8); ex[::5] = 1
    for sym in (False, True):
        a = orc.relative_positions(u, m["conn"], ex, sym)
        b = tissue.relative_positions(u, m["conn"], node_excluded=ex, symmetric_y=sym)
        assert np.array_equal(a, b)


def test_bounding_box_else_if_when_the_first_node_is_the_maximum(orc):
    """The first node only lowers the minimum: when it alone holds the largest
    x, the box of x = (5, 1, 3, 2, 1, 3) is [1, 3], not [1, 5]."""
    u = np.array([[5.0, 0, 0], [1.0, 1, 0], [3.0, 2, 0],
                  [2.0, 0, 1], [1.0, 1, 1], [3.0, 2, 1]])
    conn = np.array([[0, 1, 2, 3, 4, 5]])
    rel = orc.relative_positions(u, conn)
    cx = u[:, 0].mean()
    assert rel[0, 0] == (cx - 1.0) / (3.0 - 1.0)


# ---- growth step: closed forms

def _ellipse_prisms(m, n_layers):
    ne = len(m["conn"])
    place, _ = tissue.element_placement(m["conn"], m["node_placement"])
    rzv = tissue.initial_z_positions(m["nodes"], m["conn"], m["node_type"], m["node_placement"])
    return dict(type=np.zeros(ne, np.int32), place=place.astype(np.int32),
                flags=np.zeros(ne, np.uint8), rel_z=rzv, weights=None, explicit_ecm=False)


def _ident(n):
    return np.tile(np.eye(3).reshape(9), (n, 1))


@pytest.fixture(scope="module")
def mesh():
    return tissue.ellipse(8.0, 5.0, 1.0, 3.0, 3)


def test_unrotated_uniform_growth_is_diagonal(orc, mesh):
    m = mesh
    ne = len(m["conn"])
    P = _ellipse_prisms(m, 3)
    r, dt = (2e-5, 1e-5, -3e-6), 60.0
    fn = dict(type=1, t=(0.0, 1e9), apply=1, rate=r, angle=0.0)
    rel = orc.relative_positions(m["nodes"], m["conn"])
    Fg, inv = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P, [fn], rel, 0.0, dt, 1,
                              _ident(ne))
    D = np.diag([math.exp(x * 1.0 * dt) for x in r]).reshape(9)
    assert np.array_equal(Fg, np.tile(D, (ne, 1)))
    assert np.abs(inv.reshape(-1, 3, 3) @ Fg.reshape(-1, 3, 3) - np.eye(3)).max() < 1e-15
    # out of its time window, the function does nothing
    Fg2, _ = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P,
                             [dict(fn, t=(100.0, 200.0))], rel, 50.0, dt, 1, _ident(ne))
    assert np.array_equal(Fg2, _ident(ne))


@pytest.mark.parametrize("turn", [0.0, 0.4])
def test_uniform_growth_with_angle_in_a_turned_tissue(orc, mesh, turn):
    m = mesh
    ne = len(m["conn"])
    P = _ellipse_prisms(m, 3)
    r, dt, ang = (2e-5, -1e-5, 4e-6), 60.0, 0.6
    u = m["nodes"] @ rz(turn).T
    fn = dict(type=1, t=(0.0, 1e9), apply=1, rate=r, angle=ang)
    rel = orc.relative_positions(u, m["conn"])
    Fg, _ = orc.growth_step(m["conn"], m["refpos"], u, P, [fn], rel, 0.0, dt, 1, _ident(ne))
    D = np.diag([math.exp(x * dt) for x in r])
    G = rz(ang) @ D @ rz(ang).T
    # the lab-frame increment pulled back by the prism's turn: R^T G R
    G = rz(turn).T @ G @ rz(turn)
    assert np.abs(Fg - G.reshape(9)).max() < 1e-13


@pytest.mark.parametrize("interp", [0, 1])
def test_constant_grid_equals_uniform_growth(orc, mesh, interp):
    m = mesh
    ne = len(m["conn"])
    P = _ellipse_prisms(m, 3)
    r, dt, ang = np.array([3e-5, 1e-5, 2e-6]), 60.0, 25.0
    rel = orc.relative_positions(m["nodes"], m["conn"])
    grid = dict(type=3, t=(0.0, 1e9), apply=1, rates=np.tile(r, (5, 4, 1)),
                angles_deg=np.full((5, 4), ang))
    uni = dict(type=1, t=(0.0, 1e9), apply=1, rate=r, angle=math.radians(ang))
    Fa, _ = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P, [grid], rel, 0.0, dt,
                            interp, _ident(ne))
    Fb, _ = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P, [uni], rel, 0.0, dt,
                            interp, _ident(ne))
    if interp == 0:
        # nearest corner: no angle, the rates on the diagonal
        assert np.array_equal(Fa, np.tile(np.diag(np.exp(r * dt)).reshape(9), (ne, 1)))
    else:
        assert np.abs(Fa - Fb).max() < 1e-14


def test_weights_z_range_actin_and_ecm(orc, mesh):
    m = mesh
    ne = len(m["conn"])
    P = _ellipse_prisms(m, 3)
    rng = np.random.default_rng(1)
    P["type"] = rng.integers(0, 3, ne).astype(np.int32)
    P["weights"] = rng.uniform(0.2, 0.8, (ne, 2))
    P["flags"] = np.zeros(ne, np.uint8)
    P["flags"][::7] = 4                       # actin
    P["flags"][3::11] |= 1                    # ECM
    P["explicit_ecm"] = True
    r, dt = (2e-5, 1e-5, 3e-6), 100.0
    col = dict(type=1, t=(0.0, 1e9), apply=1, z=(0.3, 1.0), rate=r, angle=0.0)
    per = dict(type=1, t=(0.0, 1e9), apply=2, rate=(1e-5, 1e-5, 1e-5), angle=0.0)
    rel = orc.relative_positions(m["nodes"], m["conn"])
    Fg, _ = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P, [col, per], rel, 0.0, dt, 1,
                            _ident(ne))
    for e in range(ne):
        t, w, fl = P["type"][e], P["weights"][e], P["flags"][e]
        actin, ecm = bool(fl & 4), bool(fl & 1)
        gc = np.ones(3); gp = np.ones(3)
        if not ecm:
            if P["rel_z"][e] >= 0.3 and t in (0, 2):
                wc = 1.0 if t == 0 else w[0]
                gc = np.array([math.exp((0.0 if actin else x) * wc * dt) for x in r])
            if t in (1, 2):
                wp = 1.0 if t == 1 else w[1]
                gp = np.array([math.exp((0.0 if actin else 1e-5) * wp * dt)] * 3)
        expect = gc * gp if t == 2 else (gc if t == 0 else gp)
        assert np.abs(Fg[e].reshape(3, 3) - np.diag(expect)).max() < 1e-15, e


def test_column_conservation_keeps_reference_volume(orc):
    m = tissue.ellipse(8.0, 5.0, 1.0, 3.0, 3)
    conn = m["conn"]
    ne = len(conn)
    P = _ellipse_prisms(m, 3)
    place = P["place"]
    H, L = tissue.tissue_height(m["nodes"], conn, m["node_type"], m["node_placement"])
    assert L == 3
    st = tissue.element_stacks(conn, P["type"], place, L)
    cols = tissue.conserved_columns(st, P["type"], place, np.zeros(ne, bool), np.zeros(ne, bool), L)
    assert cols.shape == (np.sum(place == 1), 3) and (cols >= 0).all()
    rel = orc.relative_positions(m["nodes"], conn)
    grid = dict(type=3, t=(0.0, 1e9), apply=1, z=(0.0, 1.0),
                rates=np.random.default_rng(3).uniform(-3e-5, 5e-5, (4, 4, 3)),
                angles_deg=np.random.default_rng(4).uniform(0, 180, (4, 4)))
    Fg0 = np.tile(np.diag([1.1, 0.95, 1.02]).reshape(9), (ne, 1))
    Fg, _ = orc.growth_step(conn, m["refpos"], m["nodes"], P, [grid], rel, 0.0, 600.0, 1, Fg0,
                            columns=cols)
    V = orc.reference_volumes(m["refpos"])
    det = np.linalg.det(Fg.reshape(-1, 3, 3))
    for c in cols:
        assert abs(np.sum(det[c] * V[c]) - np.sum(V[c])) < 1e-12 * np.sum(V[c])
    # without conservation the same growth changes the column volumes
    Fn, _ = orc.growth_step(conn, m["refpos"], m["nodes"], P, [grid], rel, 0.0, 600.0, 1, Fg0)
    assert abs(np.linalg.det(Fn.reshape(-1, 3, 3))[cols[0]] @ V[cols[0]] - V[cols[0]].sum()) > 1e-6


def test_ablated_prisms_keep_their_growth(orc, mesh):
    m = mesh
    ne = len(m["conn"])
    P = _ellipse_prisms(m, 3)
    ab = np.zeros(ne, np.uint8); ab[[0, 5]] = 1
    fn = dict(type=1, t=(0.0, 1e9), apply=1, rate=(1e-5, 1e-5, 1e-5), angle=0.3)
    rel = orc.relative_positions(m["nodes"], m["conn"])
    Fg, _ = orc.growth_step(m["conn"], m["refpos"], m["nodes"], P, [fn], rel, 0.0, 60.0, 1,
                            _ident(ne), ablated=ab)
    assert np.array_equal(Fg[[0, 5]], _ident(2)) and not np.array_equal(Fg[1], _ident(1)[0])


# ---- the reference's own output

def oracle_frame(orc, d, k, P, fns, ex):
    """F_g of frame k from frame k - 1 by the oracle's growth step at t = k dt."""
    conn = d["conn"]
    ne = len(conn)
    u = replay.aligned_start(d, k)
    rel = orc.relative_positions(u, conn, ex["excluded"], ex["symmetric_y"], ex["apical_of"])
    if ex["pinned"]:  # pinned on the first step's positions (pin update at t = 0)
        rel = orc.relative_positions(replay.aligned_start(d, 0), conn, ex["excluded"],
                                     ex["symmetric_y"], ex["apical_of"])
    Fprev = _ident(ne) if k == 0 else d["Fg"][k - 1]
    return orc.growth_step(conn, d["refpos"], u, P, fns, rel, k * float(d["dt"]),
                           float(d["dt"]), ex["interp"], Fprev, columns=ex["columns"])[0]


def test_fixture_setup_helpers():
    for run, n_cols in (("run07007", 16), ("run07008", 46)):
        d = replay.load(run)
        P, fns, ex = replay.growth_setup(d)
        _, L = tissue.tissue_height(d["mesh_nodes"], d["conn"], d["mesh_node_info"][:, 1],
                                    d["mesh_node_info"][:, 0])
        assert L == 3
        # every ECM prism is basal and takes its column's apical position
        assert (P["place"][ex["ecm"]] == 0).all() and (ex["apical_of"][ex["ecm"]] >= 0).all()
        assert (ex["apical_of"] >= 0).sum() == n_cols
        if ex["columns"] is not None:
            assert ex["columns"].shape == (n_cols, 3)
            assert ((ex["columns"] >= 0).sum(axis=1) == 1).all()   # the middle prism only


def test_run07008_uniform_growth_matches_save_growth(orc):
    """Uniform growth at 45 degrees on the wing disc: every prism that is not
    ECM (whose remodelling lies outside this path) against Save_Growth, for
    the five growing steps; frame 4 has prisms whose growth is rotated."""
    d = replay.load("run07008")
    P, fns, ex = replay.growth_setup(d)
    keep = ~ex["ecm"]
    n_rot = 0
    for k in range(6):
        Fg = oracle_frame(orc, d, k, P, fns, ex)
        exp = d["Fg"][k]
        u = replay.aligned_start(d, k)
        turned = np.array([orc.growth_rotation(d["refpos"][e], u[d["conn"][e]]) is not None
                           for e in range(len(exp))])
        n_rot += int((turned & keep & ~ex["actin"]).sum())
        err = np.abs(Fg - exp).max(axis=1)
        # the saved frames hold 10 significant digits: the turn angle of a
        # rotated prism, and so its increment, carries their rounding
        assert err[keep & ~turned].max() <= 1e-15, k
        if (keep & turned).any():
            assert err[keep & turned].max() <= 1e-9, k
    assert n_rot > 0
    # frame 5: the function has ended (t >= 18000 s); a rotated prism still
    # turns its identity increments, which moves F_g in the last bits only
    assert np.abs(d["Fg"][5] - d["Fg"][4])[keep].max() < 1e-14


def test_run07007_grid_growth_with_column_conservation_matches_save_growth(orc):
    """Two grid functions (z from 120 s, x-y from 60 s, basal half of the
    tissue) with nearest-corner lookup on pinned positions and column volume
    conservation: every prism that is not ECM against Save_Growth."""
    d = replay.load("run07007")
    P, fns, ex = replay.growth_setup(d)
    keep = ~ex["ecm"]
    grew = 0
    for k in range(7):
        Fg = oracle_frame(orc, d, k, P, fns, ex)
        exp = d["Fg"][k]
        err = np.abs(Fg - exp)[keep].max()
        assert err <= 1e-13, (k, err)
        prev = _ident(len(exp)) if k == 0 else d["Fg"][k - 1]
        grew += int((np.abs(exp - prev)[keep].max(axis=1) > 1e-12).sum())
    assert grew >= 10
