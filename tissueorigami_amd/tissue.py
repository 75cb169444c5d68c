"""Host-side tissue bookkeeping and synthetic prism meshes.

Element placement and type follow the reference's rules so that the exposed
surface areas (the node "viscosity surfaces" that scale external drag) are the
ones the reference would compute:
  ShapeBase::setTissuePlacement  TissueFolding/SourceCode/ShapeBase.cpp:905-980
  ShapeBase::setTissueType       TissueFolding/SourceCode/ShapeBase.cpp:1001-1037
Node placement codes: 0 basal, 1 apical, 2 mid-line, 3 lateral, 4 apical bulk.
Node type codes: 0 columnar, 1 peripodial, 2 linker.

The synthetic generators lay nodes out layer by layer (all nodes of the basal
layer first), the numbering the reference's own meshes use (for example
ToolBox/MeshGeneration/2DEllipse/smallRectangle.mesh), and order every prism
as the reference expects after Prism::checkRotationConsistency3D
(Prism.cpp:369-417): basal triangle 0,1,2 counter-clockwise seen from the
apical side, apical triangle 3,4,5 above it.
"""
import numpy as np


def element_placement(conn, node_placement):
    """tissuePlacement and spansWholeTissue per element (ShapeBase.cpp:905-980)."""
    npl = np.asarray(node_placement)[np.asarray(conn)]
    has_ap = (npl == 1).any(1)
    has_bulk = (npl == 4).any(1)
    has_bas = (npl == 0).any(1)
    has_lat = (npl == 3).any(1)
    place = np.full(len(npl), 2, dtype=np.int32)
    spans = np.zeros(len(npl), dtype=np.uint8)
    place[~has_lat & ~has_ap & ~has_bas & has_bulk] = 4
    place[~has_lat & ~has_ap & has_bas] = 0
    place[~has_lat & has_ap & ~has_bas] = 1
    both = ~has_lat & has_ap & has_bas
    place[both] = 2
    spans[both] = 1
    place[has_lat] = 3
    return place, spans


def element_type(conn, node_type):
    """tissueType per element (ShapeBase.cpp:1001-1037)."""
    nt = np.asarray(node_type)[np.asarray(conn)]
    t = np.zeros(len(nt), dtype=np.int32)
    t[(nt == 1).any(1)] = 1
    t[(nt == 2).any(1)] = 2
    return t


def _extrude(xy, tris, thickness, n_layers, boundary):
    n2 = xy.shape[0]
    z = np.linspace(0.0, thickness, n_layers + 1)
    nodes = np.zeros(((n_layers + 1) * n2, 3))
    for k in range(n_layers + 1):
        nodes[k * n2:(k + 1) * n2, :2] = xy
        nodes[k * n2:(k + 1) * n2, 2] = z[k]
    placement = np.full((n_layers + 1) * n2, 2, dtype=np.int32)
    placement[:n2] = 0
    placement[n_layers * n2:] = 1
    circ = np.tile(boundary.astype(np.int32), n_layers + 1)
    conn = np.zeros((n_layers * len(tris), 6), dtype=np.int32)
    for k in range(n_layers):
        blk = conn[k * len(tris):(k + 1) * len(tris)]
        blk[:, :3] = tris + k * n2
        blk[:, 3:] = tris + (k + 1) * n2
    layer = np.repeat(np.arange(n_layers), len(tris)).astype(np.int32)
    refpos = nodes[conn].reshape(-1, 18).copy()
    return dict(nodes=nodes, conn=conn, refpos=refpos, node_placement=placement,
                node_type=np.zeros(len(nodes), dtype=np.int32), at_circumference=circ,
                layer=layer, n_layers=n_layers, n2d=n2)


def _ccw(xy, tris):
    a, b, c = xy[tris[:, 0]], xy[tris[:, 1]], xy[tris[:, 2]]
    cross = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    t = tris.copy()
    flip = cross < 0
    t[flip, 1], t[flip, 2] = tris[flip, 2], tris[flip, 1]
    return t


def rectangle(lx=10.0, ly=5.0, side=1.0, thickness=2.0, n_layers=2):
    """Regular right-triangle grid on an lx x ly rectangle, extruded in z."""
    nx = int(round(lx / side)); ny = int(round(ly / side))
    xs = np.linspace(0, lx, nx + 1); ys = np.linspace(0, ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    xy = np.stack([X.ravel(), Y.ravel()], 1)
    idx = np.arange((nx + 1) * (ny + 1)).reshape(ny + 1, nx + 1)
    tris = []
    for j in range(ny):
        for i in range(nx):
            p0, p1, p2, p3 = idx[j, i], idx[j, i + 1], idx[j + 1, i], idx[j + 1, i + 1]
            tris.append([p0, p1, p3])
            tris.append([p0, p3, p2])
    tris = _ccw(xy, np.array(tris, dtype=np.int32))
    boundary = (np.isclose(xy[:, 0], 0) | np.isclose(xy[:, 0], lx) |
                np.isclose(xy[:, 1], 0) | np.isclose(xy[:, 1], ly))
    m = _extrude(xy, tris, thickness, n_layers, boundary)
    m["kind"] = "rectangle"
    return m


def ellipse(a=100.0, b=70.0, side=2.0, thickness=20.0, n_layers=4):
    """Equilateral lattice clipped to an ellipse (x/a)^2 + (y/b)^2 <= 1, extruded in z.

    Rows of points at y = j h, h = side*sqrt(3)/2, odd rows shifted by side/2.
    Every lattice triangle whose three corners lie inside the ellipse is kept;
    unused points are dropped.  Deterministic for given sizes.
    """
    h = side * np.sqrt(3.0) / 2.0
    nj = int(np.floor(b / h))
    rows = []
    for j in range(-nj, nj + 1):
        y = j * h
        off = 0.5 * side if (j % 2) else 0.0
        ni = int(np.floor(a / side)) + 1
        xs = off + side * np.arange(-ni, ni + 1)
        rows.append((j, y, xs))
    pid = {}
    pts = []
    for j, y, xs in rows:
        for i, x in enumerate(xs):
            if (x / a) ** 2 + (y / b) ** 2 <= 1.0:
                pid[(j, i)] = len(pts)
                pts.append((x, y))
    xy = np.array(pts)
    ni_of = {j: len(xs) for j, _, xs in rows}
    tris = []
    for j, y, xs in rows:
        if j == nj:
            continue
        odd = (j % 2) != 0
        for i in range(len(xs)):
            # next row index whose x is +side/2 relative to (j, i)
            ir = i if not odd else i + 1
            il = ir - 1
            cand = [((j, i), (j, i + 1), (j + 1, ir)), ((j, i), (j + 1, ir), (j + 1, il))]
            for t in cand:
                if all(k in pid for k in t) and all(0 <= k[1] < ni_of[k[0]] for k in t):
                    tris.append([pid[k] for k in t])
    tris = np.array(tris, dtype=np.int32)
    used = np.unique(tris)
    remap = -np.ones(len(xy), dtype=np.int64)
    remap[used] = np.arange(len(used))
    xy = xy[used]
    tris = remap[tris].astype(np.int32)
    tris = _ccw(xy, tris)
    # boundary points: edges used by one triangle only
    e = np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), 1)
    key = e[:, 0].astype(np.int64) * len(xy) + e[:, 1]
    uniq, cnt = np.unique(key, return_counts=True)
    bnd_edges = uniq[cnt == 1]
    boundary = np.zeros(len(xy), dtype=bool)
    boundary[(bnd_edges // len(xy)).astype(np.int64)] = True
    boundary[(bnd_edges % len(xy)).astype(np.int64)] = True
    m = _extrude(xy, tris, thickness, n_layers, boundary)
    m["kind"] = "ellipse"
    m["axes"] = (a, b)
    return m


def gaussian_random_field(points, mean, sd, corr_len, seed, n_modes=256):
    """Stationary Gaussian random field with a squared-exponential covariance
    sd^2 exp(-r^2 / (2 corr_len^2)), sampled by random Fourier features."""
    rng = np.random.default_rng(seed)
    d = points.shape[1]
    k = rng.normal(0.0, 1.0 / corr_len, size=(n_modes, d))
    phase = rng.uniform(0, 2 * np.pi, size=n_modes)
    f = np.cos(points @ k.T + phase) @ np.ones(n_modes)
    return mean + sd * np.sqrt(2.0 / n_modes) * f


def growth_increment(rate_per_hour_xy, dt):
    """In-plane isotropic growth increment diag(exp(r dt), exp(r dt), 1) with r
    in 1/s, as ShapeBase::calculateFgFromRates (ShapeBase.cpp:471-505) builds it."""
    r = np.asarray(rate_per_hour_xy, dtype=np.float64) / 3600.0
    g = np.exp(r * dt)
    inc = np.zeros((len(g), 3, 3))
    inc[:, 0, 0] = g
    inc[:, 1, 1] = g
    inc[:, 2, 2] = 1.0
    return inc


def ellipse_for_prisms(target_prisms, n_layers, aspect=0.7, side=1.0, layer_thickness=2.5):
    """Ellipse mesh (semi-axes a and aspect*a, lattice side `side`) sized so
    that n_layers prism layers hold about target_prisms prisms: an equilateral
    lattice of side s has about pi a b / (sqrt(3)/4 s^2) triangles."""
    n_tri = target_prisms / n_layers
    a = np.sqrt(n_tri * np.sqrt(3.0) / 4.0 * side * side / (np.pi * aspect))
    return ellipse(a, aspect * a, side, layer_thickness * n_layers, n_layers)


def batch_ellipses(n=64, prism_range=(2000, 40000), layer_range=(2, 6), E_range=(100.0, 2000.0),
                   seed=4):
    """Independent elliptical tissues for a batched run: prism counts
    log-uniform in prism_range, layer counts uniform in layer_range, aspect
    ratio uniform in [0.6, 0.9], one Young's modulus per tissue uniform in
    E_range, nu = 0.3, internal viscosity 10 Pa s, external drag 16000
    (apical) / 10 (basal), in-plane growth from a per-tissue seeded Gaussian
    random field (mean 0.02/h, sd 0.01/h, correlation length 15 um).
    Returns one dict per tissue (mesh, per-prism E/nu/eta, per-node ext_visc,
    growth rate per prism in 1/h, surface flags, seeds)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(prism_range[0]), np.log(prism_range[1])
    out = []
    for t in range(n):
        target = float(np.exp(rng.uniform(lo, hi)))
        layers = int(rng.integers(layer_range[0], layer_range[1] + 1))
        aspect = float(rng.uniform(0.6, 0.9))
        E = float(rng.uniform(*E_range))
        m = ellipse_for_prisms(target, layers, aspect)
        ne, nn = len(m["conn"]), len(m["nodes"])
        ext = np.zeros((nn, 3))
        ext[m["node_placement"] == 1] = 16000.0
        ext[m["node_placement"] == 0] = 10.0
        cent = m["nodes"][m["conn"]].mean(1)[:, :2]
        gseed = seed * 1000 + t
        rate = gaussian_random_field(cent, 0.02, 0.01, 15.0, gseed)
        place, spans = element_placement(m["conn"], m["node_placement"])
        etype = element_type(m["conn"], m["node_type"])
        out.append(dict(mesh=m, E=np.full(ne, E), nu=np.full(ne, 0.3), eta=np.full(ne, 10.0),
                        ext_visc=ext, rate=rate, place=place, spans=spans, etype=etype,
                        layers=layers, growth_seed=gseed, target_prisms=target))
    return out


def block_count(conn, n_nodes):
    """Stored 3x3 blocks of the Jacobian of a mesh: the distinct (node, node)
    pairs that share a prism, the diagonal included."""
    c = np.asarray(conn, dtype=np.int64).reshape(-1, 6)
    a = np.repeat(c, 6, axis=1).reshape(-1)
    b = np.tile(c, (1, 6)).reshape(-1)
    return int(len(np.unique(a * int(n_nodes) + b)))


def rates_xyz(rate_per_hour):
    """Per-prism growth rates (x, y, z) in 1/s for in-plane isotropic growth."""
    r = np.zeros((len(rate_per_hour), 3))
    r[:, 0] = r[:, 1] = np.asarray(rate_per_hour) / 3600.0
    return r


def facing_slabs(m, gap, shift=(0.0, 0.0), upper=None):
    """Two extruded meshes whose apical surfaces face each other.

    The upper mesh (a copy of m unless given) is mirrored in z so that its
    apical surface lies `gap` above m's, and shifted by `shift` in (x, y); its
    prisms are renumbered (nodes 1 <-> 2 and 4 <-> 5) so that each basal
    triangle is again counter-clockwise seen from its own apical side.  The
    two meshes share no element, so their apical nodes are candidate packing
    couples (Simulation::detectPacingNodes, Simulation.cpp:7997-8078) once they
    are closer than the detection threshold."""
    u = m if upper is None else upper
    nodes = m["nodes"]
    nn = len(nodes)
    top = nodes[:, 2].max()
    mirror = u["nodes"].copy()
    mirror[:, 2] = top + gap + (u["nodes"][:, 2].max() - u["nodes"][:, 2])
    mirror[:, 0] += shift[0]
    mirror[:, 1] += shift[1]
    conn_b = u["conn"][:, [0, 2, 1, 3, 5, 4]] + nn
    out = dict(m)
    out["nodes"] = np.concatenate([nodes, mirror])
    out["conn"] = np.concatenate([m["conn"], conn_b]).astype(np.int32)
    out["refpos"] = out["nodes"][out["conn"]].reshape(-1, 18).copy()
    for key in ("node_placement", "node_type", "at_circumference", "layer"):
        out[key] = np.concatenate([m[key], u[key]])
    out["n_lower_nodes"] = nn
    out["kind"] = m.get("kind", "mesh") + "-facing-slabs"
    return out


def stiffness_applicable(where, tissue_type, placement, is_ecm=None, ecm_at_circumference=None,
                         at_basal_border_of_ecm=None):
    """Per prism: whether a Young's modulus series applies to it
    (YoungsModulusModifier::IsItApplicable, YoungsModulusModifier.cpp:74-100).

    where: the seven flags of TimeSeriesPhysicalProperties::DefineWhereToApply
    (TimeSeriesPhysicalProperties.cpp:112-120): columnar layer, peripodial
    membrane, apical, mid and basal layers, peripodial ECM, columnar ECM.
    tissue_type: 0 columnar, 1 peripodial; placement: ShapeBase::tissuePlacement
    (0 basal, 1 apical, 2 middle, 4 lateral)."""
    cl, pm, ap, mid, bas, _pecm, cecm = [bool(w) for w in where]
    tt = np.asarray(tissue_type).reshape(-1)
    pl = np.asarray(placement).reshape(-1)
    n = tt.size
    z = np.zeros(n, bool)
    ecm = z if is_ecm is None else np.asarray(is_ecm, bool).reshape(-1)
    circ = z if ecm_at_circumference is None else np.asarray(ecm_at_circumference, bool).reshape(-1)
    border = z if at_basal_border_of_ecm is None else np.asarray(at_basal_border_of_ecm, bool).reshape(-1)
    layer = (cl & (tt == 0)) | (pm & (tt == 1))
    on_ecm = (cecm & (pl == 0)) | (cecm & circ & (pl != 0))
    on_cells = (ap & ((pl == 1) | (pl == 4))) | (mid & (pl == 2)) | (bas & ((pl == 0) | border))
    return (layer & np.where(ecm, on_ecm, on_cells)).astype(np.uint8)


def relative_positions(u, conn, apical_of=None, node_excluded=None, symmetric_y=False):
    """Per prism: its centre's relative (x, y) position in the tissue's bounding
    box (ShapeBase::calculateRelativePosInBoundingBox, ShapeBase.cpp:2091-2095,
    with Simulation::calculateBoundingBox, Simulation.cpp:8388-8428).

    apical_of (optional, per prism): the apical prism of its column, whose
    position every prism of the column takes
    (Simulation::updateRelativePositionsToApicalPositioning,
    Simulation.cpp:6147-6167); -1 keeps the prism's own.  node_excluded
    (optional, per node): nodes left out of the bounding box (those all of
    whose owners mimic ECM)."""
    u = np.asarray(u, np.float64).reshape(-1, 3)
    conn = np.asarray(conn).reshape(-1, 6)
    keep = np.ones(len(u), bool) if node_excluded is None else ~np.asarray(node_excluded, bool)
    # the reference's loop: a coordinate lowers the minimum, or else raises the
    # maximum, from the initial (1e5, -1e5)
    lo = np.full(3, 100000.0)
    hi = np.full(3, -100000.0)
    for p in u[keep]:
        for i in range(3):
            if p[i] < lo[i]:
                lo[i] = p[i]
            elif p[i] > hi[i]:
                hi[i] = p[i]
    if symmetric_y:
        lo[1] = -hi[1]
    size = hi - lo
    c = np.zeros((len(conn), 3))
    for a in range(6):
        c += u[conn[:, a]]
    c /= 6
    rel = np.stack([(c[:, 0] - lo[0]) / size[0], (c[:, 1] - lo[1]) / size[1]], axis=1)
    if apical_of is not None:
        ap = np.asarray(apical_of).reshape(-1)
        sel = ap >= 0
        rel[sel] = rel[ap[sel]]
    return rel


# ---- per-prism data of the growth functions (the reference's set-up, on the host)

def tissue_height(u, conn, node_type, node_placement):
    """Simulation::calculateTissueHeight (Simulation.cpp:3275-3330) for a prism
    mesh: from the first basal columnar node, up through the first prism that
    has the current node as a basal node, adding that prism's height
    (Prism::getElementHeight, Prism.cpp:1445-1450: |node 0 - node 3|) until an
    apical node.  Returns (TissueHeight, TissueHeightDiscretisationLayers)."""
    u = np.asarray(u, np.float64).reshape(-1, 3)
    conn = np.asarray(conn).reshape(-1, 6)
    nt = np.asarray(node_type).reshape(-1)
    npl = np.asarray(node_placement).reshape(-1)
    basal = np.where((nt == 0) & (npl == 0))[0]
    if basal.size == 0:
        raise ValueError("no basal columnar node")
    node, h, layers = int(basal[0]), 0.0, 0
    while npl[node] != 1:
        owners = np.where((conn[:, :3] == node).any(axis=1))[0]
        if owners.size == 0:
            raise ValueError(f"no prism has node {node} as a basal node")
        e = int(owners[0])
        d = u[conn[e, 0]] - u[conn[e, 3]]
        h += (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) ** 0.5
        node = int(conn[e, 3 + list(conn[e, :3]).index(node)])
        layers += 1
    return h, layers


def initial_z_positions(u, conn, node_type, node_placement):
    """ShapeBase::initialRelativePositionInZ per prism
    (Simulation::assignInitialZPositions, Simulation.cpp:3365-3369, after the
    bounding box of all nodes; ShapeBase::setInitialZPosition,
    ShapeBase.cpp:812-828): 1 - (centre z - lowest z) / tissue height,
    clamped to [0, 1]; 0 apical, 1 basal."""
    u = np.asarray(u, np.float64).reshape(-1, 3)
    conn = np.asarray(conn).reshape(-1, 6)
    H, _ = tissue_height(u, conn, node_type, node_placement)
    zmin = 100000.0
    for z in u[:, 2]:
        if z < zmin:
            zmin = z
    out = np.empty(len(conn))
    for e in range(len(conn)):
        c = 0.0
        for a in range(6):
            c += u[conn[e, a], 2]
        c /= 6
        r = 1.0 - ((c - zmin) / H)
        out[e] = min(max(r, 0.0), 1.0)
    return out


def element_stacks(conn, elem_type, elem_place, layers, ablated=None):
    """Prism::constructElementStackList (Prism.cpp:1605-1647) from every
    columnar basal prism that is not ablated (Simulation::fillInElementColumnLists,
    Simulation.cpp:7657-7674): the stack, basal first, each next prism the
    first other prism whose basal nodes hold the three apical nodes reached so
    far.  Returns {prism: stack} for every prism of a stack."""
    conn = np.asarray(conn).reshape(-1, 6)
    et = np.asarray(elem_type).reshape(-1)
    ep = np.asarray(elem_place).reshape(-1)
    abl = np.zeros(len(conn), bool) if ablated is None else np.asarray(ablated, bool)
    basal_sets = [set(conn[e, :3].tolist()) for e in range(len(conn))]
    stacks = {}
    for b in range(len(conn)):
        if abl[b] or et[b] != 0 or ep[b] != 0:
            continue
        stack = [b]
        cur, top = b, [int(x) for x in conn[b, 3:]]
        for _ in range(1, layers):
            nxt = next((e for e in range(len(conn))
                        if e != cur and all(n in basal_sets[e] for n in top)), None)
            if nxt is None:
                break
            stack.append(nxt)
            top = [int(conn[nxt, 3 + list(conn[nxt, :3]).index(n)]) for n in top]
            cur = nxt
        for e in stack:
            stacks[e] = stack
    return stacks


def apical_positioning(stacks, elem_type, elem_place, n_prisms, layers):
    """Simulation::updateRelativePositionsToApicalPositioning
    (Simulation.cpp:6147-6167) as a per-prism map: every columnar apical prism
    gives its relative position to the first layers - 2 prisms of its stack
    (basal first; the prism just below the apical one keeps its own).  -1:
    the prism keeps its own."""
    et = np.asarray(elem_type).reshape(-1)
    ep = np.asarray(elem_place).reshape(-1)
    out = np.full(n_prisms, -1, np.int32)
    for e in range(n_prisms):
        if ep[e] == 1 and et[e] == 0 and e in stacks:
            for m in stacks[e][:max(layers - 2, 0)]:
                out[m] = e
    return out


def conserved_columns(stacks, elem_type, elem_place, is_ecm, is_actin, layers):
    """The prisms whose column volume Simulation::conserveColumnVolume
    (Simulation.cpp:6518-6600) conserves: per columnar apical prism, its
    stack's columnar prisms that mimic neither ECM nor actin.  Returns
    (n_columns, layers) prism ids, -1 padded."""
    et = np.asarray(elem_type).reshape(-1)
    ep = np.asarray(elem_place).reshape(-1)
    ecm = np.asarray(is_ecm, bool).reshape(-1)
    act = np.asarray(is_actin, bool).reshape(-1)
    cols = []
    for e in range(len(et)):
        if et[e] == 0 and ep[e] == 1 and e in stacks:
            m = [x for x in stacks[e][:layers] if et[x] == 0 and not ecm[x] and not act[x]]
            cols.append(m + [-1] * (layers - len(m)))
    return np.array(cols, np.int32).reshape(-1, layers)


def ecm_only_nodes(conn, n_nodes, is_ecm):
    """Node::allOwnersECMMimicing (Simulation::setUpECMMimicingElements,
    Simulation.cpp:8866-8920): nodes every owner of which mimics ECM."""
    conn = np.asarray(conn).reshape(-1, 6)
    ecm = np.asarray(is_ecm, bool).reshape(-1)
    owned = np.zeros(n_nodes, bool)
    other = np.zeros(n_nodes, bool)
    owned[conn.reshape(-1)] = True
    other[conn[~ecm].reshape(-1)] = True
    return (owned & ~other).astype(np.uint8)


def uniform_growth_input(rate_per_hour, angle_deg):
    """The rates (1/s) and angle (radians) a UniformGrowthFunction holds after
    ModelInputObject::readGrowthType1 (ModelInputObject.cpp:474-501): each
    value read into a float, scaled by 1/3600 (the angle by pi/180) and stored
    back into the float."""
    r = [float(np.float32(float(np.float32(v)) * (1.0 / 3600.0))) for v in rate_per_hour]
    a = float(np.float32(float(np.float32(angle_deg)) * (np.pi / 180.0)))
    return np.array(r), a


def grid_growth_input(rates_per_hour):
    """Grid growth rates in 1/s as GridBasedGrowthFunction holds them
    (ModelInputObject.cpp:800-812: the float read times 1/3600 in double)."""
    return np.asarray(rates_per_hour, np.float64) * (1.0 / 3600.0)
