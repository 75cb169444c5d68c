"""Per-step replay of the reference's Pardiso regression runs run07007 and
run07008.

Builds, from the committed fixtures (tests/golden/run0700x.npz), the inputs of
the NR step that produced saved frame k+1 from saved frame k (k = -1 starts
from the mesh and produces frame 0; frame j is the state after time step
j+1):
  * u_n: the start positions after the rigid pre-step alignment
    Simulation::alignTissueDVToXPositive (Simulation.cpp:8230-8272, with the
    ShapeBase rotation helpers calculateRotationAngleSinCos /
    calculateRotationAxis / constructRotationMatrix /
    rotateVectorByRotationMatrix), about the tips Simulation::assignTips
    (Simulation.cpp:8165-8188) finds on the mesh;
  * F_g, effective E (getYoungModulus, ShapeBase.cpp:179-181), internal and
    external viscosities of the produced frame;
  * the run's Poisson ratio, and 0 for ECM elements (ShapeBase::setECMMimicing,
    ShapeBase.cpp:988-995);
  * node viscosity surfaces from the aligned start positions;
  * fixed DOF: with symmetricInY, y of every node within 0.1 of y = 0
    (Simulation::setupYsymmetricity, Simulation.cpp:9557-9589);
    checkForZeroExternalViscosity (Simulation.cpp:4126-4236) fixes nothing in
    either run, since some node has external viscosity in each of x, y and z;
  * slave/master pairs: with bindCircumferenceXYToBasal, the pairs
    Simulation::bindCircumferenceXY builds (Simulation.cpp:1344-1436) on the
    circumference flags left by clearCircumferenceDataFromSymmetricityLine
    (Simulation.cpp:3057-3130).  run07007 has neither symmetry nor binding.
"""
import os

import numpy as np

from tissueorigami_amd import tissue

HERE = os.path.dirname(os.path.abspath(__file__))
SYM_LIM = 0.1          # the y window of the symmetry line (Simulation.cpp:9558, 3093)


def load(run="run07007"):
    return np.load(os.path.join(HERE, "golden", run + ".npz"))


def assign_tips(nodes, node_info):
    """Simulation::assignTips (Simulation.cpp:8165-8188): the basal (placement
    0) columnar (type 0) nodes of smallest and largest x, first index on
    ties; returns (dorsal, ventral)."""
    xmin, xmax = 10000.0, -10000.0
    dorsal = ventral = None
    for i in range(len(nodes)):
        if node_info[i, 0] == 0 and node_info[i, 1] == 0:
            if nodes[i, 0] < xmin:
                dorsal, xmin = i, nodes[i, 0]
            if nodes[i, 0] > xmax:
                ventral, xmax = i, nodes[i, 0]
    return dorsal, ventral


def _normalise(v):
    mag2 = float(v @ v)
    if abs(mag2) > 1e-14 and abs(mag2 - 1.0) > 1e-14:
        return v / mag2 ** 0.5
    return v


def align_dv_to_x(P, dorsal, ventral):
    P = P.copy()
    u = _normalise(P[ventral] - P[dorsal])
    v = np.array([1.0, 0.0, 0.0])
    c = float(u @ v)
    if c > 1.0:
        c, s = 1.0, 0.0
    elif c < -1.0:
        c, s = -1.0, 0.0
    else:
        s = np.sin(np.arccos(c))
    if c > -0.99998:
        ax = _normalise(np.cross(u, v))
    else:
        ax = np.array([1.0, 0.0, 0.0])
    R = np.array([
        [c + ax[0] * ax[0] * (1 - c), ax[0] * ax[1] * (1 - c) - ax[2] * s, ax[0] * ax[2] * (1 - c) + ax[1] * s],
        [ax[1] * ax[0] * (1 - c) + ax[2] * s, c + ax[1] * ax[1] * (1 - c), ax[1] * ax[2] * (1 - c) - ax[0] * s],
        [ax[2] * ax[0] * (1 - c) - ax[1] * s, ax[2] * ax[1] * (1 - c) + ax[0] * s, c + ax[2] * ax[2] * (1 - c)]])
    rel = P[1:] - P[0]
    P[1:] = rel @ R.T + P[0]
    return P


def symmetric_y_circumference(nodes, circ):
    """clearCircumferenceDataFromSymmetricityLine for y symmetry alone
    (Simulation.cpp:3083-3130): a node within SYM_LIM of y = 0 leaves the
    circumference unless it is within SYM_LIM of the largest or smallest x."""
    circ = np.array(circ, copy=True)
    xpos, xneg = nodes[:, 0].max(), nodes[:, 0].min()
    for i in range(len(nodes)):
        x, y = nodes[i, 0], nodes[i, 1]
        if -SYM_LIM < y < SYM_LIM:
            tip = (xpos - SYM_LIM < x < xpos + SYM_LIM) or (xneg - SYM_LIM < x < xneg + SYM_LIM)
            if not tip:
                circ[i] = 0
    return circ


def symmetric_y_fixing(nodes):
    """setupYsymmetricity (Simulation.cpp:9557-9589): y of every node strictly
    within SYM_LIM of y = 0 is fixed; (n, 3) flags.  A node below -SYM_LIM
    would be ablated, which neither run has."""
    assert not np.any(nodes[:, 1] <= -SYM_LIM), "ablated nodes are not replayed"
    fixed = np.zeros((len(nodes), 3), np.uint8)
    fixed[(nodes[:, 1] < SYM_LIM) & (nodes[:, 1] > -SYM_LIM), 1] = 1
    return fixed


def bind_circumference_xy(conn, placement, circ, fixed):
    """Simulation::bindCircumferenceXY (Simulation.cpp:1344-1436) without a
    peripodial membrane: every basal circumference node, in id order, is the
    master of the x and y DOF of the nodes above it, found by walking up
    through the prism whose basal face holds the current node
    (IsThisNodeMyBasal / getCorrecpondingApical) until an apical node is
    reached.  A fixed master fixes its slaves in that direction, and a fixed
    slave DOF is not bound.  Returns the pairs (slave DOF, master DOF) in the
    order the reference appends them, and updates `fixed` in place."""
    node_elems = [[] for _ in range(len(placement))]
    for e, c in enumerate(conn):
        for a in c:
            node_elems[a].append(e)
    pairs = []
    for m in range(len(placement)):
        if placement[m] != 0 or not circ[m]:
            continue
        cur, top, steps = m, False, 0
        while not top and steps < 10:
            steps += 1
            for e in node_elems[cur]:
                c = [int(v) for v in conn[e]]
                if cur in c[:3]:
                    s = c[c.index(cur) + 3]
                    for d in (0, 1):
                        if fixed[m, d]:
                            fixed[s, d] = 1
                    for d in (0, 1):
                        if not fixed[s, d]:
                            pairs.append((3 * s + d, 3 * m + d))
                    cur = s
                    break
            if placement[cur] == 1:
                top = True
    return np.array(pairs, np.int32).reshape(-1, 2)


def saved_binding(pairs, n_nodes):
    """What Simulation::writeNodeBinding (Simulation.cpp:8546-8565) writes for
    a pair list: node by node, x y z, the DOF whose slaveTo is above 0 (so the
    slaves of node 0 are left out)."""
    slave_to = -np.ones((n_nodes, 3), np.int64)
    for s, m in pairs:
        slave_to[s // 3, s % 3] = m // 3
    out = [(3 * i + d, 3 * slave_to[i, d] + d) for i in range(n_nodes) for d in range(3)
           if slave_to[i, d] > 0]
    return np.array(out, np.int32).reshape(-1, 2)


def constraints(d):
    """Fixed DOF flags (3n,) and slave/master pairs of a run."""
    nodes, info = d["mesh_nodes"], d["mesh_node_info"]
    fixed = symmetric_y_fixing(nodes) if int(d["symmetric_y"]) else \
        np.zeros((len(nodes), 3), np.uint8)
    circ = symmetric_y_circumference(nodes, info[:, 2]) if int(d["symmetric_y"]) else info[:, 2]
    pairs = np.zeros((0, 2), np.int32)
    if int(d["bind_circumference"]):
        pairs = bind_circumference_xy(d["conn"], info[:, 0], circ, fixed)
    return fixed.reshape(-1), pairs, circ


def step_inputs(d, k, orc):
    """Inputs of the step that produced frame k+1 (k = -1 .. n_frames-2)."""
    conn = d["conn"]
    ne = len(conn)
    nn = d["frames"].shape[1]
    place, spans = tissue.element_placement(conn, d["node_info"][:, 0])
    etype = tissue.element_type(conn, d["node_info"][:, 1])
    nu = np.full(ne, float(d["nu"]))
    nu[d["ecm"]] = 0.0
    dorsal, ventral = assign_tips(d["mesh_nodes"], d["mesh_node_info"])
    start = d["mesh_nodes"] if k < 0 else d["frames"][k]
    un = align_dv_to_x(start, dorsal, ventral).reshape(-1)
    invFg = orc.inv3_many(d["Fg"][k + 1])
    lam, mu = orc.lame(d["E"][k + 1], nu)
    A = orc.visc_surfaces(nn, conn, un, place, etype, spans)
    fixed, pairs, _ = constraints(d)
    out = dict(conn=conn, refpos=d["refpos"], invFg=invFg, lam=lam, mu=mu,
               E=d["E"][k + 1], nu=nu, eta=d["eta"][k + 1], ext_visc=d["ext_visc"][k + 1],
               visc_surf=A, un=un, dt=float(d["dt"]), expected=d["frames"][k + 1].reshape(-1),
               forces=d["forces"][k + 1].reshape(-1), fixed=fixed, pairs=pairs)
    if "du_log" in d.files:
        row = d["du_log"][k + 1]
        out["du_log"] = row[~np.isnan(row)]
    return out


def growth_setup(d):
    """Per-prism data, functions and options of a regression run, from its
    mesh and model input (tests/golden)."""
    conn, info, nodes = d["conn"], d["mesh_node_info"], d["mesh_nodes"]
    ne = len(conn)
    place, _ = tissue.element_placement(conn, info[:, 0])
    etype = tissue.element_type(conn, info[:, 1])
    ecm = np.zeros(ne, bool); ecm[d["ecm"]] = True
    act = np.zeros(ne, bool); act[d["actin"]] = True
    flags = (ecm * 1 + act * 4).astype(np.uint8)
    _, L = tissue.tissue_height(nodes, conn, info[:, 1], info[:, 0])
    stacks = tissue.element_stacks(conn, etype, place, L)
    P = dict(type=etype.astype(np.int32), place=place.astype(np.int32), flags=flags,
             rel_z=tissue.initial_z_positions(nodes, conn, info[:, 1], info[:, 0]),
             weights=None, explicit_ecm=bool(d["explicit_ecm"]))
    fns = []
    for k, t in enumerate(d["growth_type"]):
        ap = d["growth_apply"][k]
        f = dict(type=int(t), t=tuple(d["growth_time"][k]), z=tuple(d["growth_z"][k]),
                 apply=int(ap[0] + 2 * ap[1] + 4 * ap[2] + 8 * ap[3]))
        if t == 1:
            f["rate"], f["angle"] = tissue.uniform_growth_input(d["growth_rate_per_hour"][k],
                                                                d["growth_angle_deg"][k])
        else:
            f["rates"] = tissue.grid_growth_input(d[f"growth_grid{k}_rates"])
            f["angles_deg"] = d[f"growth_grid{k}_angles"]
        fns.append(f)
    cols = None
    if int(d["column_volume_conservation"]):
        cols = tissue.conserved_columns(stacks, etype, place, ecm, act, L)
    extra = dict(apical_of=tissue.apical_positioning(stacks, etype, place, ne, L),
                 excluded=tissue.ecm_only_nodes(conn, len(nodes), ecm), columns=cols,
                 ecm=ecm, actin=act, interp=int(d["growth_interp"]),
                 pinned=bool(d["growth_pinned"]), symmetric_y=bool(d["symmetric_y"]))
    return P, fns, extra


def aligned_start(d, k):
    """Positions at the start of the step that produced frame k, after the
    rigid DV alignment of Simulation::runOneStep."""
    dorsal, ventral = assign_tips(d["mesh_nodes"], d["mesh_node_info"])
    start = d["mesh_nodes"] if k == 0 else d["frames"][k - 1]
    return align_dv_to_x(start, dorsal, ventral)
