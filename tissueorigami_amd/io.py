"""Readers and writers of the reference's mesh input and save-file formats.

The formats are the reference's, restated from its writers and readers so that
files written here load in its tools (VisualiseTissueFolding, the
`tests/simOutputsToTxt.cpp` converter) and its regression outputs load here:

  .mesh input          Simulation::initiateNodesFromMeshInput /
                       initiateElementsFromMeshInput / initiatePrismFromMeshInput
                       (Simulation.cpp:2286-2321, 2431-2447)
  MeshFromEndPoint.mesh  the final state as a new .mesh input
                       (Simulation::writeRelaxedMeshFromCurrentState, Simulation.cpp:1833-1879)
  Save_Frame (text)    Simulation::writeSaveFileStepHeader / writeNodes /
                       writeElements / writeSaveFileStepFooter (Simulation.cpp:8445-8519)
  Save_Growth          F_g, 9 doubles per element per frame (writeGrowth, 8572-8587)
  Save_GrowthRate      3 doubles per element per frame (writeGrowth)
  Save_PhysicalProp    E, internal viscosity, z remodelling per element, then
                       external viscosity[3] per node (writePhysicalProp, 8623-8638)
  Save_Force           SystemForces[3] per node (writeForces, 8591-8599)
  Save_TensionCompression  6 strain components per element (writeTensionCompression)
  Save_GrowthRedistribution  2 bools per element (writeGrowthRedistribution)
  Save_Packing         int count, node pairs, packing forces (writePacking)
  Save_NodeBinding     text: count, then "slave master" DOF pairs (writeNodeBinding, 8546-8566)
  Save_SpecificElementAndNodeTypes  int32 lists (writeSpecificNodeTypes,
                       Simulation.cpp:1674-1737)
  NR log (tmp)         |du|_2 per NR iteration, printed by
                       NewtonRaphsonSolver::checkConvergenceViaDeltaU (NewtonRaphsonSolver.cpp:815-823)

Binary files are little-endian, frames appended one after another.  Text
numbers follow C++ iostream's default float formatting, which is printf's %g
at the stream's precision.
"""
import re

import numpy as np


# ---------------------------------------------------------------- .mesh input
def read_mesh(path):
    """Nodes (n, 3), node info (n, 3: placement, type, atCircumference),
    prism connectivity (m, 6) and reference positions (m, 18) as written in the
    file (the reference reorders each prism afterwards with
    Prism::checkRotationConsistency3D, Prism.cpp:369-417)."""
    tok = open(path).read().split()
    p = 0
    n = int(tok[p]); p += 1
    a = np.array(tok[p:p + 6 * n], dtype=float).reshape(n, 6)
    p += 6 * n
    m = int(tok[p]); p += 1
    conn = np.zeros((m, 6), np.int32)
    refpos = np.zeros((m, 18))
    for e in range(m):
        if int(tok[p]) != 1:
            raise ValueError(f"element {e}: shape type {tok[p]} (only prisms, type 1, are read)")
        conn[e] = np.array(tok[p + 1:p + 7], dtype=np.int32)
        refpos[e] = np.array(tok[p + 7:p + 25], dtype=float)
        p += 25
    return a[:, :3].copy(), a[:, 3:].astype(np.int32), conn, refpos


def write_mesh(path, nodes, node_info, conn, refpos):
    """Writes a .mesh input file the reference's mesh reader accepts (whitespace
    separated; full double precision)."""
    with open(path, "w") as f:
        f.write(f"{len(nodes)}\n")
        for x, i in zip(np.asarray(nodes, float), np.asarray(node_info, int)):
            f.write("%.17g\t%.17g\t%.17g    %d\t%d\t%d\n" % (x[0], x[1], x[2], i[0], i[1], i[2]))
        f.write(f"{len(conn)}\n")
        for c, r in zip(np.asarray(conn, int), np.asarray(refpos, float).reshape(-1, 18)):
            f.write("1\t" + "     ".join(str(v) for v in c) + "     " +
                    "   ".join("%.17g" % v for v in r) + "    \n")


def write_end_mesh(path, positions, node_info, conn):
    """MeshFromEndPoint.mesh as Simulation::writeRelaxedMeshFromCurrentState
    writes it (Simulation.cpp:1833-1879): the current positions become both the
    node positions and each prism's reference positions; every value at the
    stream's default precision (%g, 6 digits) followed by " \t"."""
    P = np.asarray(positions, float).reshape(-1, 3)
    info = np.asarray(node_info, int).reshape(-1, 3)
    C = np.asarray(conn, int).reshape(-1, 6)
    out = [f"{len(P)}\n"]
    for x, i in zip(P, info):
        out.append(" \t".join(["%g" % x[0], "%g" % x[1], "%g" % x[2],
                               str(i[0]), str(i[1]), str(i[2])]) + "\n")
    out.append(f"{len(C)}\n")
    for c in C:
        out.append("1 \t" + "".join(f"{v} \t" for v in c) +
                   "".join("%g \t" % P[v][j] for v in c for j in range(3)) + "\n")
    with open(path, "w") as f:
        f.write("".join(out))


# ---------------------------------------------------------------- Save_Frame
def _g(v, prec, width):
    return ("%" + str(width) + "." + str(prec) + "g") % v


def format_frame(time_s, positions, node_info, conn, refpos, ablated=None):
    """One frame of Save_Frame as the reference writes it: header, node count,
    one line per node (3 x %20.10g position, placement, type, circumference
    as %4d), element count, one line per prism (shape type 1, ablated flag,
    6 x %9d node ids, 18 x %15.4g reference positions), footer."""
    P = np.asarray(positions, float).reshape(-1, 3)
    info = np.asarray(node_info, int).reshape(-1, 3)
    C = np.asarray(conn, int).reshape(-1, 6)
    R = np.asarray(refpos, float).reshape(-1, 18)
    ab = np.zeros(len(C), int) if ablated is None else np.asarray(ablated, int)
    out = ["=============== TIME: " + _g(time_s, 6, 10) +
           "===================================================\n", f"{len(P)}\n"]
    for x, i in zip(P, info):
        out.append(_g(x[0], 10, 20) + _g(x[1], 10, 20) + _g(x[2], 10, 20) +
                   "%4d%4d%4d\n" % (i[0], i[1], i[2]))
    out.append(f"{len(C)}\n")
    for c, r, a in zip(C, R, ab):
        out.append("%4d%4d" % (1, a) + "".join("%9d" % v for v in c) +
                   "".join(_g(v, 4, 15) for v in r) + "\n")
    out.append("=============== END OF TIME: " + _g(time_s, 6, 10) +
               "============================================\n")
    return "".join(out)


def write_frames(path, times, positions, node_info, conn, refpos, mode="w"):
    """Writes (or with mode="a" appends) frames to a Save_Frame file."""
    with open(path, mode) as f:
        for t, P in zip(times, positions):
            f.write(format_frame(t, P, node_info, conn, refpos))


def read_frames(path):
    """All frames of a Save_Frame file: times (f,), positions (f, n, 3), node
    info (n, 3) of the last frame, connectivity (m, 6) and the printed
    reference positions (m, 18, 4 significant digits) of every frame, and the
    ablation flags."""
    lines = open(path).read().splitlines()
    times, frames, conns, refs, abl = [], [], [], [], []
    info = None
    i = 0
    while i < len(lines):
        m = re.match(r"=+ TIME:\s*([0-9.eE+-]+)=+", lines[i])
        if not m:
            i += 1
            continue
        times.append(float(m.group(1)))
        i += 1
        nn = int(lines[i]); i += 1
        blk = np.array(" ".join(lines[i:i + nn]).split(), dtype=float).reshape(nn, 6)
        i += nn
        ne = int(lines[i]); i += 1
        eb = np.array(" ".join(lines[i:i + ne]).split(), dtype=float).reshape(ne, 26)
        i += ne
        frames.append(blk[:, :3])
        info = blk[:, 3:].astype(np.int32)
        conns.append(eb[:, 2:8].astype(np.int32))
        refs.append(eb[:, 8:])
        abl.append(eb[:, 1].astype(np.int32))
    return dict(times=np.array(times), positions=np.array(frames), node_info=info,
                conn=conns, refpos=refs, ablated=abl)


# ---------------------------------------------------------------- binary per-frame files
def _write(path, a, dtype, mode):
    with open(path, mode + "b") as f:
        f.write(np.ascontiguousarray(a, dtype=dtype).tobytes())


def write_growth(path, Fg, mode="w"):
    """F_g per element, row-major 3x3 doubles; Fg is (frames, m, 9) or (m, 9)."""
    _write(path, Fg, "<f8", mode)


def read_growth(path, n_elements):
    return np.fromfile(path, dtype="<f8").reshape(-1, n_elements, 3, 3)


def write_growth_rate(path, rates, mode="w"):
    _write(path, rates, "<f8", mode)


def read_growth_rate(path, n_elements):
    return np.fromfile(path, dtype="<f8").reshape(-1, n_elements, 3)


def write_physical_prop(path, E, eta, z_remodelling, ext_visc, mode="w"):
    """Per frame: (E, internal viscosity, z remodelling) per element, then the
    external viscosity (x, y, z) per node.  Arrays may carry a leading frame
    axis."""
    E = np.asarray(E, float)
    frames = E.reshape(-1, E.shape[-1]).shape[0]
    el = np.stack([np.asarray(E, float), np.asarray(eta, float),
                   np.asarray(z_remodelling, float)], -1).reshape(frames, -1)
    nd = np.asarray(ext_visc, float).reshape(frames, -1)
    _write(path, np.concatenate([el, nd], 1), "<f8", mode)


def read_physical_prop(path, n_elements, n_nodes):
    a = np.fromfile(path, dtype="<f8").reshape(-1, 3 * n_elements + 3 * n_nodes)
    el = a[:, :3 * n_elements].reshape(-1, n_elements, 3)
    return dict(E=el[:, :, 0], eta=el[:, :, 1], z_remodelling=el[:, :, 2],
                ext_visc=a[:, 3 * n_elements:].reshape(-1, n_nodes, 3))


def write_forces(path, forces, mode="w"):
    """SystemForces per node (x, y, z doubles)."""
    _write(path, forces, "<f8", mode)


def read_forces(path, n_nodes):
    return np.fromfile(path, dtype="<f8").reshape(-1, n_nodes, 3)


def write_tension_compression(path, strain, mode="w"):
    """Six strain components per element (doubles)."""
    _write(path, strain, "<f8", mode)


def read_tension_compression(path, n_elements):
    return np.fromfile(path, dtype="<f8").reshape(-1, n_elements, 6)


def write_growth_redistribution(path, flags, mode="w"):
    """(there is redistribution, shrinks element) per element, one byte each."""
    _write(path, np.asarray(flags).astype(bool), "u1", mode)


def read_growth_redistribution(path, n_elements):
    return np.fromfile(path, dtype="u1").reshape(-1, n_elements, 2).astype(bool)


def write_packing(path, pairs, forces0, forces1, mode="w"):
    """One frame of Save_Packing: int32 count, int32 node pairs, then per pair
    the packing forces of its two nodes (doubles)."""
    pairs = np.asarray(pairs, np.int32).reshape(-1, 2)
    f = np.concatenate([np.asarray(forces0, float).reshape(-1, 3),
                        np.asarray(forces1, float).reshape(-1, 3)], 1)
    with open(path, mode + "b") as fh:
        fh.write(np.int32(len(pairs)).tobytes())
        fh.write(pairs.astype("<i4").tobytes())
        fh.write(f.astype("<f8").tobytes())


def read_packing(path):
    """List of frames, each (pairs (k, 2), forces (k, 6))."""
    raw = open(path, "rb").read()
    out, p = [], 0
    while p < len(raw):
        k = int(np.frombuffer(raw, "<i4", 1, p)[0]); p += 4
        pairs = np.frombuffer(raw, "<i4", 2 * k, p).reshape(k, 2); p += 8 * k
        forces = np.frombuffer(raw, "<f8", 6 * k, p).reshape(k, 6); p += 48 * k
        out.append((pairs.copy(), forces.copy()))
    return out


# ---------------------------------------------------------------- text / int files
def write_node_binding(path, pairs, mode="w"):
    """One frame of Save_NodeBinding: the slave/master DOF pair count, then one
    "slave master " line per pair."""
    pairs = np.asarray(pairs, int).reshape(-1, 2)
    with open(path, mode) as f:
        f.write(f"{len(pairs)} \n")
        for s, m in pairs:
            f.write(f"{s} {m} \n")


def read_node_binding(path):
    tok = open(path).read().split()
    out, p = [], 0
    while p < len(tok):
        n = int(tok[p]); p += 1
        out.append(np.array(tok[p:p + 2 * n], dtype=np.int32).reshape(n, 2))
        p += 2 * n
    return out


def write_specific_types(path, actin, ecm, element_bands=(), node_bands=()):
    """Save_SpecificElementAndNodeTypes: int32 actin element ids, ECM element
    ids, (element id, ellipse band) pairs, (node id, ellipse band) pairs, each
    list preceded by its count."""
    parts = []
    for ids in (actin, ecm):
        ids = np.asarray(ids, np.int32).reshape(-1)
        parts += [np.int32(len(ids)).reshape(1), ids]
    for pairs in (element_bands, node_bands):
        pairs = np.asarray(pairs, np.int32).reshape(-1, 2)
        parts += [np.int32(len(pairs)).reshape(1), pairs.reshape(-1)]
    _write(path, np.concatenate(parts), "<i4", "w")


def read_specific_types(path):
    raw = np.fromfile(path, dtype="<i4")
    p = 0
    out = {}
    for key in ("actin", "ecm"):
        n = int(raw[p]); p += 1
        out[key] = raw[p:p + n].astype(np.int32); p += n
    for key in ("element_bands", "node_bands"):
        if p >= len(raw):
            out[key] = np.zeros((0, 2), np.int32)
            continue
        n = int(raw[p]); p += 1
        out[key] = raw[p:p + 2 * n].astype(np.int32).reshape(n, 2); p += 2 * n
    return out


def read_du_log(path):
    """|du|_2 of every NR iteration, per time step, from the reference's
    console log."""
    steps, cur = [], None
    pat = re.compile(r"(?:not  yet converged via du: norm |converged with displacement: norm)"
                     r"\s*([0-9.eE+-]+)")
    for line in open(path, errors="replace"):
        if line.startswith("iteration: 0"):
            cur = []
            steps.append(cur)
        m = pat.search(line)
        if m and cur is not None:
            cur.append(float(m.group(1)))
    return steps


def read_property_grid(path):
    """A physical-property grid file of a time series (TimeSeriesPhysicalProperties::ReadGrid,
    TimeSeriesPhysicalProperties.cpp:34-92): "nX nY", then nY rows of nX values.
    Returns (ny, nx)."""
    with open(path) as f:
        tok = f.read().split()
    nx, ny = int(tok[0]), int(tok[1])
    v = np.array([float(t) for t in tok[2:2 + nx * ny]])
    if v.size != nx * ny:
        raise ValueError(f"{path}: expected {nx * ny} values, found {v.size}")
    return v.reshape(ny, nx)


def write_property_grid(path, grid):
    g = np.asarray(grid, np.float64)
    ny, nx = g.shape
    with open(path, "w") as f:
        f.write(f"{nx} {ny}\n")
        for row in g:
            f.write(" ".join(repr(float(x)) for x in row) + "\n")


def read_growth_grid(path):
    """A growth-rate grid file of a grid-based growth function
    (ModelInputObject::readGrowthType3, ModelInputObject.cpp:746-812): "nX nY",
    then the rates (fraction per hour) of rows j = nY-1 .. 0, each with x, y, z
    of the nX points, then the angles (degrees) of rows j = nY-1 .. 0.  The
    rates go through a float as the reference reads them.  Returns
    (rates (nx, ny, 3) per hour, angles (nx, ny) in degrees)."""
    with open(path) as f:
        tok = f.read().split()
    nx, ny = int(tok[0]), int(tok[1])
    p = 2
    rates = np.zeros((nx, ny, 3))
    for j in range(ny - 1, -1, -1):
        for i in range(nx):
            for k in range(3):
                rates[i, j, k] = float(np.float32(tok[p]))
                p += 1
    angles = np.zeros((nx, ny))
    for j in range(ny - 1, -1, -1):
        for i in range(nx):
            angles[i, j] = float(tok[p])
            p += 1
    return rates, angles


def write_growth_grid(path, rates_per_hour, angles_deg):
    r = np.asarray(rates_per_hour, np.float64)
    a = np.asarray(angles_deg, np.float64)
    nx, ny = a.shape
    with open(path, "w") as f:
        f.write(f"{nx} {ny}\n")
        for j in range(ny - 1, -1, -1):
            f.write(" ".join(repr(float(r[i, j, k])) for i in range(nx) for k in range(3)) + "\n")
        for j in range(ny - 1, -1, -1):
            f.write(" ".join(repr(float(a[i, j])) for i in range(nx)) + "\n")
