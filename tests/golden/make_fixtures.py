"""Extract the reference's own regression data into compact .npz fixtures.

Reads (as data only) the Pardiso regression outputs and inputs shipped with the
reference repository:

  tests/sim_pardiso_reg/run07007/expected_out/{Save_Frame, Save_Growth,
      Save_PhysicalProp, Save_Force, Save_NodeBinding,
      Save_SpecificElementAndNodeTypes, tmp}
  tests/sim_no_pardiso_reg/run07007/{smallRectangle.mesh, modelinput07007}

and writes tests/golden/run07007.npz.  Binary layouts follow the writers in
TissueFolding/SourceCode/Simulation.cpp:8430-8638 (Save_Frame text with
precision(10) positions, FP64 Save_Growth F_g per element, Save_PhysicalProp
E/eta/zRemodelling per element then external viscosity per node, Save_Force
SystemForces per node) and Simulation.cpp:1674-1720 (specific element types).

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import os
import re
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_mesh(path):
    """.mesh format: Simulation.cpp:2286-2321 (nodes) and 2431-2447 (prisms)."""
    tok = open(path).read().split()
    pos = 0
    n_nodes = int(tok[pos]); pos += 1
    nodes = np.zeros((n_nodes, 3))
    node_info = np.zeros((n_nodes, 3), dtype=np.int32)
    for i in range(n_nodes):
        nodes[i] = [float(t) for t in tok[pos:pos + 3]]
        node_info[i] = [int(t) for t in tok[pos + 3:pos + 6]]
        pos += 6
    n_el = int(tok[pos]); pos += 1
    conn = np.zeros((n_el, 6), dtype=np.int32)
    refpos = np.zeros((n_el, 18))
    for e in range(n_el):
        assert int(tok[pos]) == 1, "only prisms"
        pos += 1
        conn[e] = [int(t) for t in tok[pos:pos + 6]]; pos += 6
        refpos[e] = [float(t) for t in tok[pos:pos + 18]]; pos += 18
    return nodes, node_info, conn, refpos


def parse_frames(path):
    lines = open(path).read().splitlines()
    frames, times, conns = [], [], []
    info = None
    i = 0
    while i < len(lines):
        m = re.match(r"=+ TIME:\s+([0-9.eE+-]+)=+", lines[i])
        if not m:
            i += 1
            continue
        times.append(float(m.group(1)))
        i += 1
        nn = int(lines[i]); i += 1
        P = np.zeros((nn, 3))
        inf = np.zeros((nn, 3), dtype=np.int32)
        for k in range(nn):
            t = lines[i].split(); i += 1
            P[k] = [float(x) for x in t[:3]]
            inf[k] = [int(x) for x in t[3:6]]
        ne = int(lines[i]); i += 1
        C = np.zeros((ne, 6), dtype=np.int32)
        for k in range(ne):
            t = lines[i].split(); i += 1
            C[k] = [int(x) for x in t[2:8]]
        frames.append(P)
        conns.append(C)
        info = inf
    return np.array(times), np.array(frames), conns, info


def parse_du_log(path):
    """Per time step, the |du|_2 of every NR iteration (NewtonRaphsonSolver.cpp:815-823)."""
    steps = []
    cur = None
    for line in open(path, errors="replace"):
        if line.startswith("iteration: 0"):
            cur = []
            steps.append(cur)
        m = re.search(r"(?:not  yet converged via du: norm |converged with displacement: norm)"
                      r"\s*([0-9.eE+-]+)", line)
        if m and cur is not None:
            cur.append(float(m.group(1)))
    return steps


def read_specific_types(path):
    raw = np.fromfile(path, dtype=np.int32)
    p = 0
    n_actin = raw[p]; p += 1
    actin = raw[p:p + n_actin]; p += n_actin
    n_ecm = raw[p]; p += 1
    ecm = raw[p:p + n_ecm]; p += n_ecm
    return actin.astype(np.int32), ecm.astype(np.int32)


def read_node_binding(path):
    tok = open(path).read().split()
    out, p = [], 0
    while p < len(tok):
        n = int(tok[p]); p += 1
        pairs = np.array([int(t) for t in tok[p:p + 2 * n]], dtype=np.int32).reshape(n, 2)
        p += 2 * n
        out.append(pairs)
    return out


def modelinput_value(path, key):
    for line in open(path):
        if key in line:
            return line.split(key)[1].split()[0]
    raise KeyError(key)


def make_run07007():
    exp = os.path.join(REF, "tests/sim_pardiso_reg/run07007/expected_out")
    inp = os.path.join(REF, "tests/sim_no_pardiso_reg/run07007")
    nodes0, node_info0, conn, refpos = parse_mesh(os.path.join(inp, "smallRectangle.mesh"))
    times, frames, conns, node_info = parse_frames(os.path.join(exp, "Save_Frame"))
    for C in conns:
        assert np.array_equal(C, conn), "element node order differs from the mesh"
    n_f, n_n, n_e = len(times), nodes0.shape[0], conn.shape[0]
    Fg = np.fromfile(os.path.join(exp, "Save_Growth"), dtype="<f8").reshape(n_f, n_e, 9)
    pp = np.fromfile(os.path.join(exp, "Save_PhysicalProp"), dtype="<f8")
    per = 3 * n_e + 3 * n_n
    pp = pp.reshape(n_f, per)
    elem_pp = pp[:, :3 * n_e].reshape(n_f, n_e, 3)
    ext_visc = pp[:, 3 * n_e:].reshape(n_f, n_n, 3)
    forces = np.fromfile(os.path.join(exp, "Save_Force"), dtype="<f8").reshape(n_f, n_n, 3)
    actin, ecm = read_specific_types(os.path.join(exp, "Save_SpecificElementAndNodeTypes"))
    binding = read_node_binding(os.path.join(exp, "Save_NodeBinding"))
    assert all(len(b) == 0 for b in binding)
    du_log = parse_du_log(os.path.join(exp, "tmp"))
    nmax = max(len(s) for s in du_log)
    du = np.full((len(du_log), nmax), np.nan)
    for k, s in enumerate(du_log):
        du[k, :len(s)] = s
    mi = os.path.join(inp, "modelinput07007")
    nu = float(modelinput_value(mi, "PoissonsRatio:"))
    dt = float(modelinput_value(mi, "TimeStep(sec):"))
    out = os.path.join(HERE, "run07007.npz")
    np.savez_compressed(
        out, times=times, frames=frames, node_info=node_info, mesh_nodes=nodes0,
        mesh_node_info=node_info0, conn=conn, refpos=refpos, Fg=Fg,
        E=elem_pp[:, :, 0], eta=elem_pp[:, :, 1], zremodel=elem_pp[:, :, 2],
        ext_visc=ext_visc, forces=forces, actin=actin, ecm=ecm, du_log=du,
        nu=nu, dt=dt)
    print("wrote", out, "frames", n_f, "nodes", n_n, "elements", n_e)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference repository not present; fixtures are committed already")
    make_run07007()
