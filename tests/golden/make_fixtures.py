"""Extract the reference's own regression data into compact .npz fixtures.

Reads (as data only) the Pardiso regression outputs and inputs shipped with the
reference repository:

  tests/sim_pardiso_reg/run07007/expected_out/{Save_Frame, Save_Growth,
      Save_PhysicalProp, Save_Force, Save_NodeBinding,
      Save_SpecificElementAndNodeTypes, tmp}
  tests/sim_no_pardiso_reg/run07007/{smallRectangle.mesh, modelinput07007}

and writes tests/golden/run07007.npz with the readers of tissueorigami_amd/io.py
(the formats of the writers in TissueFolding/SourceCode/Simulation.cpp:8430-8638
and 1674-1737).  The SHA-256 of every file read is stored beside the data, so
that the writers of tissueorigami_amd/io.py can be checked byte for byte
without the reference present.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import hashlib
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", ".."))
from tissueorigami_amd import io  # noqa: E402

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def modelinput_value(path, key):
    for line in open(path):
        if key in line:
            return line.split(key)[1].split()[0]
    raise KeyError(key)


def sha256(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def make_run07007():
    exp = os.path.join(REF, "tests/sim_pardiso_reg/run07007/expected_out")
    inp = os.path.join(REF, "tests/sim_no_pardiso_reg/run07007")
    nodes0, node_info0, conn, refpos = io.read_mesh(os.path.join(inp, "smallRectangle.mesh"))
    fr = io.read_frames(os.path.join(exp, "Save_Frame"))
    times, frames, node_info = fr["times"], fr["positions"], fr["node_info"]
    for C in fr["conn"]:
        assert np.array_equal(C, conn), "element node order differs from the mesh"
    n_f, n_n, n_e = len(times), nodes0.shape[0], conn.shape[0]
    Fg = io.read_growth(os.path.join(exp, "Save_Growth"), n_e).reshape(n_f, n_e, 9)
    rate = io.read_growth_rate(os.path.join(exp, "Save_GrowthRate"), n_e)
    pp = io.read_physical_prop(os.path.join(exp, "Save_PhysicalProp"), n_e, n_n)
    forces = io.read_forces(os.path.join(exp, "Save_Force"), n_n)
    strain = io.read_tension_compression(os.path.join(exp, "Save_TensionCompression"), n_e)
    redist = io.read_growth_redistribution(os.path.join(exp, "Save_GrowthRedistribution"), n_e)
    packing = io.read_packing(os.path.join(exp, "Save_Packing"))
    assert all(len(p) == 0 for p, _ in packing)
    types = io.read_specific_types(os.path.join(exp, "Save_SpecificElementAndNodeTypes"))
    binding = io.read_node_binding(os.path.join(exp, "Save_NodeBinding"))
    assert all(len(b) == 0 for b in binding)
    du_log = io.read_du_log(os.path.join(exp, "tmp"))
    nmax = max(len(s) for s in du_log)
    du = np.full((len(du_log), nmax), np.nan)
    for k, s in enumerate(du_log):
        du[k, :len(s)] = s
    mi = os.path.join(inp, "modelinput07007")
    nu = float(modelinput_value(mi, "PoissonsRatio:"))
    dt = float(modelinput_value(mi, "TimeStep(sec):"))
    hashes = {f"sha256_{name}": sha256(os.path.join(exp, name)) for name in (
        "Save_Frame", "Save_Growth", "Save_GrowthRate", "Save_PhysicalProp", "Save_Force",
        "Save_TensionCompression", "Save_GrowthRedistribution", "Save_Packing",
        "Save_NodeBinding", "Save_SpecificElementAndNodeTypes")}
    out = os.path.join(HERE, "run07007.npz")
    np.savez_compressed(
        out, times=times, frames=frames, node_info=node_info, mesh_nodes=nodes0,
        mesh_node_info=node_info0, conn=conn, refpos=refpos, Fg=Fg, growth_rate=rate,
        E=pp["E"], eta=pp["eta"], zremodel=pp["z_remodelling"], ext_visc=pp["ext_visc"],
        forces=forces, strain=strain, redistribution=redist, n_packing_frames=len(packing),
        actin=types["actin"], ecm=types["ecm"], n_binding_frames=len(binding), du_log=du,
        nu=nu, dt=dt, **hashes)
    print("wrote", out, "frames", n_f, "nodes", n_n, "elements", n_e)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference repository not present; fixtures are committed already")
    make_run07007()
