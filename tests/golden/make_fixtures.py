"""Extract the reference's own regression data into compact .npz fixtures.

Reads (as data only) the Pardiso regression outputs and inputs shipped with the
reference repository:

  tests/sim_pardiso_reg/run0700x/expected_out/{Save_Frame, Save_Growth,
      Save_GrowthRate, Save_PhysicalProp, Save_Force, Save_NodeBinding,
      Save_GrowthRedistribution, Save_Packing, Save_TensionCompression,
      Save_SpecificElementAndNodeTypes, tmp, MeshFromEndPoint.mesh}
  tests/sim_no_pardiso_reg/run0700x/{<mesh>, modelinput0700x}

for x = 7 (smallRectangle.mesh) and x = 8 (smallWingDisc.mesh), and writes
tests/golden/run0700x.npz with the readers of tissueorigami_amd/io.py
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


def growth_options(path, inp):
    """The GrowthOptions section of a model input file
    (ModelInputObject::readGrowthOptions, ModelInputObject.cpp:290-400) and
    its functions (types 1 and 3), as data: times, parts, z range, uniform
    rates (per hour) and angle (degrees), grid files read by io.read_growth_grid;
    with the explicit-ECM and column-conservation switches."""
    tok = open(path).read().split()
    def after(key, k=0, start=0):
        return tok[tok.index(key, start) + 1 + k]
    g0 = tok.index("GrowthOptions:")
    n = int(after("NumberofGrowthFunctions(int):", 0, g0))
    out = {"growth_pinned": int(after("GridGrowthsPinnedOnInitialMesh(bool):", 0, g0)),
           "growth_interp": int(after("GridGrowthsInterpolationType(0=step,1=linear):", 0, g0))}
    npin = int(after("PinningUpdateTimes(number-times(sec)):", 0, g0))
    out["growth_pin_times"] = np.array([float(after("PinningUpdateTimes(number-times(sec)):", 1 + j, g0))
                                        for j in range(npin)])
    types, times, apply, zr, rate, angle = [], [], [], [], [], []
    p = g0
    for f in range(n):
        p = tok.index("GrowthFunctionType(int-seeDocumentation):", p) + 1
        t = int(tok[p])
        types.append(t)
        times.append([float(after("InitialTime(sec):", 0, p)), float(after("FinalTime(sec):", 0, p))])
        apply.append([int(after(k, 0, p)) for k in ("ApplyToColumnarLayer(bool):",
                                                   "ApplyToPeripodialMembrane(bool):",
                                                   "ApplyToBasalECM(bool):",
                                                   "ApplyToLateralECM(bool):")])
        if t == 1:
            q = tok.index("MaxValue(fractionPerHour-DV,AP,AB):", p)
            rate.append([float(tok[q + 1 + k]) for k in range(3)])
            angle.append(float(after("Angle(degrees):", 0, p)))
            zr.append([0.0, 1.0])
        elif t == 3:
            rate.append([0.0, 0.0, 0.0])
            angle.append(0.0)
            q = tok.index("zRange:", p)
            zr.append([float(tok[q + 1]), float(tok[q + 2])])
            gr, ga = io.read_growth_grid(os.path.join(inp, after("Filename(full-path):", 0, p)))
            out[f"growth_grid{f}_rates"], out[f"growth_grid{f}_angles"] = gr, ga
        else:
            raise ValueError(f"growth function type {t} not extracted")
    out.update(growth_type=np.array(types, np.int32), growth_time=np.array(times),
               growth_apply=np.array(apply, np.int32), growth_z=np.array(zr),
               growth_rate_per_hour=np.array(rate), growth_angle_deg=np.array(angle))
    out["explicit_ecm"] = int(after("ThereIsExplicitECM(bool):"))
    key = "ThereIsColumnViseVolumeConservation(bool):"
    out["column_volume_conservation"] = int(after(key)) if key in tok else 0
    return out


def make_run(run, mesh_name):
    """run07007 (rectangle, no symmetry, no binding) and run07008 (half wing
    disc, symmetric in y, circumference x/y bound to the basal node)."""
    exp = os.path.join(REF, f"tests/sim_pardiso_reg/{run}/expected_out")
    inp = os.path.join(REF, f"tests/sim_no_pardiso_reg/{run}")
    nodes0, node_info0, conn, refpos = io.read_mesh(os.path.join(inp, mesh_name))
    fr = io.read_frames(os.path.join(exp, "Save_Frame"))
    times, frames, node_info = fr["times"], fr["positions"], fr["node_info"]
    for C in fr["conn"]:
        assert np.array_equal(C, conn), "element node order differs from the mesh"
    n_f, n_n, n_e = len(times), nodes0.shape[0], conn.shape[0]
    Fg = io.read_growth(os.path.join(exp, "Save_Growth"), n_e).reshape(n_f, n_e, 9)
    rate = io.read_growth_rate(os.path.join(exp, "Save_GrowthRate"), n_e)
    pp = io.read_physical_prop(os.path.join(exp, "Save_PhysicalProp"), n_e, n_n)
    forces = io.read_forces(os.path.join(exp, "Save_Force"), n_n)
    redist = io.read_growth_redistribution(os.path.join(exp, "Save_GrowthRedistribution"), n_e)
    packing = io.read_packing(os.path.join(exp, "Save_Packing"))
    assert all(len(p) == 0 for p, _ in packing)
    types = io.read_specific_types(os.path.join(exp, "Save_SpecificElementAndNodeTypes"))
    binding = io.read_node_binding(os.path.join(exp, "Save_NodeBinding"))
    assert all(np.array_equal(b, binding[0]) for b in binding), "binding changes between frames"
    extra, names = {}, ["Save_Frame", "Save_Growth", "Save_GrowthRate", "Save_PhysicalProp",
                        "Save_Force", "Save_GrowthRedistribution", "Save_Packing",
                        "Save_NodeBinding", "Save_SpecificElementAndNodeTypes"]
    tc = os.path.join(exp, "Save_TensionCompression")
    if os.path.exists(tc):
        extra["strain"] = io.read_tension_compression(tc, n_e)
        names.append("Save_TensionCompression")
    tmp = os.path.join(exp, "tmp")
    if os.path.exists(tmp):
        du_log = io.read_du_log(tmp)
        nmax = max(len(s) for s in du_log)
        du = np.full((len(du_log), nmax), np.nan)
        for k, s in enumerate(du_log):
            du[k, :len(s)] = s
        extra["du_log"] = du
        for line in open(tmp):
            if "dorsalTipIndex:" in line:      # Simulation.cpp:8227
                w = line.split()
                extra["printed_tips"] = np.array([int(w[w.index("dorsalTipIndex:") + 1]),
                                                  int(w[w.index("ventralTipIndex:") + 1])])
                break
    mi = os.path.join(inp, "modelinput" + run[3:])
    nu = float(modelinput_value(mi, "PoissonsRatio:"))
    dt = float(modelinput_value(mi, "TimeStep(sec):"))
    sym_y = int(modelinput_value(mi, "symmetricInY(bool):"))
    bind_circ = int(modelinput_value(mi, "bindCircumferenceXYToBasal(bool):"))
    extra.update(growth_options(mi, inp))
    end = os.path.join(exp, "MeshFromEndPoint.mesh")
    end_nodes, end_info, end_conn, _ = io.read_mesh(end)
    assert np.array_equal(end_conn, conn)
    extra["end_nodes"], extra["end_node_info"] = end_nodes, end_info
    names.append("MeshFromEndPoint.mesh")
    hashes = {f"sha256_{name}": sha256(os.path.join(exp, name)) for name in names}
    out = os.path.join(HERE, run + ".npz")
    np.savez_compressed(
        out, times=times, frames=frames, node_info=node_info, mesh_nodes=nodes0,
        mesh_node_info=node_info0, conn=conn, refpos=refpos, Fg=Fg, growth_rate=rate,
        E=pp["E"], eta=pp["eta"], zremodel=pp["z_remodelling"], ext_visc=pp["ext_visc"],
        forces=forces, redistribution=redist, n_packing_frames=len(packing),
        actin=types["actin"], ecm=types["ecm"], binding=binding[0],
        n_binding_frames=len(binding), nu=nu, dt=dt, symmetric_y=sym_y,
        bind_circumference=bind_circ, **extra, **hashes)
    print("wrote", out, "frames", n_f, "nodes", n_n, "elements", n_e)


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference repository not present; fixtures are committed already")
    make_run("run07007", "smallRectangle.mesh")
    make_run("run07008", "smallWingDisc.mesh")
