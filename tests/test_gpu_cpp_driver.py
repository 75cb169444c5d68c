"""The C++ drop-in surface (include/tissue_nr.hpp) driven by nr_driver, which
runs Simulation::updateStepNR's loop with the reference's method names,
replays run07007 and run07008 (Node::FixedPos on the symmetry line and
NewtonRaphsonSolver::slaveMasterList from the circumference binding) against
the reference's Pardiso output."""
import os
import subprocess

import numpy as np
import pytest

import replay

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tissueorigami_amd", "nr_driver")


def write_input(path, s, nn):
    un = s["un"].reshape(nn, 3)
    fixed = s["fixed"].reshape(nn, 3)
    with open(path, "w") as f:
        f.write(f"{nn} {len(s['conn'])} {s['dt']!r} 1e-8 20\n")
        for i in range(nn):
            f.write(" ".join(repr(float(v)) for v in un[i]) + " " +
                    " ".join(repr(float(v)) for v in s["ext_visc"][i]) +
                    f" {float(s['visc_surf'][i])!r} " + " ".join(str(int(v)) for v in fixed[i]) +
                    "\n")
        for e in range(len(s["conn"])):
            vals = [str(int(v)) for v in s["conn"][e]]
            vals += [repr(float(v)) for v in s["refpos"][e]]
            vals += [repr(float(s["E"][e])), repr(float(s["nu"][e])), repr(float(s["eta"][e]))]
            vals += [repr(float(v)) for v in s["invFg"][e]]
            f.write(" ".join(vals) + "\n")
        f.write(f"{len(s['pairs'])}\n")
        for sl, ms in s["pairs"]:
            f.write(f"{int(sl)} {int(ms)}\n")


def read_output(path, nn):
    lines = open(path).read().splitlines()
    u = np.array([[float(x) for x in l.split()] for l in lines[:nn]]).reshape(-1)
    its = int(lines[nn].split()[1])
    du = np.array([float(l.split()[1]) for l in lines[nn + 1:nn + 1 + its]])
    return u, du


@pytest.mark.parametrize("mode", ["loop", "step"])
@pytest.mark.parametrize("run,k", [("run07007", 0), ("run07007", 4), ("run07007", 9),
                                   ("run07008", -1), ("run07008", 4), ("run07008", 9)])
def test_cpp_driver_replays_pardiso_runs(orc, tmp_path, mode, run, k):
    d = replay.load(run)
    s = replay.step_inputs(d, k, orc)
    nn = d["frames"].shape[1]
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input(inp, s, nn)
    subprocess.run([DRIVER, str(inp), str(out), mode], check=True, timeout=120)
    u, du = read_output(out, nn)
    # propagated Save_Frame rounding, as in test_oracle_fixtures.py
    assert np.abs(u - s["expected"]).max() < {"run07007": 1.5e-9, "run07008": 7e-8}[run]
    big = s["du_log"] > 1e-10
    assert len(du) == len(s["du_log"])
    assert np.allclose(du[big], s["du_log"][big], rtol=1e-5, atol=0)


@pytest.mark.parametrize("mode", ["loop", "step"])
def test_cpp_driver_stiffness_multiplier_and_ablation(orc, tmp_path, mode):
    """ShapeBase::stiffnessMultiplier (not 1) and laser ablation through the C++
    surface: ablated prisms (ShapeBase::IsAblated) and the ablated-node list
    of checkJacobianForAblatedNodes, called where the reference calls it.
    The result equals the oracle's dense step with E x multiplier and the same
    ablation; the node left without live prisms does not move."""
    import problems
    p = problems.perturbed_problem(seed=44, visc=True, fixed=True)
    conn = p["mesh"]["conn"]
    nn, ne = len(p["mesh"]["nodes"]), len(conn)
    node = int(conn[0, 0])
    flags = np.zeros(ne, np.uint8)
    flags[np.where((conn == node).any(axis=1))[0]] = 1
    flags[[ne // 2, ne - 2]] = 1
    ext = p["ext_visc"].copy(); ext[node] = 0.0
    area = p["area"].copy(); area[node] = 0.0
    fixed = p["fixed"].copy(); fixed[node] = 0
    mult = np.random.default_rng(3).uniform(0.5, 2.0, ne)
    s = dict(conn=conn, refpos=p["refpos"], E=p["E"], nu=p["nu"], eta=p["eta"], invFg=p["invFg"],
             un=p["un"], ext_visc=ext, visc_surf=area, fixed=fixed, dt=p["dt"],
             pairs=np.zeros((0, 2), np.int32))
    inp, out = tmp_path / "in.txt", tmp_path / "out.txt"
    write_input(inp, s, nn)
    ablated_nodes = [node, 5]
    with open(inp, "a") as f:
        f.write("stiffness " + " ".join(repr(float(v)) for v in mult) + "\n")
        ids = np.where(flags)[0]
        f.write(f"ablated_prisms {len(ids)} " + " ".join(str(int(i)) for i in ids) + "\n")
        f.write(f"ablated_nodes {len(ablated_nodes)} " + " ".join(map(str, ablated_nodes)) + "\n")
    subprocess.run([DRIVER, str(inp), str(out), mode], check=True, timeout=120)
    u, du = read_output(out, nn)
    lam, mu = orc.lame(mult * p["E"], p["nu"])
    P = orc.Problem(conn, p["refpos"], p["invFg"], lam, mu, p["eta"], ext, area, fixed)
    with orc.ablation(prisms=flags, nodes=ablated_nodes):
        uo, duo, conv = P.nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert conv and len(du) == len(duo)
    assert np.abs(u - uo).max() <= 1e-9 * np.abs(uo).max()
    assert np.array_equal(u[3 * node:3 * node + 3], p["un"][3 * node:3 * node + 3])
    # without the multiplier the result differs: it was applied
    lam1, mu1 = orc.lame(p["E"], p["nu"])
    P1 = orc.Problem(conn, p["refpos"], p["invFg"], lam1, mu1, p["eta"], ext, area, fixed)
    with orc.ablation(prisms=flags, nodes=ablated_nodes):
        u1, _, _ = P1.nr_step_dense(p["un"], p["dt"], 1e-8, 20)
    assert np.abs(u1 - uo).max() > 1e-6 * np.abs(uo).max()
