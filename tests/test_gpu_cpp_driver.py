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
