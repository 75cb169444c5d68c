"""Pins the oracle's whole NR step against the reference's own Pardiso
regression outputs (tests/sim_pardiso_reg/run0700x/expected_out in the
reference repository, extracted into tests/golden/run0700x.npz):
run07007 (rectangle, free) and run07008 (half wing disc: y fixed on the
symmetry line, circumference x/y bound to the basal node of each column)."""
import numpy as np
import pytest

import replay

# (run, k): the step that produced frame k+1; k = -1 starts from the mesh
STEPS = [("run07007", k) for k in range(10)] + [("run07008", k) for k in range(-1, 10)]

# Save_Frame holds 10 significant digits.  Perturbing a start frame by that
# rounding (uniform within half a unit of the 10th digit) moves the replayed
# step's result by about 5e-10 in run07007 (positions ~10 um) and 7e-8 in
# run07008 (positions ~100 um, dt = 3600 s).  The oracle's worst errors over
# all replayed steps are 9.0e-10 (run07007) and 4.4e-8 (run07008, 6.2e-10
# relative); the bounds below are those errors with a 1.5x margin, within the
# 1e-9 relative bar and far below what a wrong fixing or binding set gives
# (3e-2 and more).
POS_TOL = {"run07007": 1.5e-9, "run07008": 7e-8}
# Save_Force is the internal force at the saved positions; the same rounding
# of those positions moves it by about 7e-6 (run07007) and 2e-3 (run07008,
# ECM E = 1600 on 0.2 um thick prisms).
FORCE_TOL = {"run07007": 2e-5, "run07008": 5e-3}


@pytest.fixture(scope="module")
def runs():
    return {r: replay.load(r) for r in ("run07007", "run07008")}


def _problem(orc, s):
    return orc.Problem(s["conn"], s["refpos"], s["invFg"], s["lam"], s["mu"], s["eta"],
                       s["ext_visc"], s["visc_surf"], s["fixed"], s["pairs"])


@pytest.mark.parametrize("run,k", STEPS)
def test_replay_step_positions_and_du_log(orc, runs, run, k):
    s = replay.step_inputs(runs[run], k, orc)
    out, du, conv = _problem(orc, s).nr_step_dense(s["un"], s["dt"], 1e-8, 20)
    assert conv
    # the reference's own acceptance tolerance is 1e-8 absolute
    # (tests/sim_pardiso_reg/test_SimulationPardiso.py:189)
    assert np.abs(out - s["expected"]).max() < POS_TOL[run]
    assert len(du) == len(s["du_log"])
    # printed with 6 significant digits; iterates below 1e-10 are dominated by
    # the direct solver's rounding and are compared only for their magnitude
    big = s["du_log"] > 1e-10
    assert np.allclose(du[big], s["du_log"][big], rtol=1e-5, atol=0)
    assert np.all(du[~big] < 1e-10)
    if len(s["pairs"]):
        # a slave moves with its master (to the rounding of u - u_n)
        d = out - s["un"]
        assert np.abs(d[s["pairs"][:, 0]] - d[s["pairs"][:, 1]]).max() <= 1e-12
    assert np.all(out[s["fixed"] == 1] == s["un"][s["fixed"] == 1])


@pytest.mark.parametrize("run,k", [(r, k) for r, k in STEPS if k >= 0])
def test_save_force_fixture(orc, runs, run, k):
    """Save_Force holds the internal nodal forces (elastic + internal viscous,
    ShapeBase::writeInternalForcesTogeAndgv) at the last NR iterate of the
    step, evaluated at the saved positions."""
    s = replay.step_inputs(runs[run], k, orc)
    K, g, gi = _problem(orc, s).assemble_dense(s["expected"], s["un"], s["dt"])
    F = s["forces"]
    assert np.abs(F).max() > 1.0
    assert np.abs(gi - F).max() < FORCE_TOL[run]


def test_first_step_is_at_rest(runs):
    assert runs["run07007"]["du_log"][0][0] < 1e-14


@pytest.mark.parametrize("run", ["run07007", "run07008"])
def test_tips_match_printed_indexes(runs, run):
    """Simulation::assignTips on the mesh gives the dorsal and ventral tips the
    reference printed (expected_out/tmp)."""
    d = runs[run]
    assert replay.assign_tips(d["mesh_nodes"], d["mesh_node_info"]) == \
        tuple(int(v) for v in d["printed_tips"])


def test_run07008_constraints_match_saved_files(runs):
    """The fixed set and pair list rebuilt from the mesh agree with what the
    reference saved: the circumference flags of Save_Frame, and
    Save_NodeBinding, which omits the slaves of node 0 (writeNodeBinding keeps
    slaveTo > 0).  Node 0 is the ventral tip, on the symmetry line: its
    column is bound in x only."""
    d = runs["run07008"]
    fixed, pairs, circ = replay.constraints(d)
    nn = len(d["mesh_nodes"])
    assert np.array_equal(circ, d["node_info"][:, 2])
    assert np.array_equal(replay.saved_binding(pairs, nn), d["binding"])
    omitted = {tuple(p) for p in pairs} - {tuple(p) for p in d["binding"]}
    assert omitted and all(m == 0 for _, m in omitted)
    assert all(s % 3 == 0 for s, _ in omitted)
    # slave DOF are never fixed and never masters
    assert not np.any(fixed[pairs[:, 0]]) and not np.any(fixed[pairs[:, 1]])
    assert not set(pairs[:, 0]) & set(pairs[:, 1])
    # the y of every node on the symmetry line is fixed, and no other DOF is
    on_line = np.abs(d["mesh_nodes"][:, 1]) < replay.SYM_LIM
    assert np.array_equal(fixed.reshape(-1, 3)[:, 1] == 1, on_line)
    assert not np.any(fixed.reshape(-1, 3)[:, [0, 2]])
